// mcfl_main.cpp -- command-line front end, a drop-in for the reference's
// `cuda_fl <file.mci> [-S<seed>] [-A]` (CUDAMCFLmain.cu:290-668).
//
// Same inputs (.mci, bulk file, safeprimes_base32.txt in the working
// directory or the bundled table) and the same output files.  Extensions:
//   -L        excitation on the reference's 32768 per-lane MWC streams
//             (default: one stream per photon, full-GPU occupancy)
//   -F<seed>  seed of the emission stage (the reference uses time(NULL))
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <vector>

#include "../../include/mcfl.h"

int main(int argc, char* argv[]) {
  printf("\nmcfl (MI355X). Build date: %s, %s. \n", __DATE__, __TIME__);
  clock_t time0 = clock();
  if (argc < 2) {
    printf("Not enough input arguments!\n");
    return 1;
  }
  const char* filename = argv[1];
  unsigned long long seed = (unsigned long long)time(NULL);
  unsigned long long fl_seed = (unsigned long long)time(NULL);
  bool lanes = false;
  // pull the build's own flags out before the reference's interpreter sees them
  std::vector<char*> args;
  for (int i = 0; i < argc; ++i) {
    if (i >= 2 && !strcmp(argv[i], "-L")) {
      lanes = true;
      continue;
    }
    if (i >= 2 && !strncmp(argv[i], "-F", 2)) {
      fl_seed = strtoull(argv[i] + 2, nullptr, 10);
      continue;
    }
    args.push_back(argv[i]);
  }
  printf("\nExecuting %s... \n", filename);
  printf("____________________________________________________________________\n\n");
  int ignoreA = 0;
  uint64_t s64 = seed;
  if (mcfl_interpret_arg((int)args.size(), args.data(), &s64, &ignoreA)) return 1;
  seed = s64;

  mcfl_sim* sims = nullptr;
  int nsim = mcfl_read_simulation_data(filename, &sims, ignoreA);
  if (nsim == 0) {
    printf("Something wrong with read_simulation_data!\n");
    return 1;
  }
  printf("\nRead %d simulations\n\n", nsim);
  mcfl_sim* s = &sims[0];  // only the first run is simulated (CUDAMCFLmain.cu:338-366)

  std::vector<uint64_t> x(MCFL_NUM_THREADS);
  std::vector<uint32_t> a(MCFL_NUM_THREADS);
  const char* primes = "safeprimes_base32.txt";
  FILE* pf = fopen(primes, "r");
  if (pf)
    fclose(pf);
  else
    primes = mcfl_default_primes_path();
  if (mcfl_init_rng(x.data(), a.data(), MCFL_NUM_THREADS, primes, seed)) return 1;

  mcfl_ctx* ctx = nullptr;
  mcfl_status st = mcfl_ctx_create(0, primes, &ctx);
  if (st != MCFL_OK) {
    printf("No usable GPU: %s\n", mcfl_status_string(st));
    return 1;
  }
  const size_t nv = (size_t)s->nx * s->ny * s->nz;
  std::vector<double> Fx(nv, 0.0);
  printf("Running PHD simulation...\n");
  uint64_t fhd_photons = 0;
  if (lanes) {
    fhd_photons = mcfl_DoOneSimulation(ctx, s, x.data(), a.data(), Fx.data());
  } else {
    const size_t npix = (size_t)s->det.nx * (s->det.ny + 1);
    std::vector<uint64_t> Rd(npix), Tt(npix), fhd(nv);
    std::vector<uint64_t> tim(s->do_temp_sim ? (size_t)s->det.x_temp_numdets * s->det.y_temp_numdets * s->det.temp_bins : 0);
    mcfl_tallies t;
    memset(&t, 0, sizeof t);
    t.Rd_xy = Rd.data();
    t.Tt_xy = Tt.data();
    t.fhd = fhd.data();
    t.time_xyt = tim.empty() ? nullptr : tim.data();
    mcfl_rng rng;
    memset(&rng, 0, sizeof rng);
    rng.mode = MCFL_RNG_PHOTON;
    rng.seed = seed;
    clock_t t1 = clock();
    st = mcfl_run_excitation(ctx, s, &rng, &t);
    if (st != MCFL_OK) {
      printf("Excitation failed: %s\n", mcfl_status_string(st));
      mcfl_ctx_destroy(ctx);
      return 1;
    }
    double secs = (double)(clock() - t1) / CLOCKS_PER_SEC;
    printf("\nSimulation time: %.2f sec (kernel %.1f ms)\n\n", secs, t.kernel_ms);
    mcfl_write_simulation_results(s, &t, secs);
    fhd_photons = t.num_terminated_photons;
    if (s->fhd_activated == 1)
      for (size_t v = 0; v < nv; ++v) Fx[v] = ((double)fhd[v] / (0xFFFFFFFFull * fhd_photons));
  }
  if (s->fhd_activated == 1) {
    printf("Writing PHD files...\n");
    mcfl_write_phd(s, Fx.data());
  }

  uint64_t fl_photons = 0;
  bool fl_ok = true;
  if (s->do_fl_sim != 0) {
    printf("Flourescence simulation... \n");
    const size_t xy = (size_t)s->det.nx * s->det.ny;
    std::vector<double> FlR(xy), FlT(xy);
    mcfl_fl_result res;
    memset(&res, 0, sizeof res);
    st = mcfl_run_fluorescence(ctx, s, Fx.data(), s->number_of_photons_per_voxel, fl_seed, FlR.data(), FlT.data(), &res);
    if (st != MCFL_OK) {
      printf("Fluorescence failed: %s\n", mcfl_status_string(st));
      fl_ok = false;
    } else {
      fl_photons = res.photons;
      printf("\n\nFlourescence simulation finished! (%llu emitting voxels, kernel %.1f ms)\n",
             (unsigned long long)res.voxels, res.kernel_ms);
      printf("Writing results files...\n");
      mcfl_write_fl(s, FlR.data(), FlT.data());
    }
  }
  if (fhd_photons == s->number_of_photons && fl_ok)
    printf("All done, no errors! :)\n");
  else
    printf("Simulation finished, some photons were not properly simulated. \n");
  printf("Total time: %.2f sec.\n", (double)(clock() - time0) / CLOCKS_PER_SEC);
  printf("Total simulated photons:\n");
  printf("\t %llu FHD photons.\n", (unsigned long long)fhd_photons);
  printf("\t %llu Fluorescence photons.\n", (unsigned long long)fl_photons);
  printf("#############################################\n\n");
  mcfl_ctx_destroy(ctx);
  mcfl_free_simulations(sims, nsim);
  return 0;
}
