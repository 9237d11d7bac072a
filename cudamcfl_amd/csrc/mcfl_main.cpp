// mcfl_main.cpp -- command-line front end, a drop-in for the reference's
// `cuda_fl <file.mci> [-S<seed>] [-A]` (CUDAMCFLmain.cu:290-668).
//
// Same inputs (.mci, bulk file, safeprimes_base32.txt in the working
// directory or the bundled table) and the same output files.  Extensions:
//   -L        excitation and emission stages on the reference's 32768
//             per-lane MWC streams, its launch replays included (default:
//             one stream per photon, full-GPU occupancy)
//   -F<seed>  seed of the emission stage (the reference uses time(NULL))
//   -I<N>     emission stage by importance sampling: N photons launched from
//             voxels drawn in proportion to eY (1 - albedof) Fx, an unbiased
//             estimate of the same FlR/FlT images as the every-voxel loop
//             (photons per voxel x every emitting voxel) with a budget of N
//   several .mci files: their excitation simulations run as batched launches
//             (files sharing a medium share it on the device), then every
//             file's outputs are written as if it had been run alone with the
//             same seeds; this replaces one process per file in scan scripts
//             (ScriptGenerator-CeCAR-Translation.ipynb:48-55, cudamcfl-array.sh:15)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "../../include/mcfl.h"

namespace {

struct Options {
  unsigned long long seed = 0, fl_seed = 0;
  unsigned long long is_photons = 0;  // -I<N>: importance-sampled emission budget (0: every-voxel loop)
  bool lanes = false;
  int ignoreA = 0;
};

// Importance-sampled emission stage: one emission job of o.is_photons photons
// on the medium of s; images scaled to the reference's FlR/FlT units.
mcfl_status emission_is(mcfl_ctx* ctx, const mcfl_sim* s, const std::vector<double>& Fx, const Options& o,
                        std::vector<double>& FlR, std::vector<double>& FlT, double* kernel_ms) {
  const size_t np = (size_t)s->det.nx * (s->det.ny + 1), xy = (size_t)s->det.nx * s->det.ny;
  std::vector<uint64_t> R(np), T(np);
  mcfl_job j;
  memset(&j, 0, sizeof j);
  j.medium = 0;
  j.src_kind = MCFL_SOURCE_EMISSION;
  j.photons = o.is_photons;
  j.seed = o.fl_seed;
  j.Fx = Fx.data();
  j.Rd_xy = R.data();
  j.Tt_xy = T.data();
  const mcfl_sim* media[1] = {s};
  mcfl_status st = mcfl_run_batch(ctx, media, 1, &j, 1, kernel_ms);
  if (st != MCFL_OK) return st;
  const double k = j.fl_unit / ((double)s->det.dx * (double)s->det.dy);
  for (size_t i = 0; i < xy; ++i) {
    FlR[i] = (double)R[i] * k;
    FlT[i] = (double)T[i] * k;
  }
  return MCFL_OK;
}

size_t nvox(const mcfl_sim* s) { return (size_t)s->nx * s->ny * s->nz; }
size_t npix(const mcfl_sim* s) { return (size_t)s->det.nx * (s->det.ny + 1); }
size_t ntime(const mcfl_sim* s) {
  return s->do_temp_sim ? (size_t)s->det.x_temp_numdets * s->det.y_temp_numdets * s->det.temp_bins : 0;
}

// Normalised fhd of an excitation (CUDAMCFLmain.cu:181-183).
std::vector<double> fx_of(const mcfl_sim* s, const std::vector<uint64_t>& fhd, uint64_t finished) {
  std::vector<double> Fx(nvox(s), 0.0);
  if (s->fhd_activated == 1 && fhd.size() == Fx.size())
    for (size_t v = 0; v < Fx.size(); ++v) Fx[v] = ((double)fhd[v] / (0xFFFFFFFFull * finished));
  return Fx;
}

// PHD file and fluorescence stage after an excitation
// (CUDAMCFLmain.cu:368-392,409-652).  Returns false if the emission failed.
bool finish_file(mcfl_ctx* ctx, mcfl_sim* s, std::vector<double>& Fx, const Options& o, const char* primes,
                 uint64_t* fl_photons) {
  if (s->fhd_activated == 1) {
    printf("Writing PHD files...\n");
    mcfl_write_phd(s, Fx.data());
  }
  *fl_photons = 0;
  if (s->do_fl_sim == 0) return true;
  printf("Flourescence simulation... \n");
  const size_t xy = (size_t)s->det.nx * s->det.ny;
  std::vector<double> FlR(xy), FlT(xy);
  if (o.is_photons > 0 && s->bulk_method == 2) {  // (a method-1 medium emits nothing, CUDAMCFLmain.cu:520,535)
    double ms = 0.0;
    const mcfl_status st = emission_is(ctx, s, Fx, o, FlR, FlT, &ms);
    if (st != MCFL_OK) {
      printf("Fluorescence failed: %s\n", mcfl_status_string(st));
      return false;
    }
    *fl_photons = o.is_photons;
    printf("\n\nFlourescence simulation finished! (importance-sampled emission, %llu photons, kernel %.1f ms)\n",
           (unsigned long long)o.is_photons, ms);
    printf("Writing results files...\n");
    mcfl_write_fl(s, FlR.data(), FlT.data());
    return true;
  }
  mcfl_fl_result res;
  memset(&res, 0, sizeof res);
  mcfl_status st;
  if (o.lanes) {  // the reference's emission loop on its lanes, reseeded for this stage (CUDAMCFLmain.cu:412-413)
    std::vector<uint64_t> x(MCFL_NUM_THREADS);
    std::vector<uint32_t> a(MCFL_NUM_THREADS);
    if (mcfl_init_rng(x.data(), a.data(), MCFL_NUM_THREADS, primes, o.fl_seed)) return false;
    mcfl_rng rng;
    memset(&rng, 0, sizeof rng);
    rng.mode = MCFL_RNG_LANES;
    rng.x = x.data();
    rng.a = a.data();
    rng.n_lanes = MCFL_NUM_THREADS;
    st = mcfl_run_fluorescence_lanes(ctx, s, Fx.data(), s->number_of_photons_per_voxel, &rng, FlR.data(), FlT.data(),
                                     &res);
  } else {
    st = mcfl_run_fluorescence(ctx, s, Fx.data(), s->number_of_photons_per_voxel, o.fl_seed, FlR.data(), FlT.data(),
                               &res);
  }
  if (st != MCFL_OK) {
    printf("Fluorescence failed: %s\n", mcfl_status_string(st));
    return false;
  }
  *fl_photons = res.photons;
  printf("\n\nFlourescence simulation finished! (%llu emitting voxels, kernel %.1f ms)\n",
         (unsigned long long)res.voxels, res.kernel_ms);
  printf("Writing results files...\n");
  mcfl_write_fl(s, FlR.data(), FlT.data());
  return true;
}

void report(const mcfl_sim* s, uint64_t fhd_photons, uint64_t fl_photons, bool fl_ok, clock_t time0) {
  if (fhd_photons == s->number_of_photons && fl_ok)
    printf("All done, no errors! :)\n");
  else
    printf("Simulation finished, some photons were not properly simulated. \n");
  printf("Total time: %.2f sec.\n", (double)(clock() - time0) / CLOCKS_PER_SEC);
  printf("Total simulated photons:\n");
  printf("\t %llu FHD photons.\n", (unsigned long long)fhd_photons);
  printf("\t %llu Fluorescence photons.\n", (unsigned long long)fl_photons);
  printf("#############################################\n\n");
}

// One file: the reference's main (CUDAMCFLmain.cu:318-662).
int run_one(mcfl_ctx* ctx, mcfl_sim* s, const char* primes, const Options& o, clock_t time0) {
  std::vector<uint64_t> x(MCFL_NUM_THREADS);
  std::vector<uint32_t> a(MCFL_NUM_THREADS);
  if (mcfl_init_rng(x.data(), a.data(), MCFL_NUM_THREADS, primes, o.seed)) return 1;
  const size_t nv = nvox(s);
  printf("Running PHD simulation...\n");
  uint64_t fhd_photons = 0, fl_photons = 0;
  if (o.lanes) {  // the reference's DoOneSimulation on its 32768 lanes
    std::vector<double> Fx(nv, 0.0);
    fhd_photons = mcfl_DoOneSimulation(ctx, s, x.data(), a.data(), Fx.data());
    if (fhd_photons == 0) return 1;
    const bool fl_ok = finish_file(ctx, s, Fx, o, primes, &fl_photons);
    report(s, fhd_photons, fl_photons, fl_ok, time0);
    return 0;
  }
  std::vector<uint64_t> fhd(nv), Rd(npix(s)), Tt(npix(s)), tim(ntime(s));
  mcfl_tallies t;
  memset(&t, 0, sizeof t);
  t.Rd_xy = Rd.data();
  t.Tt_xy = Tt.data();
  t.fhd = fhd.data();
  t.time_xyt = tim.empty() ? nullptr : tim.data();
  mcfl_rng rng;
  memset(&rng, 0, sizeof rng);
  rng.mode = MCFL_RNG_PHOTON;
  rng.seed = o.seed;
  clock_t t1 = clock();
  mcfl_status st = mcfl_run_excitation(ctx, s, &rng, &t);
  if (st != MCFL_OK) {
    printf("Excitation failed: %s\n", mcfl_status_string(st));
    return 1;
  }
  double secs = (double)(clock() - t1) / CLOCKS_PER_SEC;
  printf("\nSimulation time: %.2f sec (kernel %.1f ms)\n\n", secs, t.kernel_ms);
  mcfl_write_simulation_results(s, &t, secs);
  fhd_photons = t.num_terminated_photons;
  std::vector<double> Fx = fx_of(s, fhd, fhd_photons);
  const bool fl_ok = finish_file(ctx, s, Fx, o, primes, &fl_photons);
  report(s, fhd_photons, fl_photons, fl_ok, time0);
  return 0;
}

// Same device medium: grid, labels or layers, materials, detector and source weight.
bool same_medium(const mcfl_sim* a, const mcfl_sim* b) {
  if (a->bulk_method != b->bulk_method || a->nx != b->nx || a->ny != b->ny || a->nz != b->nz ||
      a->grid_size != b->grid_size || a->esp != b->esp || a->halfx != b->halfx || a->halfy != b->halfy ||
      a->n_bulks != b->n_bulks || a->n_layers != b->n_layers || a->do_temp_sim != b->do_temp_sim ||
      a->start_weight != b->start_weight || a->dir != b->dir)
    return false;
  if (memcmp(&a->det, &b->det, sizeof a->det) != 0) return false;
  if (a->bulk_method == 1)  // layered slab with one inclusion
    return memcmp(a->layers, b->layers, sizeof(mcfl_layer) * (a->n_layers + 2)) == 0 &&
           memcmp(&a->inclusion, &b->inclusion, sizeof a->inclusion) == 0;
  if (memcmp(a->bulks, b->bulks, sizeof(mcfl_bulk) * (a->n_bulks + 2)) != 0) return false;
  return memcmp(a->bulk_info, b->bulk_info, nvox(a)) == 0;
}

// Several files: batched excitation, then each file's outputs.
int run_many(mcfl_ctx* ctx, std::vector<mcfl_sim*>& sims, const std::vector<std::string>& names, const char* primes,
             const Options& o, clock_t time0) {
  const int n = (int)sims.size();
  std::vector<int> single;
  std::vector<int> todo;
  // files share batched launches (a batch splits a file above the photon
  // split of mcfl_run_excitation the same way, so the results are a single
  // run's; a layered medium shares a launch only with files on the same
  // medium); lane streams and files above a batch job's 2^32 - 1 photons
  // run alone
  for (int i = 0; i < n; ++i)
    (!o.lanes && sims[i]->number_of_photons <= 0xFFFFFFFFull ? todo : single).push_back(i);
  int rc = 0;
  size_t next = 0;
  while (next < todo.size()) {
    // one batch: at most MCFL_MAX_BATCH_JOBS jobs and 320 material entries
    std::vector<const mcfl_sim*> media;
    std::vector<int> medium_of, files;
    int mats = 0;
    while (next < todo.size() && (int)files.size() < MCFL_MAX_BATCH_JOBS) {
      const mcfl_sim* s = sims[todo[next]];
      int m = -1;
      for (size_t k = 0; k < media.size(); ++k)
        if (same_medium(media[k], s)) m = (int)k;
      if (m < 0) {
        const int n_mat = s->bulk_method == 1 ? (int)s->n_layers + 3 : (int)s->n_bulks + 2;
        if (!media.empty() && (mats + n_mat > 320 || s->bulk_method == 1 || media[0]->bulk_method == 1)) break;
        media.push_back(s);
        mats += n_mat;
        m = (int)media.size() - 1;
      }
      medium_of.push_back(m);
      files.push_back(todo[next++]);
    }
    const int nj = (int)files.size();
    std::vector<mcfl_job> jobs(nj);
    std::vector<std::vector<uint64_t>> Rd(nj), Tt(nj), fhd(nj), tim(nj);
    for (int j = 0; j < nj; ++j) {
      mcfl_sim* s = sims[files[j]];
      mcfl_job& J = jobs[j];
      memset(&J, 0, sizeof J);
      J.medium = medium_of[j];
      J.src_kind = s->dir == 0.0f ? MCFL_SOURCE_ISOTROPIC : s->src_kind;
      J.xi = s->xi;
      J.yi = s->yi;
      J.zi = s->zi;
      J.src_r = s->src_r;
      J.photons = s->number_of_photons;
      J.seed = o.seed;
      Rd[j].resize(npix(s));
      Tt[j].resize(npix(s));
      J.Rd_xy = Rd[j].data();
      J.Tt_xy = Tt[j].data();
      if (s->fhd_activated == 1) {
        J.tallies |= MCFL_TALLY_FHD;
        fhd[j].resize(nvox(s));
        J.fhd = fhd[j].data();
      }
      if (s->do_temp_sim == 1) {
        J.tallies |= MCFL_TALLY_TEMP;
        tim[j].resize(ntime(s));
        J.time_xyt = tim[j].data();
      }
    }
    printf("Running %d excitation simulations (%zu media) in one batch...\n", nj, media.size());
    clock_t t1 = clock();
    double ms = 0.0;
    mcfl_status st = mcfl_run_batch(ctx, media.data(), (int)media.size(), jobs.data(), nj, &ms);
    if (st != MCFL_OK) {
      printf("Batch failed: %s\n", mcfl_status_string(st));
      return 1;
    }
    const double secs = (double)(clock() - t1) / CLOCKS_PER_SEC;
    printf("\nBatch time: %.2f sec (kernel %.1f ms)\n\n", secs, ms);
    for (int j = 0; j < nj; ++j) {
      mcfl_sim* s = sims[files[j]];
      printf("\n%s\n", names[files[j]].c_str());
      mcfl_tallies t;
      memset(&t, 0, sizeof t);
      t.Rd_xy = Rd[j].data();
      t.Tt_xy = Tt[j].data();
      t.fhd = fhd[j].empty() ? nullptr : fhd[j].data();
      t.time_xyt = tim[j].empty() ? nullptr : tim[j].data();
      memcpy(t.totals, jobs[j].totals, sizeof t.totals);
      t.num_terminated_photons = jobs[j].totals[MCFL_TOT_PHOTONS];
      mcfl_write_simulation_results(s, &t, secs);
      uint64_t fl_photons = 0;
      std::vector<double> Fx = fx_of(s, fhd[j], t.num_terminated_photons);
      const bool fl_ok = finish_file(ctx, s, Fx, o, primes, &fl_photons);
      report(s, t.num_terminated_photons, fl_photons, fl_ok, time0);
    }
  }
  for (int i : single) {
    printf("\n%s\n", names[i].c_str());
    rc |= run_one(ctx, sims[i], primes, o, time0);
  }
  return rc;
}

}  // namespace

int main(int argc, char* argv[]) {
  printf("\nmcfl (MI355X). Build date: %s, %s. \n", __DATE__, __TIME__);
  clock_t time0 = clock();
  if (argc < 2) {
    printf("Not enough input arguments!\n");
    return 1;
  }
  Options o;
  o.seed = (unsigned long long)time(NULL);
  o.fl_seed = (unsigned long long)time(NULL);
  // .mci files first, then flags; the build's own flags are taken out before
  // the reference's interpreter sees the rest
  std::vector<std::string> files;
  std::vector<char*> args{argv[0]};
  int i = 1;
  for (; i < argc && argv[i][0] != '-'; ++i) files.push_back(argv[i]);
  if (files.empty()) {  // the reference takes argv[1] as the input file whatever it is (CUDAMCFLmain.cu:305-310)
    files.push_back(argv[1]);
    i = 2;
  }
  args.push_back(argv[1]);
  for (; i < argc; ++i) {
    if (!strcmp(argv[i], "-L")) {
      o.lanes = true;
      continue;
    }
    if (!strncmp(argv[i], "-F", 2)) {
      o.fl_seed = strtoull(argv[i] + 2, nullptr, 10);
      continue;
    }
    if (!strncmp(argv[i], "-I", 2)) {
      o.is_photons = strtoull(argv[i] + 2, nullptr, 10);
      continue;
    }
    args.push_back(argv[i]);
  }
  uint64_t s64 = o.seed;
  if (mcfl_interpret_arg((int)args.size(), args.data(), &s64, &o.ignoreA)) return 1;
  o.seed = s64;

  std::vector<mcfl_sim*> sims;
  std::vector<mcfl_sim*> owned;
  std::vector<int> counts;
  for (const std::string& f : files) {
    printf("\nExecuting %s... \n", f.c_str());
    printf("____________________________________________________________________\n\n");
    mcfl_sim* s = nullptr;
    int nsim = mcfl_read_simulation_data(f.c_str(), &s, o.ignoreA);
    if (nsim == 0) {
      printf("Something wrong with read_simulation_data!\n");
      for (size_t k = 0; k < owned.size(); ++k) mcfl_free_simulations(owned[k], counts[k]);
      return 1;
    }
    printf("\nRead %d simulations\n\n", nsim);
    owned.push_back(s);
    counts.push_back(nsim);
    sims.push_back(&s[0]);  // only the first run is simulated (CUDAMCFLmain.cu:338-366)
  }

  const char* primes = "safeprimes_base32.txt";
  FILE* pf = fopen(primes, "r");
  if (pf)
    fclose(pf);
  else
    primes = mcfl_default_primes_path();

  mcfl_ctx* ctx = nullptr;
  mcfl_status st = mcfl_ctx_create(0, primes, &ctx);
  int rc = 1;
  if (st != MCFL_OK) {
    printf("No usable GPU: %s\n", mcfl_status_string(st));
  } else {
    rc = sims.size() == 1 ? run_one(ctx, sims[0], primes, o, time0) : run_many(ctx, sims, files, primes, o, time0);
    mcfl_ctx_destroy(ctx);
  }
  for (size_t k = 0; k < owned.size(); ++k) mcfl_free_simulations(owned[k], counts[k]);
  return rc;
}
