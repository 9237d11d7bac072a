// Parser/writer driver for the sanitizer test: parses each .mci argument with
// mcfl_read_simulation_data (CUDAMCFLio.cu:243-659), writes the output files
// of tallies of the parsed shapes, and frees the result.
// Built by tests/test_host_sanitized.py with AddressSanitizer and
// UndefinedBehaviorSanitizer (host code only; no device code is involved).
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "mcfl.h"
int main(int argc, char** argv) {
  bool write = false;  // "-w": also write the output files of the files that follow
  for (int i = 1; i < argc; ++i) {
    if (argv[i][0] == '-' && argv[i][1] == 'w') {
      write = true;
      continue;
    }
    mcfl_sim* s = nullptr;
    int n = mcfl_read_simulation_data(argv[i], &s, 0);
    printf("%s: %d runs\n", argv[i], n);
    if (n > 0) {
      const mcfl_sim& m = s[0];
      printf("  grid %d %d %d bulks %u photons %llu\n", m.nx, m.ny, m.nz, m.n_bulks,
             (unsigned long long)m.number_of_photons);
      // the writers on tallies of the parsed shapes (Write_Simulation_Results, PHD, FlR/FlT)
      // the parser, like the reference's, takes any integers; the run-time
      // entry points refuse the shapes skipped here
      const bool ok = m.det.nx > 0 && m.det.ny > 0 && m.det.nx < 100000 && m.det.ny < 100000 && m.nx > 0 &&
                      m.ny > 0 && m.nz > 0 && (double)m.nx * m.ny * m.nz < 1e9 &&
                      (double)m.det.x_temp_numdets * m.det.y_temp_numdets * m.det.temp_bins < 1e8;
      if (!ok || !write) {
        mcfl_free_simulations(s, n);
        continue;
      }
      const size_t npix = (size_t)m.det.nx * (m.det.ny + 1), nvox = (size_t)m.nx * m.ny * m.nz;
      const size_t ntime = m.do_temp_sim ? (size_t)m.det.x_temp_numdets * m.det.y_temp_numdets * m.det.temp_bins : 0;
      std::vector<uint64_t> R(npix, 3), T(npix, 5), F(nvox, 7), tim(ntime + 1, 1);
      std::vector<double> Fx(nvox, 1e-6), FlR((size_t)m.det.nx * m.det.ny, 2e-9), FlT(FlR.size(), 1e-9);
      mcfl_tallies t = {};
      t.Rd_xy = R.data();
      t.Tt_xy = T.data();
      t.fhd = F.data();
      t.time_xyt = ntime ? tim.data() : nullptr;
      t.num_terminated_photons = m.number_of_photons;
      mcfl_write_simulation_results(&m, &t, 1.0);
      mcfl_write_phd(&m, Fx.data());
      mcfl_write_fl(&m, FlR.data(), FlT.data());
      mcfl_free_simulations(s, n);
    }
  }
  return 0;
}
