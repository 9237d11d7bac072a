// Parser driver for the sanitizer test: parses each .mci argument with
// mcfl_read_simulation_data (CUDAMCFLio.cu:243-659) and frees the result.
// Built by tests/test_host_sanitized.py with AddressSanitizer and
// UndefinedBehaviorSanitizer (host code only; no device code is involved).
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "mcfl.h"
int main(int argc, char** argv) {
  for (int i = 1; i < argc; ++i) {
    mcfl_sim* s = nullptr;
    int n = mcfl_read_simulation_data(argv[i], &s, 0);
    printf("%s: %d runs\n", argv[i], n);
    if (n > 0) {
      printf("  grid %d %d %d bulks %u photons %llu\n", s[0].nx, s[0].ny, s[0].nz, s[0].n_bulks, (unsigned long long)s[0].number_of_photons);
      mcfl_free_simulations(s, n);
    }
  }
  return 0;
}
