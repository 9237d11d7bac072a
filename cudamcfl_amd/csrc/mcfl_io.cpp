// mcfl_io.cpp -- .mci parser, command-line arguments and output writers.
//
// Restates CUDAMCFLio.cu (read_simulation_data :243-659, readfloats :199-216,
// readints :218-235, interpret_arg :30-55, Write_Simulation_Results :57-191)
// and the PHD / fluorescence writers of CUDAMCFLmain.cu:368-392,604-647, with
// the same grammar, derived optics, file names and byte layouts.  Differences
// (all on paths where the reference has undefined behaviour) are noted inline.
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "../../include/mcfl.h"

namespace {

constexpr int kNFloats = 12;  // CUDAMCFLio.cu:17
constexpr int kNInts = 5;     // CUDAMCFLio.cu:18

bool is_letter(char c) { return (c >= 'A' && c <= 'Z') || (c >= 'a' && c <= 'z'); }

// readfloats: next line with at least one number; more than n numbers fails.
bool read_floats(int n, float* v, FILE* f) {
  if (n > kNFloats) return false;
  int got = 0;
  char line[200];
  while (got <= 0) {
    if (feof(f)) return false;
    if (!fgets(line, sizeof line, f)) line[0] = 0;
    memset(v, 0, kNFloats * sizeof(float));
    got = sscanf(line, "%f %f %f %f %f %f %f %f %f %f %f %f", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6], &v[7],
                 &v[8], &v[9], &v[10], &v[11]);
    if (got > n) return false;
  }
  return true;
}

bool read_ints(int n, int* v, FILE* f) {
  if (n > kNInts) return false;
  int got = 0;
  char line[MCFL_STR_LEN];
  while (got <= 0) {
    if (feof(f)) return false;
    if (!fgets(line, sizeof line, f)) line[0] = 0;
    memset(v, 0, kNInts * sizeof(int));
    got = sscanf(line, "%d %d %d %d %d", &v[0], &v[1], &v[2], &v[3], &v[4]);
    if (got > n) return false;
  }
  return true;
}

int fail(const char* what) {
  fprintf(stderr, "%s\n", what);
  return 0;
}

// Derived optics of one medium line "n mua mus flc AbsCExi AbsCEmi eY g"
// (CUDAMCFLio.cu:560-568).
void set_bulk(mcfl_bulk* b, const float* t) {
  b->n = t[0];
  b->mua = t[1] + t[3] * t[4];
  b->muaf = t[1] + t[3] * t[5];
  b->eY = t[6];
  b->flc = t[3];
  b->g = t[7];
  b->mutr = (t[2] == 0.0f) ? FLT_MAX : 1.0f / (t[1] + t[2]);
  b->albedof = t[2] / (t[3] * t[4] + t[2]);
}

void set_outer(mcfl_bulk* b, float n) {  // CUDAMCFLio.cu:546-555 / :585-594
  b->n = n;
  b->mua = 0.f;
  b->muaf = 0.f;
  b->eY = 0.f;
  b->flc = 0.f;
  b->g = 1.f;
  b->mutr = FLT_MAX;
  b->albedof = 0.f;
}

// Reads the bulk-descriptor file character by character (CUDAMCFLio.cu:618-636).
bool read_bulk_file(mcfl_sim* s, size_t nvox) {
  s->bulk_info = static_cast<uint8_t*>(malloc(nvox ? nvox : 1));
  if (!s->bulk_info) return false;
  std::ifstream in(s->bulkinfo_filename);
  if (!in) {
    fprintf(stderr, "Error opening 3D bulk file %s\n", s->bulkinfo_filename);
    return false;
  }
  size_t count = 0;
  bool bad_label = false;
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream iss(line);
    char c;
    while (iss >> c) {
      int v = c - '0';
      // The reference stores one element past the end when the file is too
      // long (read_count <= fhd_size); this build stops at the end.
      if (count < nvox) {
        if (v < 0 || v > 255) bad_label = true;
        s->bulk_info[count] = static_cast<uint8_t>(v < 0 ? 255 : (v > 255 ? 255 : v));
      }
      ++count;
    }
  }
  printf("Number of voxels of bulk descritor file: %zu\n", count);
  if (count != nvox) {
    printf("Error reading 3D bulk file: size mismatch.\n");
    // The reference continues with uninitialised labels; this build fills
    // the missing ones with medium 1 (the background).
    for (size_t i = count; i < nvox; ++i) s->bulk_info[i] = 1;
  }
  if (bad_label) fprintf(stderr, "warning: bulk file holds non-digit labels\n");
  return true;
}

}  // namespace

extern "C" int mcfl_interpret_arg(int argc, char** argv, uint64_t* seed, int* ignoreAdetection) {
  for (int i = 2; i < argc; i++) {
    int unknown = 1;
    if (!strcmp(argv[i], "-A")) {
      unknown = 0;
      *ignoreAdetection = 1;
      printf("Ignoring A-detection!\n");
    }
    unsigned long long sv;
    if (!strncmp(argv[i], "-S", 2) && sscanf(argv[i], "%*2c %llu", &sv)) {
      *seed = sv;
      unknown = 0;
      printf("Seed=%llu\n", sv);
    }
    if (unknown) {
      printf("Unknown argument %s!\n", argv[i]);
      return 1;
    }
  }
  return 0;
}

extern "C" void mcfl_free_simulations(mcfl_sim* sims, int n) {
  if (!sims) return;
  for (int i = 0; i < n; ++i) {
    free(sims[i].layers);
    free(sims[i].bulks);
    free(sims[i].bulk_info);
  }
  free(sims);
}

extern "C" int mcfl_read_simulation_data(const char* filename, mcfl_sim** simulations, int ignoreAdetection) {
  (void)ignoreAdetection;  // "-A" has no effect in the reference either (CUDAMCFLio.cu:40)
  *simulations = nullptr;
  FILE* f = fopen(filename, "r");
  if (!f) {
    perror("Error opening file");
    return 0;
  }
  float ft[kNFloats];
  int it[kNInts];
  auto bail = [&](const char* msg, mcfl_sim* sims, int n) {
    fclose(f);
    mcfl_free_simulations(sims, n);
    *simulations = nullptr;
    return fail(msg);
  };
  if (!read_floats(1, ft, f)) return bail("Error reading file version", nullptr, 0);
  if (!read_ints(1, it, f)) return bail("Error reading number of runs", nullptr, 0);
  const int nsim = it[0];
  if (nsim <= 0) return bail("Error reading number of runs", nullptr, 0);
  mcfl_sim* sims = static_cast<mcfl_sim*>(calloc((size_t)nsim, sizeof(mcfl_sim)));
  if (!sims) return bail("Failed to malloc simulations.", nullptr, 0);

  for (int i = 0; i < nsim; ++i) {
    mcfl_sim* s = &sims[i];
    snprintf(s->inp_filename, MCFL_STR_LEN, "%s", filename);
    if (!read_ints(1, it, f)) return bail("Error bulk description method", sims, nsim);
    s->bulk_method = it[0];

    // output file name and A|B (CUDAMCFLio.cu:290-301)
    char line[MCFL_STR_LEN], str[MCFL_STR_LEN];
    char AorB = 0;
    int ii = 0;
    while (ii <= 0) {
      s->begin = ftell(f);
      if (!fgets(line, sizeof line, f)) line[0] = 0;
      ii = sscanf(line, "%499s %c/n", str, &AorB);
      if (feof(f) || ii > 2) return bail("Error reading output filename", sims, nsim);
      if (ii > 0) ii = is_letter(str[0]);
    }
    snprintf(s->outp_filename, MCFL_STR_LEN, "%s", str);
    s->AorB = AorB;

    // photons (CUDAMCFLio.cu:304-317)
    unsigned long long nph = 0;
    ii = 0;
    while (ii <= 0) {
      if (!fgets(line, sizeof line, f)) line[0] = 0;
      nph = 0;
      ii = sscanf(line, "%llu", &nph);
      if (feof(f) || ii > 1) return bail("Error reading number of photons", sims, nsim);
    }
    if (nph <= (unsigned long long)MCFL_NUM_THREADS) {
      printf("WARNING: minimum number of photons that can be simulated is NUM_THREADS. Setting excitation photons to: %i\n\n",
             MCFL_NUM_THREADS - 1);
      nph = MCFL_NUM_THREADS - 1;
    }
    s->number_of_photons = nph;

    if (!read_ints(1, it, f)) return bail("Error reading number of photons per voxel", sims, nsim);
    if (it[0] <= MCFL_NUM_THREADS) {
      printf("WARNING: minimum number of photons that can be simulated is NUM_THREADS. Setting photons per voxel to: %i\n\n",
             MCFL_NUM_THREADS - 1);
      s->number_of_photons_per_voxel = MCFL_NUM_THREADS - 1;
    } else {
      s->number_of_photons_per_voxel = (uint32_t)it[0];
    }

    if (!read_ints(1, it, f)) return bail("Error reading grid size", sims, nsim);
    s->grid_size = it[0];

    if (!read_ints(1, it, f)) return bail("Error reading fluorescence simulaion flag", sims, nsim);
    s->do_fl_sim = it[0];
    s->fhd_activated = (it[0] == 1 || it[0] == 2) ? 1 : 0;
    if (s->do_fl_sim != 0 && s->do_fl_sim != 1 && s->do_fl_sim != 2)
      return bail("Error reading fluorescence simulaion flag", sims, nsim);

    if (!read_floats(2, ft, f)) return bail("Error reading dx and dy", sims, nsim);
    s->det.dx = ft[0];
    s->det.dy = ft[1];
    if (!read_ints(2, it, f)) return bail("Error reading No. of nx and ny", sims, nsim);
    s->det.nx = it[0];
    s->det.ny = it[1];
    if (!read_floats(2, ft, f)) return bail("Error reading detector x0 and y0", sims, nsim);
    s->det.x0 = ft[0];
    s->det.y0 = ft[1];
    if (!read_floats(3, ft, f)) return bail("Error source position", sims, nsim);
    s->xi = ft[0];
    s->yi = ft[1];
    s->zi = ft[2];

    if (!read_ints(1, it, f)) return bail("Error reading temporal simulaion flag", sims, nsim);
    s->do_temp_sim = it[0];
    if (s->do_temp_sim != 0 && s->do_temp_sim != 1) return bail("Error reading temporal simulaion flag", sims, nsim);
    if (!read_floats(2, ft, f)) return bail("Error reading temporal detector array center", sims, nsim);
    s->det.x0_temp_det = ft[0];
    s->det.y0_temp_det = ft[1];
    if (!read_ints(2, it, f)) return bail("Error reading No. of temporal detectors", sims, nsim);
    s->det.x_temp_numdets = (uint32_t)it[0];
    s->det.y_temp_numdets = (uint32_t)it[1];
    if (!read_floats(2, ft, f)) return bail("Error reading temporal detector separation", sims, nsim);
    s->det.x_temp_sepdets = ft[0];
    s->det.y_temp_sepdets = ft[1];
    if (!read_floats(1, ft, f)) return bail("Error temporal detector radious", sims, nsim);
    s->det.temp_det_r = ft[0];
    if (!read_ints(1, it, f)) return bail("Error reading location of detector array flag", sims, nsim);
    s->det.temp_rort = (uint32_t)it[0];
    if (it[0] != 0 && it[0] != 1) return bail("Error reading location of detector array flag", sims, nsim);
    if (!read_ints(1, it, f)) return bail("Error reading No. of temporal bins", sims, nsim);
    s->det.temp_bins = (uint64_t)(int64_t)it[0];
    if (!read_ints(1, it, f)) return bail("Error reading max temp", sims, nsim);
    s->det.max_temp = (uint64_t)(int64_t)it[0];

    double n1 = 1.0, n2 = 1.0;
    if (s->bulk_method == 1) {  // CUDAMCFLio.cu:427-526
      if (!read_ints(1, it, f)) return bail("Error reading No. of layers", sims, nsim);
      const int nl = it[0];
      if (nl < 0 || nl > 98) return bail("Error reading No. of layers", sims, nsim);
      s->n_layers = (uint32_t)nl;
      s->layers = static_cast<mcfl_layer*>(calloc((size_t)nl + 2, sizeof(mcfl_layer)));
      if (!s->layers) return bail("Failed to malloc layers.", sims, nsim);
      if (!read_floats(1, ft, f)) return bail("Error reading upper refractive index", sims, nsim);
      s->layers[0].n = ft[0];
      float dtot = 0.f;
      for (int l = 1; l <= nl; ++l) {
        if (!read_floats(9, ft, f)) return bail("Error reading layer data", sims, nsim);
        mcfl_layer* L = &s->layers[l];
        L->n = ft[0];
        L->mua = ft[1] + ft[3] * ft[4];
        L->muaf = ft[1] + ft[3] * ft[5];
        L->eY = ft[6];
        L->flc = ft[3];
        L->g = ft[7];
        L->z_min = dtot;
        dtot += ft[8];
        L->z_max = dtot;
        L->mutr = (ft[2] == 0.0f) ? FLT_MAX : 1.0f / (ft[1] + ft[2]);
        L->albedof = ft[2] / (ft[3] * ft[4] + ft[2]);
      }
      s->esp = dtot;
      if (!read_floats(1, ft, f)) return bail("Error reading lower refractive index", sims, nsim);
      s->layers[nl + 1].n = ft[0];
      if (!read_floats(12, ft, f)) return bail("Error leyendo datos de inclusion", sims, nsim);
      mcfl_inclusion* I = &s->inclusion;
      I->x = ft[0];
      I->y = ft[1];
      I->z = ft[2];
      I->r = ft[3];
      I->n = ft[4];
      I->mua = ft[5] + ft[7] * ft[8];
      I->muaf = ft[5] + ft[7] * ft[9];
      I->flc = ft[7];
      I->eY = ft[10];
      I->g = ft[11];
      I->mutr = (ft[6] == 0.0f) ? FLT_MAX : 1.0f / (ft[5] + ft[6]);
      I->albedof = ft[6] / (ft[7] * ft[8] + ft[6]);
      n1 = s->layers[0].n;
      n2 = s->layers[1].n;
      // photon hitting density grid, as for method 2 (CUDAMCFLmem.cu:21-23)
      s->nx = (int)(4 * s->esp * s->grid_size);
      s->ny = (int)(4 * s->esp * s->grid_size);
      s->nz = (int)(s->esp * s->grid_size);
      s->halfx = 2 * s->esp;
      s->halfy = 2 * s->esp;
      s->n_bulks = 0;
      s->bulks = static_cast<mcfl_bulk*>(calloc(1, sizeof(mcfl_bulk)));
      s->bulk_info = static_cast<uint8_t*>(calloc(1, 1));
    } else if (s->bulk_method == 2) {  // CUDAMCFLio.cu:528-646
      if (!read_ints(1, it, f)) return bail("Error reading No. of bulk descriptors", sims, nsim);
      const int nb = it[0];
      if (nb < 1 || nb > 98) return bail("Error reading No. of bulk descriptors", sims, nsim);
      s->n_bulks = (uint32_t)nb;
      s->bulks = static_cast<mcfl_bulk*>(calloc((size_t)nb + 2, sizeof(mcfl_bulk)));
      if (!s->bulks) return bail("Failed to malloc bulk descriptors.", sims, nsim);
      if (!read_floats(1, ft, f)) return bail("Error bulk thickness", sims, nsim);
      s->esp = ft[0];
      if (!read_floats(1, ft, f)) return bail("Error reading outer refractive index", sims, nsim);
      set_outer(&s->bulks[0], ft[0]);
      for (int b = 1; b <= nb; ++b) {
        if (!read_floats(8, ft, f)) return bail("Error reading bulk descriptors data", sims, nsim);
        set_bulk(&s->bulks[b], ft);
      }
      if (!read_floats(1, ft, f)) return bail("Error reading lower refractive index", sims, nsim);
      set_outer(&s->bulks[nb + 1], ft[0]);
      ii = 0;
      while (ii <= 0) {  // CUDAMCFLio.cu:597-604
        if (feof(f)) return bail("Error bulk matrix filename", sims, nsim);
        if (!fgets(line, sizeof line, f)) line[0] = 0;
        ii = sscanf(line, "%499s", str);
      }
      snprintf(s->bulkinfo_filename, MCFL_STR_LEN, "%s", str);
      printf("Bulk Descriptor file: %s\n", s->bulkinfo_filename);
      // Grid (CUDAMCFLio.cu:609-611), or the dims extension line that may follow
      // the filename (a comment to the reference parser, which stops here).
      s->nx = (int)(4 * s->esp * s->grid_size);
      s->ny = (int)(4 * s->esp * s->grid_size);
      s->nz = (int)(s->esp * s->grid_size);
      s->halfx = 2 * s->esp;
      s->halfy = 2 * s->esp;
      long here = ftell(f);
      int ex, ey, ez;
      if (fgets(line, sizeof line, f) && sscanf(line, " # mcfl-dims %d %d %d", &ex, &ey, &ez) == 3 && ex > 0 &&
          ey > 0 && ez > 0) {
        s->nx = ex;
        s->ny = ey;
        s->nz = ez;
        s->halfx = (float)ex / (2.0f * (float)s->grid_size);
        s->halfy = (float)ey / (2.0f * (float)s->grid_size);
      } else {
        fseek(f, here, SEEK_SET);
      }
      if (s->nx <= 0 || s->ny <= 0 || s->nz <= 0) return bail("Error: empty voxel grid", sims, nsim);
      if (!read_bulk_file(s, (size_t)s->nx * (size_t)s->ny * (size_t)s->nz))
        return bail("Error reading 3D bulk file", sims, nsim);
      n1 = s->bulks[0].n;
      n2 = s->bulks[1].n;
      s->layers = static_cast<mcfl_layer*>(calloc(1, sizeof(mcfl_layer)));
    } else {
      return bail("Error bulk description method", sims, nsim);
    }
    // start weight: specular reflection at normal incidence (CUDAMCFLio.cu:648-652)
    double r = (n1 - n2) / (n1 + n2);
    r = r * r;
    s->start_weight = (unsigned int)((double)0xFFFFFFFF * (1 - r));
    s->dir = 1.0f;
    s->src_kind = MCFL_SOURCE_PENCIL;
    s->end = ftell(f);
  }
  fclose(f);
  *simulations = sims;
  return nsim;
}

// ----------------------------------------------------------------- writers

namespace {
FILE* open_out(const mcfl_sim* s, const char* suffix, const char* what) {
  char name[MCFL_STR_LEN + 32];
  snprintf(name, sizeof name, "%s%s", s->outp_filename, suffix);
  FILE* f = fopen(name, "w");
  if (!f) perror(what);
  return f;
}

void write_image(FILE* f, const uint64_t* img, int nx, int ny, double scale) {
  for (int y = 0; y < ny; y++) {
    for (int x = 0; x < nx; x++) fprintf(f, " %E ", (double)img[y * nx + x] / scale);
    fprintf(f, " \n ");
  }
}
}  // namespace

extern "C" int mcfl_write_simulation_results(const mcfl_sim* sim, const mcfl_tallies* t, double sim_time_s) {
  const double dx = (double)sim->det.dx, dy = (double)sim->det.dy;
  const int nx = sim->det.nx, ny = sim->det.ny;
  const double scale1 = (double)0xFFFFFFFF * (double)sim->number_of_photons;  // requested N (CUDAMCFLio.cu:82)
  FILE* in = fopen(sim->inp_filename, "r");
  if (!in) {
    perror("Error opening input file");
    return 1;
  }
  FILE* pf = open_out(sim, "_param.dat", "Error opening parameters output file");
  if (!pf) {
    fclose(in);
    return 1;
  }
  fprintf(pf, "Compilation date: %s, %s. \n\n", __DATE__, __TIME__);
  fprintf(pf, "A1 \t# Version number of the file format.\n\n");
  fprintf(pf, "####\n");
  fprintf(pf, "# Data categories include: \n");
  fprintf(pf, "# InParm, RAT, \n");
  fprintf(pf, "# A_l, A_z, Rd_r, Rd_a, Tt_r, Tt_a, \n");
  fprintf(pf, "# A_rz, Rd_ra, Tt_ra \n");
  fprintf(pf, "####\n\n");
  fprintf(pf, "# User time: %.2f sec\n\n", sim_time_s);
  fprintf(pf, "InParam\t\t# Input parameters:\n");
  fseek(in, (long)sim->begin, SEEK_SET);
  char buf[MCFL_STR_LEN];
  while (sim->end > ftell(in)) {
    if (!fgets(buf, sizeof buf, in)) break;
    fputs(buf, pf);
  }
  fclose(in);
  fclose(pf);

  FILE* tf = open_out(sim, "_Trans.dat", "Error opening transmission output file");
  if (!tf) return 1;
  write_image(tf, t->Tt_xy, nx, ny, scale1 * dx * dy);
  fclose(tf);
  FILE* rf = open_out(sim, "_Refl.dat", "Error opening reflection output file");
  if (!rf) return 1;
  write_image(rf, t->Rd_xy, nx, ny, scale1 * dx * dy);
  fclose(rf);

  if (sim->do_temp_sim == 1 && t->time_xyt) {
    FILE* mf = open_out(sim, "_time.dat", "Error opening time output file");
    if (!mf) return 1;
    const int xt = (int)sim->det.x_temp_numdets, yt = (int)sim->det.y_temp_numdets;
    const long tt = (long)sim->det.temp_bins;
    fprintf(mf, "#x y timebins[numbins]\n");
    for (int ix = 0; ix < xt; ix++)
      for (int iy = 0; iy < yt; iy++) {
        // detector positions as DoOneSimulation computes them (CUDAMCFLmain.cu:95-100)
        float px = ix * sim->det.x_temp_sepdets - ((xt - 1) * sim->det.x_temp_sepdets) / 2 + sim->det.x0_temp_det;
        float py = iy * sim->det.y_temp_sepdets - ((yt - 1) * sim->det.y_temp_sepdets) / 2 + sim->det.y0_temp_det;
        fprintf(mf, "%f %f ", px, py);
        for (long it = 0; it < tt; it++) fprintf(mf, " %E ", (double)t->time_xyt[ix + xt * (iy + it * yt)] / scale1);
        fprintf(mf, " \n ");
      }
    fclose(mf);
  }
  return 0;
}

extern "C" int mcfl_write_phd(const mcfl_sim* sim, const double* Fx) {
  FILE* f = open_out(sim, "_PHD-Ascii.dat", "Error opening output file");
  if (!f) return 1;
  fprintf(f, "%llu\t%llu\t%llu\n", (unsigned long long)sim->nx, (unsigned long long)sim->ny,
          (unsigned long long)sim->nz);
  const size_t n = (size_t)sim->nx * sim->ny * sim->nz;
  for (size_t v = 0; v < n; ++v) fprintf(f, "%.10E\n", Fx[v]);
  fclose(f);
  return 0;
}

extern "C" int mcfl_write_fl(const mcfl_sim* sim, const double* FlR, const double* FlT) {
  const int nx = sim->det.nx, ny = sim->det.ny;
  const char* suf[2] = {"_FlR.dat", "_FlT.dat"};
  const double* img[2] = {FlR, FlT};
  for (int k = 0; k < 2; ++k) {
    FILE* f = open_out(sim, suf[k], "Error opening output file");
    if (!f) return 1;
    for (int y = 0; y < ny; y++) {
      for (int x = 0; x < nx; x++) fprintf(f, " %E ", img[k][y * nx + x]);
      fprintf(f, " \n ");
    }
    fclose(f);
  }
  return 0;
}
