// mcfl_host.cpp -- device contexts, RNG initialisation and the run entry
// points of the C ABI (include/mcfl.h).
//
// Replaces the host halves of CUDAMCFLmain.cu (DoOneSimulation :60-193, the
// fluorescence voxel loop :409-651), CUDAMCFLmem.cu (allocation, constant
// upload and copy-back, :17-217) and CUDAMCFLrng.cu:31-90 (init_RNG).  Device
// buffers live in the context and are reused across runs; one simulation is
// one persistent kernel launch with no host polling.
#include <dlfcn.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <new>
#include <string>
#include <vector>

#include "mcfl_internal.h"
#include "mcfl_mwc.h"

#define HIPCHK MCFL_HIPCHK
#define STCHK MCFL_STCHK

namespace mcfl_detail {

mcfl_status ensure(DevBuf& b, size_t bytes) {
  if (bytes == 0) bytes = 8;
  if (b.cap >= bytes) return MCFL_OK;
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.cap = 0;
  if (hipMalloc(&b.p, bytes) != hipSuccess) return MCFL_ERR_NOMEM;
  b.cap = bytes;
  return MCFL_OK;
}

void release(DevBuf& b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.cap = 0;
}

// Per-material constants of the kernel (mcfl_device.h MatDev).  emission:
// mua <- muaf (CUDAMCFLmain.cu:458-461).
static void mat_of(float mutr, float mua, float g, float n, MatDev& m) {
  const float g2 = g * g;
  m.mutr = mutr;
  m.mua_mutr = mua * mutr;
  m.n = n;
  m.onemg = 1.0f - g;
  m.onemg2 = 1.0f - g2;
  m.onepg2 = 1.0f + g2;
  m.twog32 = 2.0f * g * 0x1p-32f;
  m.inv2g = (g != 0.0f) ? 1.0f / (2.0f * g) : 0.0f;
}

// Reflect's quotients against the outer media (material 0 above, `last` below).
static void surface_quotients(std::vector<MatDev>& out, uint32_t last) {
  const float n0 = out[0].n, nl = out[last].n;
  for (MatDev& m : out) {
    m.q_top = (m.n * m.n) / (n0 * n0);
    m.r_top = m.n / n0;
    m.q_bot = (m.n * m.n) / (nl * nl);
    m.r_bot = m.n / nl;
  }
}

// Layered medium (bulk method 1): layers 0..n_layers+1, then the inclusion.
void build_mats(const mcfl_sim* s, bool emission, std::vector<MatDev>& out) {
  if (s->bulk_method == 1) {
    const uint32_t n = s->n_layers + 2;
    out.resize(n + 1);
    for (uint32_t i = 0; i < n; ++i) {
      const mcfl_layer& l = s->layers[i];
      mat_of(l.mutr, emission ? l.muaf : l.mua, l.g, l.n, out[i]);
    }
    const mcfl_inclusion& c = s->inclusion;
    mat_of(c.mutr, emission ? c.muaf : c.mua, c.g, c.n, out[n]);
    surface_quotients(out, n - 1);
    return;
  }
  const uint32_t n = s->n_bulks + 2;
  out.resize(n);
  for (uint32_t i = 0; i < n; ++i) {
    const mcfl_bulk& b = s->bulks[i];
    mat_of(b.mutr, emission ? b.muaf : b.mua, b.g, b.n, out[i]);
  }
  surface_quotients(out, n - 1);
}

// Layer bounds (z_min, z_max) of a layered medium, 0 above and n_layers+1 below.
void layer_bounds(const mcfl_sim* s, std::vector<float>& zb) {
  zb.assign(2 * ((size_t)s->n_layers + 2), 0.0f);
  for (uint32_t i = 0; i < s->n_layers + 2; ++i) {
    zb[2 * i] = s->layers[i].z_min;
    zb[2 * i + 1] = s->layers[i].z_max;
  }
}

// The layer LaunchPhoton starts a photon in (CUDAMCFLtransport.cu:547-560),
// 0 when the source depth is in none of the layers 1..n_layers.
uint32_t start_layer(const mcfl_sim* s) {
  for (uint32_t l = 1; l <= s->n_layers; ++l)
    if (s->zi < s->layers[l].z_max && s->zi >= s->layers[l].z_min) return l;
  return 0;
}

size_t nvox_of(const mcfl_sim* s) { return (size_t)s->nx * (size_t)s->ny * (size_t)s->nz; }
size_t npix_of(const mcfl_sim* s) { return (size_t)s->det.nx * (size_t)(s->det.ny + 1); }
size_t ntime_of(const mcfl_sim* s) {
  return (size_t)s->det.x_temp_numdets * (size_t)s->det.y_temp_numdets * (size_t)s->det.temp_bins;
}

mcfl_status validate(const mcfl_sim* s) {
  if (!s) return MCFL_ERR_ARG;
  if (s->bulk_method != 1 && s->bulk_method != 2) return MCFL_ERR_UNSUPPORTED;
  if (s->nx <= 0 || s->ny <= 0 || s->nz <= 0) return MCFL_ERR_ARG;
  if (s->grid_size <= 0 || !(s->esp > 0.0f)) return MCFL_ERR_ARG;
  if (s->det.nx <= 0 || s->det.ny <= 0) return MCFL_ERR_ARG;
  if (s->do_temp_sim && (s->det.temp_bins == 0 || s->det.max_temp == 0)) return MCFL_ERR_ARG;
  const size_t nv = nvox_of(s);
  if (nv > 0xFFFFFFFFull) return MCFL_ERR_CAPACITY;
  if (s->nx >= (1 << 24) || s->ny >= (1 << 24)) return MCFL_ERR_CAPACITY;  // 24-bit index products in the kernel
  if (s->bulk_method == 1) {
    if (!s->layers) return MCFL_ERR_ARG;
    if (s->n_layers < 1 || s->n_layers > 98) return MCFL_ERR_CAPACITY;
    if (start_layer(s) == 0) return MCFL_ERR_ARG;  // the reference leaves the layer undefined
    return MCFL_OK;
  }
  if (!s->bulk_info || !s->bulks) return MCFL_ERR_ARG;
  if (s->n_bulks < 1 || s->n_bulks > 98) return MCFL_ERR_CAPACITY;
  const uint8_t last = (uint8_t)(s->n_bulks + 1);
  for (size_t i = 0; i < nv; ++i)
    if (s->bulk_info[i] > last) return MCFL_ERR_LABEL;
  return MCFL_OK;
}

void fill_medium(const mcfl_sim* s, const uint8_t* d_labels, MediumDev* md, const float* d_zb) {
  memset(md, 0, sizeof *md);
  md->labels = d_labels;
  md->nx = s->nx;
  md->ny = s->ny;
  md->nz = s->nz;
  md->nxy = (uint32_t)s->nx * (uint32_t)s->ny;
  md->last = s->n_bulks + 1;
  if (s->bulk_method == 1) {
    md->last = s->n_layers + 1;
    md->n_layers = s->n_layers;
    md->zb = d_zb;
    md->inc_x = s->inclusion.x;
    md->inc_y = s->inclusion.y;
    md->inc_z = s->inclusion.z;
    md->inc_r2 = s->inclusion.r * s->inclusion.r;
  }
  md->esp = s->esp;
  md->gs = (float)s->grid_size;
  md->halfx = s->halfx;
  md->halfy = s->halfy;
}

void fill_det(const mcfl_sim* s, DetDev* d, const float* d_tpos) {
  d->dx = s->det.dx;
  d->dy = s->det.dy;
  d->x0 = s->det.x0;
  d->y0 = s->det.y0;
  d->size_x = (s->det.dx * (float)s->det.nx) * 0.5f;  // CUDAMCFLtransport.cu:259-260
  d->size_y = (s->det.dy * (float)s->det.ny) * 0.5f;
  d->nx = s->det.nx;
  d->ny = s->det.ny;
  d->npix = (uint32_t)npix_of(s);
  d->t_ndx = s->det.x_temp_numdets;
  d->t_ndy = s->det.y_temp_numdets;
  d->t_rort = s->det.temp_rort;
  d->t_r = s->det.temp_det_r;
  d->t_bins = s->det.temp_bins;
  d->t_max = s->det.max_temp;
  d->t_posx = d_tpos;
  d->t_posy = d_tpos ? d_tpos + s->det.x_temp_numdets : nullptr;
}

// Temporal detector positions as DoOneSimulation computes them (CUDAMCFLmain.cu:94-100).
void tdet_positions(const mcfl_sim* s, std::vector<float>& pos) {
  const int nx = (int)s->det.x_temp_numdets, ny = (int)s->det.y_temp_numdets;
  pos.resize((size_t)nx + (size_t)ny);
  for (int i = 0; i < nx; i++)
    pos[i] = i * s->det.x_temp_sepdets - ((nx - 1) * s->det.x_temp_sepdets) / 2 + s->det.x0_temp_det;
  for (int i = 0; i < ny; i++)
    pos[nx + i] = i * s->det.y_temp_sepdets - ((ny - 1) * s->det.y_temp_sepdets) / 2 + s->det.y0_temp_det;
}

int src_kind_of(const mcfl_sim* s) {
  if (s->dir == 0.0f) return MCFL_SRC_ISOTROPIC;
  return s->src_kind == MCFL_SOURCE_DISK ? MCFL_SRC_DISK : MCFL_SRC_PENCIL;
}

// Layer-bound entries of a launch's dynamic LDS (layered media only).
static uint32_t n_zb_of(const LaunchArgs& A) { return A.med.n_layers ? A.med.n_layers + 2u : 0u; }

int grid_blocks(const mcfl_ctx* c, const LaunchArgs& A, int temporal, int perjob) {
  const uint32_t nslots = A.work == MCFL_WORK_JOBS ? A.njobs : 1u;
  int bpc = 0;
  if (mcfl_transport_occupancy(A.work, temporal, perjob, A.med.n_layers ? 1 : 0, mcfl_wg_of(&A),
                               mcfl_lds_bytes(nslots, A.tile_n, A.n_mats, n_zb_of(A)), &bpc) != 0 ||
      bpc <= 0)
    bpc = 1;
  return c->n_cu * bpc;
}

// Register-limited resident workgroups per CU of A's kernel variant.
static int bpc_of(const mcfl_ctx* c, const LaunchArgs& A, int temporal, int perjob) {
  const int layered = A.med.n_layers ? 1 : 0;
  if (mcfl_wg_of(&A) == MCFL_WG) return c->bpc[A.work][temporal ? 1 : 0][perjob ? 1 : 0][layered];
  int bpc = 0;
  if (mcfl_transport_occupancy(A.work, temporal, perjob, layered, mcfl_wg_of(&A), 0, &bpc) != 0) return 0;
  return bpc;
}

void choose_tile(const mcfl_ctx* c, LaunchArgs& A, const MediumDev& m, int temporal, int perjob, int side_req) {
  A.tile_n = 0;
  const uint32_t nslots = A.work == MCFL_WORK_JOBS ? A.njobs : 1u;
  const int layered = A.med.n_layers ? 1 : 0;
  const int bpc = bpc_of(c, A, temporal, perjob);
  if (bpc <= 0) return;
  const size_t per_wg = c->lds_per_cu / (size_t)bpc;
  const size_t other = mcfl_lds_bytes(nslots, 0, A.n_mats, n_zb_of(A));
  if (per_wg <= other + 64 * 8) return;
  const uint32_t cap = (uint32_t)std::min<size_t>((per_wg - other) / 8, 1u << 15);
  // box s x s x 2s/3 around the source: of the boxes of this volume it keeps
  // the most voxel-change flushes out of global memory (the photon cloud is
  // deeper than it is wide; 24x24x16 keeps 2% more than 27x27x13 at 9.2k
  // entries, measured by replaying the transport on the CPU)
  int side = std::max(2, (int)std::lround(std::cbrt(1.5 * cap)));
  const int want = side_req > 0 ? side_req : c->tile_side;
  if (want > 0) side = std::max(2, std::min(want, (int)std::sqrt((double)cap)));
  const uint32_t tnx = (uint32_t)std::min(m.nx, side), tny = (uint32_t)std::min(m.ny, side);
  const uint32_t tnz = std::min<uint32_t>((uint32_t)m.nz, cap / (tnx * tny));
  if (tnz == 0) return;
  A.tnx = tnx;
  A.tny = tny;
  A.tnz = tnz;
  A.tile_n = tnx * tny * tnz;
  // keep the register-limited occupancy
  int got = 0;
  while (A.tile_n > 0 &&
         (mcfl_transport_occupancy(A.work, temporal, perjob, layered, mcfl_wg_of(&A),
                                   mcfl_lds_bytes(nslots, A.tile_n, A.n_mats, n_zb_of(A)), &got) != 0 ||
          got < bpc)) {
    if (A.tnz <= 1) {
      A.tile_n = 0;
      break;
    }
    A.tnz -= 1;
    A.tile_n = A.tnx * A.tny * A.tnz;
  }
}

void place_tile(const LaunchArgs& A, const MediumDev& m, float xi, float yi, JobDev& J) {
  J.tile = 0;
  J.tx0 = J.ty0 = 0;
  if (A.tile_n == 0 || A.tnx > (uint32_t)m.nx || A.tny > (uint32_t)m.ny || A.tnz > (uint32_t)m.nz) return;
  auto clampi = [](long v, long lo, long hi) { return v < lo ? lo : (v > hi ? hi : v); };
  const long ix = clampi((long)std::floor((xi + m.halfx) * m.gs), 0, m.nx - 1);
  const long iy = clampi((long)std::floor((yi + m.halfy) * m.gs), 0, m.ny - 1);
  J.tx0 = (uint32_t)clampi(ix - (long)A.tnx / 2, 0, m.nx - (long)A.tnx);
  J.ty0 = (uint32_t)clampi(iy - (long)A.tny / 2, 0, m.ny - (long)A.tny);
  J.tile = 1;
}

void plan_tiles(const mcfl_ctx* c, LaunchArgs& A, JobDev* jd, size_t n, int temporal, int perjob, int side_req) {
  A.tile_n = 0;
  const JobDev* first = nullptr;
  for (size_t j = 0; j < n; ++j) {
    jd[j].tile = 0;
    if (jd[j].fhd != nullptr && !first) first = &jd[j];
  }
  if (!first) return;
  choose_tile(c, A, first->med, temporal, perjob, side_req);
  for (size_t j = 0; j < n; ++j)
    if (jd[j].fhd != nullptr) place_tile(A, jd[j].med, jd[j].xi, jd[j].yi, jd[j]);
}

mcfl_status launch_and_time(mcfl_ctx* c, LaunchArgs& A, int temporal, int perjob, double* ms,
                            const std::function<int()>& post) {
  const int nblocks = grid_blocks(c, A, temporal, perjob);
  A.n_waves = (uint32_t)nblocks * (mcfl_wg_of(&A) / 64u);
  if (A.chunk_max == 0) A.chunk_max = c->chunk_max;
  A.refill_min = c->refill_min;
  A.max_steps = c->max_steps;
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  if (mcfl_transport_launch(&A, temporal, perjob, nblocks, c->stream) != 0) return MCFL_ERR_DEVICE;
  if (post && post() != 0) return MCFL_ERR_DEVICE;
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  HIPCHK(hipEventSynchronize(c->ev1));
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, c->ev0, c->ev1));
  if (ms) *ms = t;
  return MCFL_OK;
}

}  // namespace mcfl_detail

using namespace mcfl_detail;

namespace {

std::string lib_dir() {
  Dl_info info;
  if (dladdr(reinterpret_cast<void*>(&mcfl_default_primes_path), &info) && info.dli_fname) {
    std::string p(info.dli_fname);
    size_t s = p.rfind('/');
    return s == std::string::npos ? std::string(".") : p.substr(0, s);
  }
  return ".";
}

// Reads the first column of n lines after the seeding line.
bool read_multipliers(const char* path, uint32_t n, std::vector<uint32_t>& out) {
  FILE* f = fopen(path, "r");
  if (!f) return false;
  unsigned a0, t1, t2;
  if (fscanf(f, "%u %u %u", &a0, &t1, &t2) != 3) {
    fclose(f);
    return false;
  }
  out.resize(n);
  for (uint32_t i = 0; i < n; ++i) {
    unsigned a;
    if (fscanf(f, "%u %u %u", &a, &t1, &t2) != 3) {
      fclose(f);
      return false;
    }
    out[i] = a;
  }
  fclose(f);
  return true;
}

mcfl_status medium_upload(mcfl_ctx* c, const mcfl_sim* s, MediumDev* md) {
  if (s->bulk_method == 1) {  // layered: no labels, layer bounds instead
    std::vector<float> zb;
    layer_bounds(s, zb);
    STCHK(ensure(c->zb, zb.size() * sizeof(float)));
    HIPCHK(hipMemcpy(c->zb.p, zb.data(), zb.size() * sizeof(float), hipMemcpyHostToDevice));
    fill_medium(s, nullptr, md, c->zb.as<const float>());
    return MCFL_OK;
  }
  const size_t nv = nvox_of(s);
  STCHK(ensure(c->labels, nv));
  HIPCHK(hipMemcpyAsync(c->labels.p, s->bulk_info, nv, hipMemcpyHostToDevice, c->stream));
  fill_medium(s, c->labels.as<const uint8_t>(), md, nullptr);
  return MCFL_OK;
}

}  // namespace

// ------------------------------------------------------------------ C ABI

extern "C" const char* mcfl_default_primes_path(void) {
  static std::string path;
  if (path.empty()) path = lib_dir() + "/../data/safeprimes_base32.txt";
  return path.c_str();
}

extern "C" uint64_t mcfl_stream_key(uint64_t seed, uint64_t stream) { return ::mcfl_key(seed, stream); }

namespace {

// One line of the multipliers file: the multiplier in its first column.
bool read_multiplier(FILE* fp, uint32_t* a) {
  unsigned int m, p1, p2;
  if (fscanf(fp, "%u %u %u", &m, &p1, &p2) != 3) return false;
  *a = m;
  return true;
}

// Lane state drawn from the seeding generator g (multiplier `seeder`) for a
// lane of multiplier a: one step gives the carry as floor(u / 2^32 * a),
// evaluated in double precision as CUDAMCFLrng.cu:78 does, the next step's
// low word is the state's low word (:79-81).  Draws repeat until the state
// is valid for a (CUDAMCFLrng.cu:72).
uint64_t seed_lane(uint64_t* g, uint32_t seeder, uint32_t a) {
  uint64_t x;
  do {
    const uint32_t u = mcfl_mwc_next(g, seeder);
    const uint32_t carry = (uint32_t)floor((double)u / 4294967296.0 * (double)a);
    const uint32_t lo = mcfl_mwc_next(g, seeder);
    x = ((uint64_t)carry << 32) | lo;
  } while (!mcfl_mwc_state_ok(x, a));
  return x;
}

}  // namespace

// init_RNG, CUDAMCFLrng.cu:31-90: the file's first multiplier drives a
// seeding generator started at xinit (which must itself be a valid state for
// it, :56-65); lane i takes the multiplier on line i+2 and a state drawn from
// the seeding generator.
extern "C" int mcfl_init_rng(uint64_t* x, uint32_t* a, uint32_t n_rng, const char* file, uint64_t xinit) {
  if (!file || !file[0]) file = mcfl_default_primes_path();
  FILE* fp = fopen(file, "r");
  if (!fp) {
    fprintf(stderr, "mcfl: cannot open the multipliers file %s\n", file);
    return 1;
  }
  uint32_t seeder = 0;
  int rc = 0;
  if (!read_multiplier(fp, &seeder)) {
    rc = 1;
  } else if (!mcfl_mwc_state_ok(xinit, seeder)) {
    fprintf(stderr, "mcfl: seed %llu is not a valid state for the seeding multiplier\n", (unsigned long long)xinit);
    rc = 1;
  } else {
    uint64_t g = xinit;
    // The reference keeps the previous multiplier when the file runs out; a
    // short file is an error here.
    for (uint32_t i = 0; i < n_rng && rc == 0; i++) {
      if (!read_multiplier(fp, &a[i])) {
        rc = 1;
        break;
      }
      x[i] = seed_lane(&g, seeder, a[i]);
    }
  }
  fclose(fp);
  return rc;
}

extern "C" const char* mcfl_status_string(mcfl_status s) {
  switch (s) {
    case MCFL_OK: return "ok";
    case MCFL_ERR_ARG: return "invalid argument";
    case MCFL_ERR_IO: return "i/o error";
    case MCFL_ERR_PARSE: return "malformed input";
    case MCFL_ERR_SEED: return "invalid seed";
    case MCFL_ERR_PHOTONS: return "too few photons for the lane streams";
    case MCFL_ERR_LABEL: return "voxel label outside 0..n_bulks+1";
    case MCFL_ERR_NOMEM: return "out of memory";
    case MCFL_ERR_DEVICE: return "HIP runtime error";
    case MCFL_ERR_UNSUPPORTED: return "unsupported";
    case MCFL_ERR_CAPACITY: return "capacity exceeded";
  }
  return "unknown";
}

extern "C" mcfl_status mcfl_ctx_create(int device, const char* primes, mcfl_ctx** out) {
  if (!out) return MCFL_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) return MCFL_ERR_DEVICE;
  mcfl_ctx* c = new (std::nothrow) mcfl_ctx();
  if (!c) return MCFL_ERR_NOMEM;
  c->device = device;
  auto fail = [&](mcfl_status st) {
    mcfl_ctx_destroy(c);
    return st;
  };
  if (hipSetDevice(device) != hipSuccess) return fail(MCFL_ERR_DEVICE);
  // this device's allocation-free scratch (mcfl_run_excitation_noalloc)
  if (mcfl_scratch_address(device, &c->scratch, nullptr) != 0) return fail(MCFL_ERR_DEVICE);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return fail(MCFL_ERR_DEVICE);
  c->n_cu = prop.multiProcessorCount;
  if (const char* e = getenv("MCFL_TALLY_COPIES")) c->copies = atoi(e) > 1 ? MCFL_TALLY_COPIES : 1;
  // the default split keeps one launch's share of a voxel below the u64 range
  // of one copy for F < 256 full weights per photon: with a single copy the
  // launches are 8 times shorter
  if (c->copies == 1) c->run_photons /= MCFL_TALLY_COPIES;
  if (const char* e = getenv("MCFL_TILE_SIDE")) c->tile_side = atoi(e);
  if (const char* e = getenv("MCFL_REFILL_MIN")) c->refill_min = (uint32_t)std::min(64, std::max(1, atoi(e)));
  if (const char* e = getenv("MCFL_CHUNK_MAX")) c->chunk_max = (uint32_t)std::min(1 << 20, std::max(64, atoi(e)));
  if (const char* e = getenv("MCFL_RUN_PHOTONS")) c->run_photons = std::max(1ull, strtoull(e, nullptr, 10));
  if (const char* e = getenv("MCFL_MAX_STEPS")) c->max_steps = (uint32_t)std::min(1ull << 31, strtoull(e, nullptr, 10));
  snprintf(c->name, sizeof c->name, "%s", prop.gcnArchName);
  for (int w = 0; w < 4; ++w)
    for (int t = 0; t < 2; ++t)
      for (int pj = 0; pj < 2; ++pj)
        for (int ly = 0; ly < 2; ++ly) {
          if (ly && (pj || w == MCFL_WORK_VOXELS)) continue;  // no such variant
          if (w == MCFL_WORK_VLANES && (t || pj || ly)) continue;
          if (mcfl_transport_occupancy(w, t, pj, ly, 0, 0, &c->bpc[w][t][pj][ly]) != 0) return fail(MCFL_ERR_DEVICE);
        }
  if (prop.maxSharedMemoryPerMultiProcessor > 0) c->lds_per_cu = prop.maxSharedMemoryPerMultiProcessor;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return fail(MCFL_ERR_DEVICE);
  if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) return fail(MCFL_ERR_DEVICE);
  if (!primes || !primes[0]) primes = mcfl_default_primes_path();
  if (!read_multipliers(primes, kTablePhotons, c->atab)) return fail(MCFL_ERR_IO);
  if (hipMalloc(&c->d_atab, kTablePhotons * sizeof(uint32_t)) != hipSuccess) return fail(MCFL_ERR_NOMEM);
  if (hipMemcpy(c->d_atab, c->atab.data(), kTablePhotons * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess)
    return fail(MCFL_ERR_DEVICE);
  *out = c;
  return MCFL_OK;
}

extern "C" void mcfl_ctx_destroy(mcfl_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (DevBuf* b : {&c->labels, &c->zb, &c->mats, &c->jobs, &c->counter, &c->tot, &c->Rd, &c->Tt, &c->fhd, &c->axyz,
                    &c->time, &c->Rd2, &c->Tt2, &c->tpos, &c->xs, &c->as, &c->vox, &c->vscale, &c->fhd_acc,
                    &c->axyz_acc})
    release(*b);
  if (c->d_atab) (void)hipFree(c->d_atab);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

extern "C" mcfl_status mcfl_ctx_info(mcfl_ctx* c, int* n_cu, int* bpc, char* name, int name_len) {
  if (!c) return MCFL_ERR_ARG;
  if (n_cu) *n_cu = c->n_cu;
  if (bpc) *bpc = c->bpc[MCFL_WORK_JOBS][0][0][0];
  if (name && name_len > 0) snprintf(name, (size_t)name_len, "%s", c->name);
  return MCFL_OK;
}

namespace {

// Device buffers of one excitation run: the context's own, grown on demand,
// or (noalloc) slices of this GPU's scratch array at the MCFL_NA_OFF_*
// offsets, after a check against the slice's size.
enum Slot { B_LABELS, B_ZB, B_MATS, B_JOBS, B_COUNTER, B_TOT, B_RD, B_TT, B_FHD, B_AXYZ, B_TIME, B_RD2, B_TT2,
            B_TPOS, B_XS, B_AS, B_NSLOTS };

struct RunBufs {
  mcfl_ctx* c;
  bool noalloc;
  void* p[B_NSLOTS] = {};
  uint64_t vol_used = 0;  // noalloc: bytes of the volume-copy region taken by fhd
  RunBufs(mcfl_ctx* ctx, bool na) : c(ctx), noalloc(na) {}
  mcfl_status get(Slot k, size_t bytes) {
    if (!noalloc) {
      DevBuf* own[B_NSLOTS] = {&c->labels, &c->zb, &c->mats, &c->jobs, &c->counter, &c->tot, &c->Rd, &c->Tt,
                               &c->fhd, &c->axyz, &c->time, &c->Rd2, &c->Tt2, &c->tpos, &c->xs, &c->as};
      STCHK(ensure(*own[k], bytes));
      p[k] = own[k]->p;
      return MCFL_OK;
    }
    static const uint64_t off[B_NSLOTS] = {MCFL_NA_OFF_LABELS, MCFL_NA_OFF_ZB, MCFL_NA_OFF_MATS, MCFL_NA_OFF_JOBS,
                                           MCFL_NA_OFF_COUNTER, MCFL_NA_OFF_TOT, MCFL_NA_OFF_RD, MCFL_NA_OFF_TT,
                                           MCFL_NA_OFF_VOL, MCFL_NA_OFF_VOL, MCFL_NA_OFF_TIME, MCFL_NA_OFF_RD2,
                                           MCFL_NA_OFF_TT2, MCFL_NA_OFF_TPOS, MCFL_NA_OFF_XS, MCFL_NA_OFF_AS};
    static const uint64_t cap[B_NSLOTS] = {MCFL_NA_MAX_VOXELS, 1024, 16384, 1536, 8, 8 * MCFL_NTOT,
                                           8ull * MCFL_NA_MAX_PIX, 8ull * MCFL_NA_MAX_PIX, MCFL_NA_VOL_BYTES,
                                           MCFL_NA_VOL_BYTES, 8ull * MCFL_NA_MAX_TIME, 8ull * MCFL_NA_MAX_PIX,
                                           8ull * MCFL_NA_MAX_PIX, 4ull * MCFL_NA_MAX_TDET, 8ull * MCFL_NA_MAX_LANES,
                                           4ull * MCFL_NA_MAX_LANES};
    uint64_t o = off[k], limit = cap[k];
    if (k == B_AXYZ) {  // the A_xyz copies follow the fhd copies in the volume region
      o += vol_used;
      limit -= vol_used;
    }
    if (bytes > limit) return MCFL_ERR_CAPACITY;
    if (k == B_FHD) vol_used = (bytes + 255) & ~255ull;
    p[k] = static_cast<char*>(c->scratch) + o;
    return MCFL_OK;
  }
  template <class T>
  T* at(Slot k) const {
    return static_cast<T*>(p[k]);
  }
};

// mcfl_run_excitation and its allocation-free variant: buffers from b.
mcfl_status run_excitation(mcfl_ctx* c, const mcfl_sim* s, const mcfl_rng* rng, mcfl_tallies* out, RunBufs& b) {
  if (!c || !rng || !out || !out->Rd_xy || !out->Tt_xy) return MCFL_ERR_ARG;
  STCHK(validate(s));
  if (s->number_of_photons == 0) return MCFL_ERR_ARG;
  const bool lanes = rng->mode == MCFL_RNG_LANES;
  if (lanes) {
    if (!rng->x || !rng->a || rng->n_lanes == 0) return MCFL_ERR_ARG;
    if (s->number_of_photons < rng->n_lanes) return MCFL_ERR_PHOTONS;  // the reference never terminates
    if (rng->first_photon != 0) return MCFL_ERR_ARG;
  } else if (rng->mode != MCFL_RNG_PHOTON) {
    return MCFL_ERR_ARG;
  }
  if (hipSetDevice(c->device) != hipSuccess) return MCFL_ERR_DEVICE;
  const int temporal = s->do_temp_sim == 1 && out->time_xyt ? 1 : 0;
  const bool fhd_on = s->fhd_activated == 1 && out->fhd;
  const size_t nv = nvox_of(s), npix = npix_of(s);
  const size_t ntime = (size_t)s->det.x_temp_numdets * s->det.y_temp_numdets * s->det.temp_bins;
  const uint64_t N = s->number_of_photons;
  const bool split = !lanes && N > c->run_photons;
  if (b.noalloc && split) return MCFL_ERR_CAPACITY;  // a split run needs accumulators beyond the scratch

  LaunchArgs A;
  memset(&A, 0, sizeof A);
  if (s->bulk_method == 1) {  // layered: no labels, layer bounds instead
    std::vector<float> zb;
    layer_bounds(s, zb);
    STCHK(b.get(B_ZB, zb.size() * sizeof(float)));
    HIPCHK(hipMemcpyAsync(b.p[B_ZB], zb.data(), zb.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    fill_medium(s, nullptr, &A.med, b.at<const float>(B_ZB));
  } else {
    STCHK(b.get(B_LABELS, nv));
    HIPCHK(hipMemcpyAsync(b.p[B_LABELS], s->bulk_info, nv, hipMemcpyHostToDevice, c->stream));
    fill_medium(s, b.at<const uint8_t>(B_LABELS), &A.med, nullptr);
  }
  std::vector<MatDev> mats;
  build_mats(s, false, mats);
  STCHK(b.get(B_MATS, mats.size() * sizeof(MatDev)));
  HIPCHK(hipMemcpyAsync(b.p[B_MATS], mats.data(), mats.size() * sizeof(MatDev), hipMemcpyHostToDevice, c->stream));

  // tallies cleared with the fill kernel
  auto clear = [&](Slot k, size_t words) -> mcfl_status {
    return mcfl_fill_u64(b.at<unsigned long long>(k), words, c->stream) == 0 ? MCFL_OK : MCFL_ERR_DEVICE;
  };
  STCHK(b.get(B_RD, npix * 8));
  STCHK(b.get(B_TT, npix * 8));
  STCHK(b.get(B_TOT, MCFL_NTOT * 8));
  STCHK(b.get(B_COUNTER, 8));
  STCHK(clear(B_RD, npix));
  STCHK(clear(B_TT, npix));
  STCHK(clear(B_TOT, MCFL_NTOT));
  STCHK(clear(B_COUNTER, 1));

  JobDev J;
  memset(&J, 0, sizeof J);
  J.begin = 0;
  J.count = lanes ? rng->n_lanes : N;
  J.k0 = lanes ? 0 : rng->first_photon;
  J.key = ::mcfl_key(rng->seed, 0);
  J.src_kind = src_kind_of(s);
  J.xi = s->xi;
  J.yi = s->yi;
  J.zi = s->zi;
  J.src_r = s->src_r;
  J.start_weight = s->start_weight;
  J.mat_base = 0;
  J.start_lab = s->bulk_method == 1 ? start_layer(s) : 0u;
  J.Rd = b.at<unsigned long long>(B_RD);
  J.Tt = b.at<unsigned long long>(B_TT);
  J.tot = b.at<unsigned long long>(B_TOT);
  const uint32_t copies = c->copies;
  J.copy_stride = copies > 1 ? nv : 0;
  if (fhd_on) {
    STCHK(b.get(B_FHD, nv * 8 * copies));
    STCHK(clear(B_FHD, nv * copies));
    J.fhd = b.at<unsigned long long>(B_FHD);
    J.flags |= MCFL_F_FHD;
  }
  if (out->A_xyz) {
    STCHK(b.get(B_AXYZ, nv * 8 * copies));
    STCHK(clear(B_AXYZ, nv * copies));
    J.axyz = b.at<unsigned long long>(B_AXYZ);
    J.flags |= MCFL_F_AXYZ;
  }
  if (out->Rd2_xy && out->Tt2_xy) {
    STCHK(b.get(B_RD2, npix * 8));
    STCHK(b.get(B_TT2, npix * 8));
    STCHK(clear(B_RD2, npix));
    STCHK(clear(B_TT2, npix));
    J.Rd2 = b.at<unsigned long long>(B_RD2);
    J.Tt2 = b.at<unsigned long long>(B_TT2);
    J.flags |= MCFL_F_SQ;
  }
  const float* d_tpos = nullptr;
  if (temporal) {
    std::vector<float> pos;
    tdet_positions(s, pos);
    STCHK(b.get(B_TPOS, pos.size() * sizeof(float)));
    HIPCHK(hipMemcpyAsync(b.p[B_TPOS], pos.data(), pos.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    STCHK(b.get(B_TIME, ntime * 8));
    STCHK(clear(B_TIME, ntime));
    J.time = b.at<unsigned long long>(B_TIME);
    J.flags |= MCFL_F_TEMP;
    d_tpos = b.at<const float>(B_TPOS);
  }
  fill_det(s, &J.det, d_tpos);
  static_assert(sizeof(JobDev) <= 1536, "the scratch holds one job of at most 1536 bytes");
  STCHK(b.get(B_JOBS, sizeof(JobDev)));  // uploaded below, once its tile is placed

  A.mats = b.at<const MatDev>(B_MATS);
  A.n_mats = (uint32_t)mats.size();
  A.counter = b.at<unsigned long long>(B_COUNTER);
  A.a_tab = c->d_atab;
  A.a_mask = kTablePhotons - 1;
  A.jobs = b.at<const JobDev>(B_JOBS);
  A.njobs = 1;
  if (lanes) {
    const size_t L = rng->n_lanes;
    STCHK(b.get(B_XS, L * 8));
    STCHK(b.get(B_AS, L * 4));
    HIPCHK(hipMemcpyAsync(b.p[B_XS], rng->x, L * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(b.p[B_AS], rng->a, L * 4, hipMemcpyHostToDevice, c->stream));
    A.work = MCFL_WORK_LANES;
    A.total = L;
    A.xs = b.at<unsigned long long>(B_XS);
    A.as = b.at<const uint32_t>(B_AS);
    A.lane_q = N / L;
    A.lane_rem = N % L;
  } else {
    A.work = MCFL_WORK_JOBS;
    A.total = N;
  }
  A.vol = mcfl_vol_of(J.flags);
  if (fhd_on) {
    choose_tile(c, A, A.med, temporal, 0);
    place_tile(A, A.med, s->xi, s->yi, J);
  }
  unsigned long long* fhd_out = b.at<unsigned long long>(B_FHD);
  unsigned long long* axyz_out = b.at<unsigned long long>(B_AXYZ);
  if (!split) {
    HIPCHK(hipMemcpyAsync(b.p[B_JOBS], &J, sizeof J, hipMemcpyHostToDevice, c->stream));
    auto reduce = [&]() -> int {
      int e = 0;
      if (fhd_on) e |= mcfl_reduce_copies(fhd_out, nv, copies, c->stream);
      if (out->A_xyz) e |= mcfl_reduce_copies(axyz_out, nv, copies, c->stream);
      return e;
    };
    STCHK(launch_and_time(c, A, temporal, 0, &out->kernel_ms, reduce));
    HIPCHK(hipMemcpy(out->totals, b.p[B_TOT], MCFL_NTOT * 8, hipMemcpyDeviceToHost));
  } else {
    /* Photons [k0, k0 + n) per launch: photon k draws from stream k whatever
       launch runs it, so the tallies equal one launch's, except that a
       voxel's fhd / A_xyz and the weight totals saturate at 2^64-1 instead of
       wrapping (each launch's copies stay far below the u64 range). */
    if (fhd_on) {
      STCHK(ensure(c->fhd_acc, nv * 8));
      HIPCHK(hipMemsetAsync(c->fhd_acc.p, 0, nv * 8, c->stream));
      fhd_out = c->fhd_acc.as<unsigned long long>();
    }
    if (out->A_xyz) {
      STCHK(ensure(c->axyz_acc, nv * 8));
      HIPCHK(hipMemsetAsync(c->axyz_acc.p, 0, nv * 8, c->stream));
      axyz_out = c->axyz_acc.as<unsigned long long>();
    }
    memset(out->totals, 0, sizeof out->totals);
    out->kernel_ms = 0.0;
    for (uint64_t k0 = 0; k0 < N; k0 += c->run_photons) {
      const uint64_t n = std::min(c->run_photons, N - k0);
      J.k0 = rng->first_photon + k0;
      J.count = n;
      A.total = n;
      HIPCHK(hipMemcpyAsync(b.p[B_JOBS], &J, sizeof J, hipMemcpyHostToDevice, c->stream));
      STCHK(clear(B_COUNTER, 1));
      STCHK(clear(B_TOT, MCFL_NTOT));
      auto accumulate = [&]() -> int {
        int e = 0;
        if (fhd_on) e |= mcfl_accumulate_copies(b.at<unsigned long long>(B_FHD), nv, copies, fhd_out, c->stream);
        if (out->A_xyz)
          e |= mcfl_accumulate_copies(b.at<unsigned long long>(B_AXYZ), nv, copies, axyz_out, c->stream);
        return e;
      };
      double ms = 0.0;
      STCHK(launch_and_time(c, A, temporal, 0, &ms, accumulate));
      out->kernel_ms += ms;
      uint64_t t[MCFL_NTOT];
      HIPCHK(hipMemcpy(t, b.p[B_TOT], sizeof t, hipMemcpyDeviceToHost));
      for (int i = 0; i < MCFL_NTOT; ++i) {
        const uint64_t a = out->totals[i] + t[i];
        out->totals[i] = a < t[i] ? ~0ull : a;
      }
    }
  }
  // The copies below run on the null stream, which does not wait for the
  // context's non-blocking stream: finish its work first.
  HIPCHK(hipStreamSynchronize(c->stream));

  HIPCHK(hipMemcpy(out->Rd_xy, b.p[B_RD], npix * 8, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(out->Tt_xy, b.p[B_TT], npix * 8, hipMemcpyDeviceToHost));
  if (fhd_on) HIPCHK(hipMemcpy(out->fhd, fhd_out, nv * 8, hipMemcpyDeviceToHost));
  if (out->A_xyz) HIPCHK(hipMemcpy(out->A_xyz, axyz_out, nv * 8, hipMemcpyDeviceToHost));
  if (J.flags & MCFL_F_SQ) {
    HIPCHK(hipMemcpy(out->Rd2_xy, b.p[B_RD2], npix * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(out->Tt2_xy, b.p[B_TT2], npix * 8, hipMemcpyDeviceToHost));
  }
  if (temporal) HIPCHK(hipMemcpy(out->time_xyt, b.p[B_TIME], ntime * 8, hipMemcpyDeviceToHost));
  if (lanes) HIPCHK(hipMemcpy(rng->x, b.p[B_XS], (size_t)rng->n_lanes * 8, hipMemcpyDeviceToHost));
  out->num_terminated_photons = out->totals[MCFL_TOT_PHOTONS];
  return MCFL_OK;
}

}  // namespace

extern "C" mcfl_status mcfl_run_excitation(mcfl_ctx* c, const mcfl_sim* s, const mcfl_rng* rng, mcfl_tallies* out) {
  if (!c) return MCFL_ERR_ARG;
  RunBufs b(c, false);
  return run_excitation(c, s, rng, out, b);
}

extern "C" mcfl_status mcfl_run_excitation_noalloc(mcfl_ctx* c, const mcfl_sim* s, const mcfl_rng* rng,
                                                   mcfl_tallies* out) {
  if (!c || !c->scratch) return MCFL_ERR_ARG;
  RunBufs b(c, true);
  return run_excitation(c, s, rng, out, b);
}

extern "C" uint64_t mcfl_DoOneSimulation(mcfl_ctx* c, mcfl_sim* s, uint64_t* x, uint32_t* a, double* Fx) {
  const size_t nv = s ? nvox_of(s) : 0, npix = s ? npix_of(s) : 0;
  if (!s) return 0;
  std::vector<uint64_t> Rd(npix), Tt(npix), fhd(s->fhd_activated == 1 ? nv : 0);
  std::vector<uint64_t> tim(s->do_temp_sim == 1 ? (size_t)s->det.x_temp_numdets * s->det.y_temp_numdets * s->det.temp_bins : 0);
  mcfl_tallies t;
  memset(&t, 0, sizeof t);
  t.Rd_xy = Rd.data();
  t.Tt_xy = Tt.data();
  t.fhd = fhd.empty() ? nullptr : fhd.data();
  t.time_xyt = tim.empty() ? nullptr : tim.data();
  mcfl_rng rng;
  memset(&rng, 0, sizeof rng);
  rng.mode = MCFL_RNG_LANES;
  rng.x = x;
  rng.a = a;
  rng.n_lanes = MCFL_NUM_THREADS;
  clock_t t0 = clock();
  mcfl_status st = mcfl_run_excitation(c, s, &rng, &t);
  if (st != MCFL_OK) {
    fprintf(stderr, "excitation failed: %s\n", mcfl_status_string(st));
    return 0;
  }
  double secs = (double)(clock() - t0) / CLOCKS_PER_SEC;
  printf("\nSimulation time: %.2f sec\n\n", secs);
  printf("Writing excitation results...\n");
  mcfl_write_simulation_results(s, &t, secs);
  const uint64_t fin = t.num_terminated_photons;
  if (Fx && !fhd.empty())
    for (size_t v = 0; v < nv; ++v) Fx[v] = ((double)fhd[v] / (0xFFFFFFFFull * fin));  // CUDAMCFLmain.cu:181-183
  printf("Photons simulated: %llu\n\n", (unsigned long long)fin);
  return fin;
}

extern "C" uint64_t mcfl_DoOneSimulationFl(mcfl_ctx* c, mcfl_sim* s, uint64_t* x, uint32_t* a, uint64_t* R_out,
                                           uint64_t* T_out) {
  if (!c || !s || !x || !a || !R_out || !T_out) return 0;
  const size_t npix = npix_of(s);
  std::vector<uint64_t> tim(s->do_temp_sim == 1 ? ntime_of(s) : 0);
  mcfl_tallies t;
  memset(&t, 0, sizeof t);
  t.Rd_xy = R_out;  // det.nx + det.ny*det.nx entries, the reference's tempvoxelR/T (CUDAMCFLmain.cu:203-205,275-281)
  t.Tt_xy = T_out;
  t.time_xyt = tim.empty() ? nullptr : tim.data();
  mcfl_rng rng;
  memset(&rng, 0, sizeof rng);
  rng.mode = MCFL_RNG_LANES;
  rng.x = x;
  rng.a = a;
  rng.n_lanes = MCFL_NUM_THREADS;
  // the caller's simulation as it stands: the reference's voxel loop has set
  // mua <- muaf, dir = 0, fhd off and the source at the voxel corner
  // (CUDAMCFLmain.cu:453-466,481-487); fhd is not tallied here
  mcfl_sim sv = *s;
  sv.fhd_activated = 0;
  const mcfl_status st = mcfl_run_excitation(c, &sv, &rng, &t);
  if (st != MCFL_OK) {
    fprintf(stderr, "fluorescence voxel simulation failed: %s\n", mcfl_status_string(st));
    for (size_t i = 0; i < npix; ++i) R_out[i] = T_out[i] = 0;
    return 0;  // the reference's failure value (CUDAMCFLmain.cu:268-272)
  }
  return t.num_terminated_photons;
}

extern "C" mcfl_status mcfl_fl_plan(const mcfl_sim* s, const double* Fx, uint64_t nper, uint32_t* vox, double* scale,
                                    uint64_t* n_vox, double* unit) {
  if (!s || !Fx || !n_vox || !unit || nper == 0) return MCFL_ERR_ARG;
  STCHK(validate(s));
  if (s->bulk_method == 1) {  // method 1 emits only where albedof < 0, i.e. nowhere (CUDAMCFLmain.cu:520,535)
    *n_vox = 0;
    *unit = 1.0;
    return MCFL_OK;
  }
  const size_t nv = nvox_of(s);
  const uint64_t cap = *n_vox;
  uint64_t n = 0;
  double S = 0.0;
  for (size_t v = 0; v < nv; ++v) {
    const mcfl_bulk& b = s->bulks[s->bulk_info[v]];
    if (!(b.albedof < 1)) continue;  // CUDAMCFLmain.cu:551
    // voxelw, CUDAMCFLmain.cu:553-556
    const double w = ((double)b.eY * (double)(1 - b.albedof) * Fx[v]) / (double)(nper * 0xFFFFFFFFull);
    if (n < cap) {
      if (vox) vox[n] = (uint32_t)v;
      if (scale) scale[n] = w;
    }
    S += w * (double)nper * 4294967296.0;
    ++n;
  }
  *n_vox = n;
  // One raw unit: the largest possible image sum maps below 2^62.
  *unit = S > 0.0 ? S / 4611686018427387904.0 : 1.0;
  if (scale)
    for (uint64_t j = 0; j < n && j < cap; ++j) scale[j] /= *unit;
  return n <= cap ? MCFL_OK : MCFL_ERR_CAPACITY;
}

// The emission stage on photon streams (rng == nullptr: one job per emitting
// voxel, MCFL_WORK_VOXELS) or on the reference's lanes (MCFL_WORK_VLANES:
// voxels in the reference loop's order, lane states advanced in place).
static mcfl_status fluorescence(mcfl_ctx* c, const mcfl_sim* s, const double* Fx, uint64_t nper, uint64_t seed,
                                const mcfl_rng* rng, double* FlR, double* FlT, mcfl_fl_result* res) {
  if (!c || !s || !Fx || nper == 0) return MCFL_ERR_ARG;
  STCHK(validate(s));
  const bool lanes = rng != nullptr;
  if (lanes) {
    if (!rng->x || !rng->a || rng->n_lanes == 0) return MCFL_ERR_ARG;
    if (rng->first_photon != 0) return MCFL_ERR_ARG;  // photon offsets exist for photon streams only
    if (nper < rng->n_lanes) return MCFL_ERR_PHOTONS;  // DoOneSimulationFl fails (CUDAMCFLmain.cu:268-272)
  }
  if (hipSetDevice(c->device) != hipSuccess) return MCFL_ERR_DEVICE;
  const size_t nv = nvox_of(s), npix = npix_of(s);
  if (s->bulk_method == 1) {  // no emitting voxel (see mcfl_fl_plan): zero images
    const size_t xy = (size_t)s->det.nx * (size_t)s->det.ny;
    if (FlR) memset(FlR, 0, xy * sizeof(double));
    if (FlT) memset(FlT, 0, xy * sizeof(double));
    if (res) {
      if (res->FlR_raw) memset(res->FlR_raw, 0, npix * 8);
      if (res->FlT_raw) memset(res->FlT_raw, 0, npix * 8);
      res->unit = 1.0;
      res->voxels = res->photons = 0;
      res->kernel_ms = 0.0;
      memset(res->totals, 0, sizeof res->totals);
    }
    return MCFL_OK;
  }
  std::vector<uint32_t> vox(nv);
  std::vector<double> scale(nv);
  uint64_t nvx = nv;
  double unit = 1.0;
  STCHK(mcfl_fl_plan(s, Fx, nper, vox.data(), scale.data(), &nvx, &unit));
  vox.resize(nvx);
  scale.resize(nvx);
  if (lanes) {  // the reference loop's order: x outermost, then y, then z (CUDAMCFLmain.cu:471-473)
    std::vector<size_t> ord(nvx);
    for (size_t j = 0; j < nvx; ++j) ord[j] = j;
    const uint32_t nx = (uint32_t)s->nx, nxy = (uint32_t)s->nx * (uint32_t)s->ny;
    std::sort(ord.begin(), ord.end(), [&](size_t p, size_t q) {
      const uint32_t u = vox[p], w = vox[q];
      const uint32_t ux = u % nx, wx = w % nx, uy = (u % nxy) / nx, wy = (w % nxy) / nx;
      if (ux != wx) return ux < wx;
      if (uy != wy) return uy < wy;
      return u / nxy < w / nxy;
    });
    std::vector<uint32_t> v2(nvx);
    std::vector<double> s2(nvx);
    for (size_t j = 0; j < nvx; ++j) {
      v2[j] = vox[ord[j]];
      s2[j] = scale[ord[j]];
    }
    vox.swap(v2);
    scale.swap(s2);
  }

  LaunchArgs A;
  memset(&A, 0, sizeof A);
  STCHK(medium_upload(c, s, &A.med));
  std::vector<MatDev> mats;
  build_mats(s, true, mats);
  STCHK(ensure(c->mats, mats.size() * sizeof(MatDev)));
  HIPCHK(hipMemcpyAsync(c->mats.p, mats.data(), mats.size() * sizeof(MatDev), hipMemcpyHostToDevice, c->stream));
  STCHK(ensure(c->Rd, npix * 8));
  STCHK(ensure(c->Tt, npix * 8));
  STCHK(ensure(c->tot, MCFL_MAX_JOBS * MCFL_NTOT * 8));
  STCHK(ensure(c->counter, 8));
  STCHK(ensure(c->vox, (nvx ? nvx : 1) * 4));
  STCHK(ensure(c->vscale, (nvx ? nvx : 1) * 8));
  HIPCHK(hipMemsetAsync(c->Rd.p, 0, npix * 8, c->stream));
  HIPCHK(hipMemsetAsync(c->Tt.p, 0, npix * 8, c->stream));
  HIPCHK(hipMemsetAsync(c->tot.p, 0, MCFL_NTOT * 8, c->stream));
  HIPCHK(hipMemsetAsync(c->counter.p, 0, 8, c->stream));
  if (nvx) {
    HIPCHK(hipMemcpyAsync(c->vox.p, vox.data(), nvx * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->vscale.p, scale.data(), nvx * 8, hipMemcpyHostToDevice, c->stream));
  }
  JobDev J;
  memset(&J, 0, sizeof J);
  J.count = nvx * nper;
  J.src_kind = MCFL_SRC_ISOTROPIC;
  J.start_weight = 0xFFFFFFFFu;
  J.flags = MCFL_F_SCALED;
  J.Rd = static_cast<unsigned long long*>(c->Rd.p);
  J.Tt = static_cast<unsigned long long*>(c->Tt.p);
  J.tot = static_cast<unsigned long long*>(c->tot.p);
  fill_det(s, &J.det, nullptr);
  STCHK(ensure(c->jobs, sizeof(JobDev)));
  HIPCHK(hipMemcpyAsync(c->jobs.p, &J, sizeof J, hipMemcpyHostToDevice, c->stream));

  A.mats = static_cast<const MatDev*>(c->mats.p);
  A.n_mats = (uint32_t)mats.size();
  A.work = lanes ? MCFL_WORK_VLANES : MCFL_WORK_VOXELS;
  A.total = lanes ? (nvx ? rng->n_lanes : 0) : nvx * nper;
  if (lanes) {
    const size_t L = rng->n_lanes;
    STCHK(ensure(c->xs, L * 8));
    STCHK(ensure(c->as, L * 4));
    HIPCHK(hipMemcpyAsync(c->xs.p, rng->x, L * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->as.p, rng->a, L * 4, hipMemcpyHostToDevice, c->stream));
    A.xs = static_cast<unsigned long long*>(c->xs.p);
    A.as = static_cast<const uint32_t*>(c->as.p);
    A.lane_q = nper / L;
    A.lane_rem = nper % L;
    A.n_vox = (uint32_t)nvx;
  }
  A.counter = static_cast<unsigned long long*>(c->counter.p);
  A.a_tab = c->d_atab;
  A.a_mask = kTablePhotons - 1;
  A.jobs = static_cast<const JobDev*>(c->jobs.p);
  A.njobs = 1;
  A.vox = static_cast<const uint32_t*>(c->vox.p);
  A.vscale = static_cast<const double*>(c->vscale.p);
  A.per_job = nper;
  A.vseed = seed;
  double ms = 0.0;
  if (A.total > 0) STCHK(launch_and_time(c, A, 0, 0, &ms));
  HIPCHK(hipStreamSynchronize(c->stream));  // the clears above when no launch ran

  std::vector<uint64_t> R(npix), T(npix);
  HIPCHK(hipMemcpy(R.data(), c->Rd.p, npix * 8, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(T.data(), c->Tt.p, npix * 8, hipMemcpyDeviceToHost));
  const double dxdy = (double)s->det.dx * (double)s->det.dy;
  const size_t xy = (size_t)s->det.nx * (size_t)s->det.ny;
  for (size_t i = 0; i < xy; ++i) {
    if (FlR) FlR[i] = (double)R[i] * unit / dxdy;
    if (FlT) FlT[i] = (double)T[i] * unit / dxdy;
  }
  if (res) {
    if (res->FlR_raw) memcpy(res->FlR_raw, R.data(), npix * 8);
    if (res->FlT_raw) memcpy(res->FlT_raw, T.data(), npix * 8);
    res->unit = unit;
    res->voxels = nvx;
    res->photons = nvx * nper;
    res->kernel_ms = ms;
    HIPCHK(hipMemcpy(res->totals, c->tot.p, MCFL_NTOT * 8, hipMemcpyDeviceToHost));
  }
  if (lanes && A.total > 0) HIPCHK(hipMemcpy(rng->x, c->xs.p, (size_t)rng->n_lanes * 8, hipMemcpyDeviceToHost));
  return MCFL_OK;
}

extern "C" mcfl_status mcfl_run_fluorescence(mcfl_ctx* c, const mcfl_sim* s, const double* Fx, uint64_t nper,
                                             uint64_t seed, double* FlR, double* FlT, mcfl_fl_result* res) {
  return fluorescence(c, s, Fx, nper, seed, nullptr, FlR, FlT, res);
}

extern "C" mcfl_status mcfl_run_fluorescence_lanes(mcfl_ctx* c, const mcfl_sim* s, const double* Fx, uint64_t nper,
                                                   const mcfl_rng* rng, double* FlR, double* FlT, mcfl_fl_result* res) {
  if (!rng || rng->mode != MCFL_RNG_LANES) return MCFL_ERR_ARG;
  return fluorescence(c, s, Fx, nper, 0, rng, FlR, FlT, res);
}
