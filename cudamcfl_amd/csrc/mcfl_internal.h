// mcfl_internal.h -- host-side internals shared by the single-run entry
// points (mcfl_host.cpp) and the batch / serving engine (mcfl_engine.cpp).
#ifndef MCFL_INTERNAL_H
#define MCFL_INTERNAL_H

#include <cstddef>
#include <cstdint>
#include <functional>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "../../include/mcfl.h"
#include "mcfl_device.h"

extern "C" int mcfl_transport_occupancy(int work, int temporal, int perjob, int layered, uint32_t wg, uint32_t lds_bytes,
                                        int* blocks_per_cu);
extern "C" int mcfl_transport_launch(const LaunchArgs* A, int temporal, int perjob, int nblocks, hipStream_t stream);
extern "C" int mcfl_fill_u64(unsigned long long* p, uint64_t n, hipStream_t stream);
// Device address of the current device's allocation-free scratch (MCFL_NA_BYTES).
extern "C" int mcfl_scratch_address(int device, void** out, uint64_t* bytes);
extern "C" int mcfl_reduce_copies(unsigned long long* p, uint64_t n, uint32_t ncopies, hipStream_t stream);
// Test support (no counterpart in the reference): mcfl_selfcheck counts, over
// inputs [lo, hi) (32-bit patterns), where a form of the per-step arithmetic
// the transport kernel uses differs on the GPU from an independent form
// (which: MCFL_SC_*; p the fixed operand's bits); mcfl_math_eval evaluates the
// shared FP32 functions (fn: MCFL_ME_*) on the GPU for comparison with the host.
extern "C" int mcfl_selfcheck(int which, uint32_t p, uint64_t lo, uint64_t hi, uint64_t* mismatches, uint64_t* first);
extern "C" int mcfl_math_eval(int fn, const uint32_t* in, uint32_t* out, uint64_t n);
// acc[i] += the sum of the ncopies copies of element i, both saturating at
// 2^64-1; the copies are zeroed (ready for the next launch of a chunked run).
extern "C" int mcfl_accumulate_copies(unsigned long long* p, uint64_t n, uint32_t ncopies, unsigned long long* acc,
                                      hipStream_t stream);

namespace mcfl_detail {

constexpr uint32_t kTablePhotons = 32768;  // photon-stream multipliers: lines 2..32769 of the primes file

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

mcfl_status ensure(DevBuf& b, size_t bytes);
void release(DevBuf& b);

void build_mats(const mcfl_sim* s, bool emission, std::vector<MatDev>& out);
size_t nvox_of(const mcfl_sim* s);
size_t npix_of(const mcfl_sim* s);
size_t ntime_of(const mcfl_sim* s);
mcfl_status validate(const mcfl_sim* s);
void fill_medium(const mcfl_sim* s, const uint8_t* d_labels, MediumDev* md, const float* d_zb = nullptr);
void layer_bounds(const mcfl_sim* s, std::vector<float>& zb);
uint32_t start_layer(const mcfl_sim* s);
void fill_det(const mcfl_sim* s, DetDev* d, const float* d_tpos);
void tdet_positions(const mcfl_sim* s, std::vector<float>& pos);
int src_kind_of(const mcfl_sim* s);

}  // namespace mcfl_detail

struct mcfl_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int n_cu = 0;
  int bpc[4][2][2][2] = {};  // resident workgroups per CU [work][temporal][perjob][layered], register-limited
  size_t lds_per_cu = 160 * 1024;
  uint32_t copies = MCFL_TALLY_COPIES;  // copies of the per-step volume tallies (1 or MCFL_TALLY_COPIES)
  int tile_side = 0;                    // lateral size of the fhd LDS tile, 0: from the LDS budget (env MCFL_TILE_SIDE)
  uint32_t refill_min = 4;              // idle lanes that trigger a wave's refill (env MCFL_REFILL_MIN)
  uint32_t chunk_max = 1024;            // largest photon chunk a wave takes from the counter (env MCFL_CHUNK_MAX)
  uint32_t max_steps = 0;               // per-photon step cap, 0: MCFL_MAX_STEPS_PER_PHOTON (env MCFL_MAX_STEPS)
  uint64_t run_photons = 1ull << 27;    // photon-stream excitations above this run as several launches
                                        // of at most this many photons (env MCFL_RUN_PHOTONS)
  char name[256] = {};
  uint32_t* d_atab = nullptr;
  void* scratch = nullptr;  // this device's allocation-free scratch (MCFL_NA_BYTES), looked up at creation
  std::vector<uint32_t> atab;
  mcfl_detail::DevBuf labels, zb, mats, jobs, counter, tot, Rd, Tt, fhd, axyz, time, Rd2, Tt2, tpos, xs, as, vox, vscale, fhd_acc, axyz_acc;
};

namespace mcfl_detail {
// Launch geometry, launch, HIP-event timing on the context stream.
// `post` (may be empty) queues follow-up kernels (copy reductions) inside the timed region.
mcfl_status launch_and_time(mcfl_ctx* c, LaunchArgs& A, int temporal, int perjob, double* ms,
                            const std::function<int()>& post = {});
int grid_blocks(const mcfl_ctx* c, const LaunchArgs& A, int temporal, int perjob);
// Chooses the dimensions of the fhd LDS tile (A.tnx, tny, tnz, tile_n) for a
// medium of m's size within the LDS left free at the kernel's
// register-limited occupancy.  side > 0 forces the tile's lateral size
// (clamped to the budget).
void choose_tile(const mcfl_ctx* c, LaunchArgs& A, const MediumDev& m, int temporal, int perjob, int side = 0);
// Places job J's tile (dimensions from A) around its source voxel in medium m;
// J.tile = 0 when the tile does not fit that medium.
void place_tile(const LaunchArgs& A, const MediumDev& m, float xi, float yi, JobDev& J);
// Tiles of a job launch: dimensions for the first job tallying fhd, then one
// tile per fhd job (the jobs' med fields hold their media).
void plan_tiles(const mcfl_ctx* c, LaunchArgs& A, JobDev* jd, size_t n, int temporal, int perjob, int side = 0);
}  // namespace mcfl_detail

#define MCFL_HIPCHK(expr)                           \
  do {                                              \
    if ((expr) != hipSuccess) return MCFL_ERR_DEVICE; \
  } while (0)
#define MCFL_STCHK(expr)            \
  do {                              \
    mcfl_status st_ = (expr);       \
    if (st_ != MCFL_OK) return st_; \
  } while (0)

#endif
