/*
 * mcfl_oracle.h -- CPU restatement of CUDAMCFL's voxelized ("bulk method 2")
 * photon transport, used ONLY as test infrastructure: the parity tests, the
 * smoke check and the CPU-baseline leg of the benchmark load it as the
 * checker.  No product code path links or calls it.
 *
 * It is written independently of the HIP kernel.  The only code the two share
 * is ../cudamcfl_amd/csrc/mcfl_fmath.h and mcfl_mwc.h (the FP32 math and the
 * MWC generator), which are pinned by their own tests against libm and the
 * generator's closed form.
 *
 * Reference functions restated (file:line in the reference repository):
 *   orc_init_rng        CUDAMCFLrng.cu:31-90       (init_RNG)
 *   orc_launch          CUDAMCFLtransport.cu:460-562 (LaunchPhoton, method 2)
 *   orc_spin            CUDAMCFLtransport.cu:585-625 (Spin)
 *   orc_reflect         CUDAMCFLtransport.cu:628-700 (Reflect)
 *   orc_move_to_boundary CUDAMCFLtransport.cu:702-847 (MoveToFirstBoundary)
 *   survive             CUDAMCFLtransport.cu:850-864 (PhotonSurvive)
 *   orc_step            CUDAMCFLtransport.cu:286-447 (one MCd3D iteration)
 *   orc_step_layered    CUDAMCFLtransport.cu:79-226  (one MCd iteration, method 1)
 *   orc_run_lanes       CUDAMCFLmain.cu:60-165 + CUDAMCFLtransport.cu:565-582
 *                       (DoOneSimulation relaunch loop, per-lane streams)
 *   orc_fluor_reference CUDAMCFLmain.cu:409-594 (per-voxel emission loop)
 *   orc_run_fluor_lanes CUDAMCFLmain.cu:196-288,471-594 (the same loop, fixed-point images)
 */
#ifndef MCFL_ORACLE_H
#define MCFL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* BulkStruct field order (CUDAMCFL.h:116-127). */
typedef struct {
  float mutr, mua, g, n, flc, muaf, eY, albedof;
} orc_bulk;

enum { ORC_SRC_ISOTROPIC = 0, ORC_SRC_PENCIL = 1, ORC_SRC_DISK = 2 };

typedef struct {
  /* medium: labels x-fastest, index ix + nx*(iy + ny*iz) */
  int32_t nx, ny, nz;
  float esp;   /* slab thickness [cm] */
  float gs;    /* voxels per cm, as float */
  float halfx; /* lateral half extents [cm]; 2*esp for reference-shaped grids */
  float halfy;
  uint32_t n_bulks;     /* labels 1..n_bulks are media, 0 above, n_bulks+1 below */
  const orc_bulk* bulks; /* n_bulks + 2 entries */
  const uint8_t* labels;
  /* reflectance / transmittance image */
  float det_dx, det_dy, det_x0, det_y0;
  int32_t det_nx, det_ny;
  /* temporal detectors */
  int32_t do_temp;
  uint32_t t_ndx, t_ndy;
  const float* t_posx;
  const float* t_posy;
  float t_r;
  uint32_t t_rort;
  uint64_t t_bins, t_max;
  /* source */
  int32_t src_kind;
  float xi, yi, zi, src_r;
  uint32_t start_weight;
  /* tallies */
  int32_t fhd_on, axyz_on, sq_on;
  /* bulk method 1 (layered slab with one spherical inclusion): bulks[] are
     the layers 0..n_bulks+1 (n_bulks = number of layers, 0 above and
     n_bulks+1 below), z_min/z_max their bounds, labels unused */
  int32_t layered;
  const float* z_min;
  const float* z_max;
  orc_bulk inc;
  float inc_x, inc_y, inc_z, inc_r;
  uint32_t start_layer; /* layer containing the source depth */
} orc_sim;

enum {
  ORC_T_LAUNCHED_W = 0, /* sum of start weights */
  ORC_T_R,              /* weight leaving through the top (all of it) */
  ORC_T_T,              /* weight leaving through the bottom */
  ORC_T_A,              /* absorbed weight */
  ORC_T_ROUL_GAIN,      /* roulette survivors' added weight */
  ORC_T_ROUL_LOSS,      /* roulette victims' weight */
  ORC_T_ZERO_KILL,      /* weight dropped when the absorbed quantum is 0 */
  ORC_T_PHOTONS,        /* photons launched */
  ORC_T_STEPS,          /* transport loop iterations */
  ORC_T_ABORTED,        /* photons stopped by the per-photon step cap */
  ORC_NTOT = 16
};

typedef struct {
  uint64_t* Rd;   /* det_nx*(det_ny+1) */
  uint64_t* Tt;   /* det_nx*(det_ny+1) */
  uint64_t* fhd;  /* nx*ny*nz (if fhd_on) */
  uint64_t* axyz; /* nx*ny*nz (if axyz_on) */
  uint64_t* time; /* t_ndx*t_ndy*t_bins (if do_temp) */
  uint64_t* Rd2;  /* per-pixel sum of (w*w)>>32 (if sq_on) */
  uint64_t* Tt2;
  uint64_t tot[ORC_NTOT];
} orc_tallies;

/* Per-photon step cap of the transport loop (bounded-work watchdog). */
#define ORC_MAX_STEPS_PER_PHOTON (1u << 28)
/* Sets the cap for later runs (0: ORC_MAX_STEPS_PER_PHOTON); not thread-safe
   against runs in progress. */
void orc_set_max_steps(uint32_t n);

/* init_RNG restatement: reads n_rng+1 lines of the primes file.  0 = OK. */
int orc_init_rng(uint64_t* x, uint32_t* a, uint32_t n_rng, const char* primes_file, uint64_t xinit);

/* Photon streams: photons [k0,k1) of a job with key `key`, multipliers
 * a_tab[k & (P-1)] (P a power of two).  Tallies are added (not cleared).
 * exit_scale > 0 replaces each image increment w by floor(w*exit_scale+0.5). */
int orc_run_photons(const orc_sim* sim, const uint32_t* a_tab, uint32_t P, uint64_t key, uint64_t k0,
                    uint64_t k1, double exit_scale, orc_tallies* t, int nthreads);

/* Reference per-lane streams: lane i runs quota_i photons back to back on
 * (x[i], a[i]); the first photon replays its launch draws (the reference's
 * LaunchPhoton_Global does not store the advanced state).  x is advanced in
 * place.  Quotas: N/L each, the first N%L lanes one more. */
int orc_run_lanes(const orc_sim* sim, uint64_t* x, const uint32_t* a, uint32_t L, uint64_t N, orc_tallies* t,
                  int nthreads);

/* Voxel emission jobs (photon streams): job j emits n_per_job photons
 * isotropically from the corner of voxel vox[j], keys keys[j], image
 * increments scaled by scale[j].  sim->src_kind is ignored. */
int orc_run_fluor_jobs(const orc_sim* sim_em, const uint32_t* a_tab, uint32_t P, uint32_t njobs,
                       const uint32_t* vox, const uint64_t* keys, const double* scale, uint64_t n_per_job,
                       orc_tallies* t, int nthreads);

/* The reference's emission loop on its lane streams, images in fixed point:
 * voxels vox[0..nvox) in the loop order, each one simulation of nv photons on
 * the continuing lanes (first photon of each lane replays its launch draws);
 * exits add floor(w*scale[j]+0.5).  x advanced in place.  -2 if nv < L. */
int orc_run_fluor_lanes(const orc_sim* sim, uint64_t* x, const uint32_t* a, uint32_t L, uint32_t nvox,
                        const uint32_t* vox, const double* scale, uint64_t nv, orc_tallies* t, int nthreads);

/* Reference emission loop (CUDAMCFLmain.cu:471-594): every voxel whose
 * material has albedof < 1 runs one isotropic simulation of n_per_voxel
 * photons on the continuing lane streams (x advanced in place); images are
 * accumulated in double exactly as the reference does.  Returns the number of
 * emission photons simulated; fails (returns voxel status 0) for
 * n_per_voxel <= L like the reference. */
uint64_t orc_fluor_reference(const orc_sim* sim_em, const double* Fx, uint64_t* x, const uint32_t* a, uint32_t L,
                             uint64_t n_per_voxel, double* FlR, double* FlT, int nthreads);

/* Unit pieces, exposed for the pinning tests. */
typedef struct {
  float x, y, z, dx, dy, dz;
  uint32_t weight, bulkpos;
  uint64_t tof;
} orc_photon;

int orc_reflect_test(float n1, float n2, float dz_in, float xi, float* dz_out, float* dx_scale);
void orc_spin_test(float g, const float* dir_in, uint64_t* x, uint32_t a, float* dir_out);
void orc_backstep_test(const float* p, const float* d, const float* c, float r2, float* t1, float* t2);
float orc_fresnel_test(float n1, float n2, float cos_i);

/* orc_launch on photon streams (key, a_tab of P entries): n photons' position
   and direction (6 floats each), weight and entry label. */
void orc_launch_sample(const orc_sim* s, int kind, const uint32_t* a_tab, uint32_t P, uint64_t key, uint64_t n,
                       float* pos_dir, uint32_t* weight, uint32_t* label);
/* survive (PhotonSurvive): 1 if the photon of weight w lives; its weight after. */
int orc_survive_test(uint32_t w, uint64_t* x, uint32_t a, uint32_t* w_out, uint64_t* gain, uint64_t* loss);
/* survive on n photon streams of weight w: photons that live, sum of their weights. */
void orc_survive_batch(uint32_t w, const uint32_t* a_tab, uint32_t P, uint64_t key, uint64_t n, uint64_t* lived,
                       uint64_t* weight_sum);
/* orc_move_to_boundary from position/direction pd: label found, position after, time of flight added. */
uint32_t orc_march_test(const orc_sim* s, const float* pd, uint32_t old_label, float max_s, float* out,
                        uint64_t* tof);
/* fn: 0 log, 1 sin(2 pi t), 2 cos(2 pi t), 3 uint32 (bits of in) -> float toward zero, 4 1/sqrt,
   5 x*rsqrt(x), 6 renormalisation factor */
void orc_math_eval(int fn, const float* in, float* out, size_t n);
/* n draws of the [0,1) generator from state x: states after each step and the draws */
void orc_mwc_draws(uint64_t x, uint32_t a, size_t n, uint64_t* states, float* co);
uint64_t orc_photon_state(uint64_t key, uint64_t k, uint32_t a);
/* float bit patterns in [lo, hi) where mcfl_tof_div(q) != q / 2.99792e-5, or
   where the kernel's increment mcfl_tof_inc(q) differs from the quotient
   truncated by mcfl_d2u64_rz */
uint64_t orc_tof_div_mismatches(uint32_t lo, uint32_t hi);

#ifdef __cplusplus
}
#endif
#endif
