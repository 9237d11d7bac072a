/*
 * mcfl_device.h -- device-side data layout of the transport launch, shared by
 * the HIP kernel (mcfl_transport.hip) and the host code that fills it.
 *
 * A launch runs one flattened photon index space [0, total) over a single
 * voxel medium.  The index space is cut either into explicit jobs
 * (MCFL_WORK_JOBS: one excitation simulation, a multi-source batch, the
 * current epoch of several served requests), into voxel emission jobs
 * (MCFL_WORK_VOXELS: the fluorescence stage, one job per emitting voxel), or
 * into the reference's per-lane streams (MCFL_WORK_LANES).
 */
#ifndef MCFL_DEVICE_H
#define MCFL_DEVICE_H

#include <stdint.h>

#define MCFL_MAX_MATERIALS 320   /* all material tables of one launch (LDS); one medium holds at most 100 (CUDAMCFL.h:24) */
#define MCFL_MAX_JOBS 64         /* explicit jobs per launch (per-job totals live in LDS) */
#define MCFL_NTOT 16             /* totals per job, see enum mcfl_total in mcfl.h */
#define MCFL_MAX_STEPS_PER_PHOTON (1u << 28)
/* Transport workgroup: MCFL_WG threads (whole wavefronts), register
   allocation for MCFL_WAVES_PER_EU resident waves per SIMD.  The LDS tile of
   the photon hitting density is per workgroup, so larger workgroups at the
   same occupancy share a larger tile. */
#ifndef MCFL_WG
#define MCFL_WG 768 /* 12 waves: 3 per SIMD, two workgroups (80 KB of LDS each) per CU */
#endif
#ifndef MCFL_WAVES_PER_EU
#define MCFL_WAVES_PER_EU 6 /* 80 VGPRs */
#endif
/* Workgroup sizes of the job kernels the batch autotuner chooses from: 512
   (3 workgroups per CU at 6 waves/SIMD, the smallest tiles), 768 (the
   default), 1024 (one workgroup per CU, 4 waves/SIMD, one 160 KB tile). */
#define MCFL_WG_ALT1 512
#define MCFL_WG_ALT2 1024
#ifndef MCFL_TALLY_COPIES
#define MCFL_TALLY_COPIES 8      /* per-step volume tallies: one copy per XCD, summed after the launch */
#endif

enum { MCFL_WORK_JOBS = 0, MCFL_WORK_VOXELS = 1, MCFL_WORK_LANES = 2, MCFL_WORK_VLANES = 3 };
/* Kernel-internal: a MCFL_WORK_JOBS launch of one job (the job's tally
   pointers, flags and material base are then wave-uniform scalars). */
enum { MCFL_WORK_ONE = 4 };
/* Volume tallies compiled into a job kernel: fhd and A_xyz, fhd alone, none. */
enum { MCFL_VOL_ALL = 0, MCFL_VOL_FHD = 1, MCFL_VOL_NONE = 2 };
enum { MCFL_SRC_ISOTROPIC = 0, MCFL_SRC_PENCIL = 1, MCFL_SRC_DISK = 2, MCFL_SRC_EMISSION = 3 };
enum {
  MCFL_DT_LAUNCHED_W = 0,
  MCFL_DT_R,
  MCFL_DT_T,
  MCFL_DT_A,
  MCFL_DT_ROUL_GAIN,
  MCFL_DT_ROUL_LOSS,
  MCFL_DT_ZERO_KILL,
  MCFL_DT_PHOTONS,
  MCFL_DT_STEPS,
  MCFL_DT_ABORTED
};

/* Per-material constants (48 bytes, LDS resident).  inv2g == 0 <=> g == 0.
   The surface quotients are the two divisions of Reflect against the media
   above (material 0) and below (the table's last outer material), rounded as
   the kernel would round them (n*n)/(n'*n') and n/n'. */
struct alignas(16) MatDev {
  float mutr;     /* 1/(mua_base + mus), FLT_MAX for glass */
  float mua_mutr; /* mua*mutr: absorbed fraction per interaction */
  float n;        /* refractive index */
  float onemg;    /* 1 - g */
  float onemg2;   /* 1 - g^2 */
  float onepg2;   /* 1 + g^2 */
  float twog32;   /* 2g * 2^-32: Spin's coefficient of the raw 32-bit draw */
  float inv2g;    /* 1/(2g), 0 when g == 0 */
  float q_top, r_top; /* (n*n)/(n0*n0), n/n0 */
  float q_bot, r_bot; /* the same against the medium below */
};

struct MediumDev {
  const uint8_t* labels; /* nx*ny*nz, x fastest */
  int32_t nx, ny, nz;
  uint32_t nxy;  /* nx*ny */
  uint32_t last; /* n_bulks + 1 (voxel medium), n_layers + 1 (layered) */
  float esp, gs, halfx, halfy;
  /* bulk method 1 (layered slab, n_layers > 0): materials 0..n_layers+1 are
     the layers, material n_layers+2 the spherical inclusion; zb holds z_min,
     z_max of each layer */
  const float* zb;
  uint32_t n_layers;
  float inc_x, inc_y, inc_z, inc_r2;
};

struct DetDev {
  float dx, dy, x0, y0, size_x, size_y;
  int32_t nx, ny;
  uint32_t npix; /* nx*(ny+1) */
  /* temporal detectors */
  uint32_t t_ndx, t_ndy, t_rort;
  float t_r;
  uint64_t t_bins, t_max;
  const float* t_posx;
  const float* t_posy;
};

/* Self-checks and evaluations of the per-step arithmetic (test support) */
enum { MCFL_SC_U2F_RZ = 0, MCFL_SC_F2U_RZ = 1, MCFL_SC_DIV = 2, MCFL_SC_PIXEL = 3 };
enum { MCFL_ME_LOG = 0, MCFL_ME_RSQRT = 1, MCFL_ME_RENORM = 2, MCFL_ME_SINCOS = 3, MCFL_ME_DRAW_OC = 4 };

/* Tally flags */
enum { MCFL_F_FHD = 1, MCFL_F_AXYZ = 2, MCFL_F_SQ = 4, MCFL_F_TEMP = 8, MCFL_F_SCALED = 16 };
/* The volume tallies a set of job flags needs. */
static inline uint32_t mcfl_vol_of(uint32_t flags) {
  return (flags & MCFL_F_AXYZ) ? (uint32_t)MCFL_VOL_ALL : ((flags & MCFL_F_FHD) ? (uint32_t)MCFL_VOL_FHD : (uint32_t)MCFL_VOL_NONE);
}

/* What the reflectance / transmittance image tally of an exit reads, one
   per job, copied from the job table into LDS when a workgroup starts: an
   exit then reads its detector from LDS instead of a chain of global loads. */
struct ExitDev {
  float x0, y0, size_x, size_y, dx, dy;
  uint32_t nx, npix, flags, pad;
  unsigned long long* Rd;
  unsigned long long* Tt;
  unsigned long long* Rd2;
  unsigned long long* Tt2;
  double rdx, rdy;   /* 1/dx, 1/dy rounded to double */
  double exit_scale; /* the job's image increment scale (MCFL_F_SCALED) */
};

struct JobDev {
  uint64_t begin;  /* first flattened photon index */
  uint64_t count;  /* photons */
  uint64_t key;    /* photon-stream key */
  uint64_t k0;     /* photon number of the job's first index (epoch slices) */
  int32_t src_kind;
  float xi, yi, zi, src_r;
  uint32_t start_weight;
  uint32_t mat_base; /* first entry of this job's material table in the launch table */
  uint32_t start_lab; /* layered medium: the layer holding zi */
  uint32_t flags;
  uint32_t tile;     /* 1: workgroups at home on this job privatise its fhd in an LDS tile */
  uint32_t tx0, ty0; /* the tile's first voxel in x and y (it starts at z = 0) */
  double exit_scale;
  unsigned long long* Rd;
  unsigned long long* Tt;
  unsigned long long* fhd;
  unsigned long long* axyz;
  unsigned long long* time;
  unsigned long long* Rd2;
  unsigned long long* Tt2;
  unsigned long long* tot; /* MCFL_NTOT counters */
  uint64_t copy_stride; /* fhd/axyz: elements between the per-XCD copies (0: one copy) */
  DetDev det;
  MediumDev med; /* used when the launch mixes media */
  /* MCFL_SRC_EMISSION: Walker alias table over emitting voxels */
  const uint32_t* al_prob; /* acceptance threshold, 32-bit fixed point */
  const uint32_t* al_alias;
  const uint32_t* al_vox;
  uint32_t al_n;
};

struct LaunchArgs {
  MediumDev med;
  const MatDev* mats; /* n_mats entries (all jobs' tables, concatenated) */
  uint32_t n_mats;
  int32_t work; /* MCFL_WORK_* */
  uint64_t total; /* size of the flattened index space */
  unsigned long long* counter; /* one work counter per job (zeroed before launch) */
  const uint32_t* a_tab; /* photon-stream multipliers */
  uint32_t a_mask;       /* table length - 1 (power of two) */
  /* MCFL_WORK_JOBS */
  const JobDev* jobs;
  uint32_t njobs;
  /* MCFL_WORK_VOXELS: job 0 is the template; voxel j emits per_job photons */
  const uint32_t* vox;
  const double* vscale;
  uint64_t per_job;
  uint64_t vseed;
  /* MCFL_WORK_LANES: lane i runs q (+1 if i < rem) photons on (xs[i], as[i]);
     MCFL_WORK_VLANES: lane i runs that many photons from each voxel vox[0..n_vox)
     in turn (the reference's emission loop), exits scaled by vscale[j] */
  unsigned long long* xs;
  const uint32_t* as;
  uint64_t lane_q, lane_rem;
  uint32_t n_vox;
  /* LDS tile of a workgroup's home job's photon hitting density: voxels
     [t0, t0+tn) per axis, t0 per job (tz0 = 0) */
  uint32_t tile_n; /* entries (0: no tile) */
  uint32_t tnx, tny, tnz;
  /* launch geometry */
  uint32_t n_waves;
  uint32_t chunk_max;
  uint32_t refill_min; /* a wave refills when this many of its lanes are idle (1..64) */
  uint32_t wg;         /* threads per workgroup: 0 (MCFL_WG), or MCFL_WG_ALT1/2 for job launches */
  uint32_t max_steps;  /* per-photon step cap; 0: MCFL_MAX_STEPS_PER_PHOTON */
  /* the volume tallies the launch's jobs need (MCFL_VOL_*): job launches run
     kernels compiled with just those; set from the job table by whoever
     builds it (MCFL_VOL_ALL, 0, is always safe) */
  uint32_t vol;
};

/* Scratch of the allocation-free excitation (mcfl_run_excitation_noalloc):
   one file-scope device array per GPU (mcfl_transport.hip), cut into the
   run's device buffers at these byte offsets (each a multiple of 256).  Its
   limits: MCFL_NA_MAX_VOXELS labels and fhd voxels (fhd and A_xyz together:
   half as many), MCFL_NA_MAX_PIX entries per image, MCFL_NA_MAX_TIME
   temporal-tally entries, MCFL_NA_MAX_TDET detector positions,
   MCFL_NA_MAX_LANES reference lanes, MCFL_MAX_MATERIALS materials. */
#define MCFL_NA_MAX_VOXELS (1u << 20)
#define MCFL_NA_MAX_PIX (1u << 17)
#define MCFL_NA_MAX_TIME (1u << 16)
#define MCFL_NA_MAX_TDET 2048u
#define MCFL_NA_MAX_LANES 32768u
enum : uint64_t {
  MCFL_NA_OFF_COUNTER = 0,
  MCFL_NA_OFF_TOT = 256,
  MCFL_NA_OFF_JOBS = 512,   /* one JobDev, at most 1536 bytes */
  MCFL_NA_OFF_MATS = 2048,  /* MCFL_MAX_MATERIALS MatDev entries */
  MCFL_NA_OFF_ZB = MCFL_NA_OFF_MATS + 16384,
  MCFL_NA_OFF_TPOS = MCFL_NA_OFF_ZB + 1024,
  MCFL_NA_OFF_XS = MCFL_NA_OFF_TPOS + 4ull * MCFL_NA_MAX_TDET,
  MCFL_NA_OFF_AS = MCFL_NA_OFF_XS + 8ull * MCFL_NA_MAX_LANES,
  MCFL_NA_OFF_LABELS = MCFL_NA_OFF_AS + 4ull * MCFL_NA_MAX_LANES,
  MCFL_NA_OFF_RD = MCFL_NA_OFF_LABELS + MCFL_NA_MAX_VOXELS,
  MCFL_NA_OFF_TT = MCFL_NA_OFF_RD + 8ull * MCFL_NA_MAX_PIX,
  MCFL_NA_OFF_RD2 = MCFL_NA_OFF_TT + 8ull * MCFL_NA_MAX_PIX,
  MCFL_NA_OFF_TT2 = MCFL_NA_OFF_RD2 + 8ull * MCFL_NA_MAX_PIX,
  MCFL_NA_OFF_TIME = MCFL_NA_OFF_TT2 + 8ull * MCFL_NA_MAX_PIX,
  MCFL_NA_OFF_VOL = MCFL_NA_OFF_TIME + 8ull * MCFL_NA_MAX_TIME, /* per-XCD copies of fhd, then of A_xyz */
  MCFL_NA_VOL_BYTES = 8ull * MCFL_TALLY_COPIES * MCFL_NA_MAX_VOXELS,
  MCFL_NA_BYTES = MCFL_NA_OFF_VOL + MCFL_NA_VOL_BYTES /* 73296896 bytes, 69.9 MiB */
};

/* Threads per workgroup of a launch. */
static inline uint32_t mcfl_wg_of(const LaunchArgs* A) { return A->wg ? A->wg : (uint32_t)MCFL_WG; }

/* Dynamic LDS of a launch: per-job totals, fhd tile, material table, layer
   bounds, per-job exit constants. */
/* fhd tile entries rounded up to an even count: the material table after the
   tile starts on a 16-byte boundary (the kernel reads a material's fields
   with 16-byte loads). */
#define MCFL_TILE_SLOTS(tile_n) (((tile_n) + 1u) & ~1u)

static inline uint32_t mcfl_lds_bytes(uint32_t nslots, uint32_t tile_n, uint32_t n_mats, uint32_t n_zb) {
  return (nslots * MCFL_NTOT + MCFL_TILE_SLOTS(tile_n)) * 8u + n_mats * (uint32_t)sizeof(MatDev) + n_zb * 8u +
         nslots * (uint32_t)sizeof(ExitDev);
}

#endif
