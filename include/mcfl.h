/*
 * mcfl.h -- C ABI of the MI355X voxelized fluorescence Monte Carlo library
 * (libmcfl.so).  Plain C types, plain pointers and sizes; no exceptions or
 * C++ types cross it.  Every entry point names the reference interface it
 * replaces (file:line in the CUDAMCFL repository).
 *
 * Units follow the reference: lengths in cm, coefficients in 1/cm, photon
 * weights as 32-bit fixed point (0xFFFFFFFF = one photon), tallies as 64-bit
 * sums of weights (CUDAMCFL.h:184-186).
 *
 * Ownership: the caller owns every host buffer it passes (RNG states,
 * output images); the library owns device memory per context.  A context is
 * used by one host thread at a time.  Failures are returned as mcfl_status;
 * the library never exits the process.
 */
#ifndef MCFL_H
#define MCFL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MCFL_STR_LEN 500        /* CUDAMCFL.h:25 */
#define MCFL_NUM_THREADS 32768  /* reference lanes, CUDAMCFL.h:18-20 */
#define MCFL_NTOT 16

typedef enum mcfl_status {
  MCFL_OK = 0,
  MCFL_ERR_ARG = 1,         /* invalid argument or inconsistent simulation */
  MCFL_ERR_IO = 2,          /* file could not be opened / written */
  MCFL_ERR_PARSE = 3,       /* malformed .mci input */
  MCFL_ERR_SEED = 4,        /* invalid RNG seed (CUDAMCFLrng.cu:60-65) */
  MCFL_ERR_PHOTONS = 5,     /* too few photons for the reference lane streams */
  MCFL_ERR_LABEL = 6,       /* voxel label outside 0..n_bulks+1 */
  MCFL_ERR_NOMEM = 7,       /* host or device allocation failed */
  MCFL_ERR_DEVICE = 8,      /* HIP runtime error / no device */
  MCFL_ERR_UNSUPPORTED = 9, /* feature outside this build (e.g. a layered medium in a batch) */
  MCFL_ERR_CAPACITY = 10    /* request exceeds a stated limit */
} mcfl_status;

/* ---- reference types (CUDAMCFL.h), same fields and meaning ---- */

typedef struct mcfl_layer { /* LayerStruct, CUDAMCFL.h:38-52 */
  float z_min, z_max, mutr, mua, g, n, flc, muaf, eY, albedof;
} mcfl_layer;

typedef struct mcfl_bulk { /* BulkStruct, CUDAMCFL.h:116-127 */
  float mutr, mua, g, n, flc, muaf, eY, albedof;
} mcfl_bulk;

typedef struct mcfl_inclusion { /* IncStruct, CUDAMCFL.h:98-114 */
  float x, y, z, r, mutr, mua, g, n, flc, muaf, eY, albedof;
} mcfl_inclusion;

typedef struct mcfl_det { /* DetStruct, CUDAMCFL.h:71-96 */
  float dx, dy, dz;
  int32_t nx, ny, nz;
  float x0, y0, sep;
  float x0_temp_det, y0_temp_det;
  uint32_t x_temp_numdets, y_temp_numdets;
  float x_temp_sepdets, y_temp_sepdets, temp_det_r;
  uint32_t temp_rort;
  uint64_t temp_bins, max_temp;
} mcfl_det;

enum {
  MCFL_SOURCE_ISOTROPIC = 0, /* isotropic point source at a voxel corner (the reference's emission source) */
  MCFL_SOURCE_PENCIL = 1,    /* collimated Gaussian spot, w = 0.03 cm (the reference's excitation source) */
  MCFL_SOURCE_DISK = 2,      /* collimated uniform disk of radius src_r (extension) */
  MCFL_SOURCE_EMISSION = 3   /* importance-sampled fluorescence emission of a whole medium (extension) */
};

typedef struct mcfl_sim { /* SimulationStruct, CUDAMCFL.h:129-170 */
  uint64_t number_of_photons;
  uint32_t number_of_photons_per_voxel;
  uint32_t n_layers;
  uint32_t n_bulks;
  uint32_t start_weight;
  char outp_filename[MCFL_STR_LEN];
  char inp_filename[MCFL_STR_LEN];
  int64_t begin, end; /* byte range of the run in the input file (echoed to _param.dat) */
  char AorB;
  mcfl_det det;
  mcfl_layer* layers;
  mcfl_bulk* bulks; /* n_bulks + 2 entries: 0 above, n_bulks+1 below */
  mcfl_inclusion inclusion;
  int32_t grid_size; /* voxels per cm */
  float esp;         /* slab thickness [cm] */
  float xi, yi, zi;  /* source position [cm] */
  float dir;         /* 0 isotropic, 1 collimated */
  float n_down;
  int32_t fhd_activated;
  int32_t do_fl_sim;
  int32_t bulk_method; /* 1: layered slab + one sphere (MCd), 2: voxel labels (MCd3D) */
  uint8_t* bulk_info; /* nx*ny*nz labels, x fastest (the reference stores int16) */
  char bulkinfo_filename[MCFL_STR_LEN];
  int32_t do_temp_sim;
  /* ---- extensions; the parser fills them so that reference inputs behave as before ---- */
  int32_t nx, ny, nz; /* voxel grid: (4 esp gs)^2 x (esp gs) unless "# mcfl-dims nx ny nz" follows the bulk filename */
  float halfx, halfy; /* lateral half extents [cm]: 2 esp, or nx/(2 gs) with the dims extension */
  int32_t src_kind;   /* MCFL_SOURCE_*; collimated sources are pencil beams unless set to DISK */
  float src_r;        /* disk radius [cm] */
} mcfl_sim;

/* ---- parser, RNG, writers (host only) ---- */

/* read_simulation_data, CUDAMCFLio.cu:243-659.  Returns the number of runs,
 * 0 on error.  Release with mcfl_free_simulations. */
int mcfl_read_simulation_data(const char* filename, mcfl_sim** simulations, int ignoreAdetection);

/* FreeSimulationStruct, CUDAMCFLmem.cu:219-223 (also frees bulks and labels). */
void mcfl_free_simulations(mcfl_sim* simulations, int n_simulations);

/* interpret_arg, CUDAMCFLio.cu:30-55 (-S<seed>, -A).  0 = OK, 1 = unknown argument. */
int mcfl_interpret_arg(int argc, char** argv, uint64_t* seed, int* ignoreAdetection);

/* init_RNG, CUDAMCFLrng.cu:31-90: per-lane MWC states and multipliers from the
 * safe-primes file (NULL or "" = the library's bundled table).  0 = OK, 1 =
 * error (missing file, invalid seed). */
int mcfl_init_rng(uint64_t* x, uint32_t* a, uint32_t n_rng, const char* safeprimes_file, uint64_t xinit);

/* Path of the bundled safeprimes_base32.txt (static string). */
const char* mcfl_default_primes_path(void);

/* Photon-stream key used by MCFL_RNG_PHOTON for stream `stream` of `seed`
 * (no counterpart in the reference; exposed so results can be reproduced). */
uint64_t mcfl_stream_key(uint64_t seed, uint64_t stream);

/* ---- device context ---- */

typedef struct mcfl_ctx mcfl_ctx;

/* Opens HIP device `device`, loads the MWC multiplier table (NULL/"" = bundled).
 * Destroy every batch and server of a context before the context. */
mcfl_status mcfl_ctx_create(int device, const char* safeprimes_file, mcfl_ctx** out);
void mcfl_ctx_destroy(mcfl_ctx* ctx);
const char* mcfl_status_string(mcfl_status s);
/* Number of compute units and resident transport workgroups per CU. */
mcfl_status mcfl_ctx_info(mcfl_ctx* ctx, int* n_cu, int* blocks_per_cu, char* name, int name_len);

/* ---- excitation ---- */

enum { MCFL_RNG_PHOTON = 0, MCFL_RNG_LANES = 1 };

typedef struct mcfl_rng {
  int32_t mode;      /* MCFL_RNG_PHOTON: one MWC stream per photon, keyed by seed; MCFL_RNG_LANES: reference lanes */
  uint64_t seed;     /* MCFL_RNG_PHOTON */
  uint64_t* x;       /* MCFL_RNG_LANES: per-lane states (advanced in place, CUDAMCFLmem.cu:42) */
  const uint32_t* a; /* MCFL_RNG_LANES: per-lane multipliers */
  uint32_t n_lanes;  /* MCFL_RNG_LANES: number of lanes (32768 in the reference) */
  /* MCFL_RNG_PHOTON (extension): the run simulates photons [first_photon,
   * first_photon + number_of_photons) of the seed's streams, 0 for a whole
   * run.  Tallies of disjoint ranges add up to those of their union: a run
   * can be split across processes or GPUs, or resumed, and summed exactly.
   * Must be 0 for MCFL_RNG_LANES. */
  uint64_t first_photon;
} mcfl_rng;

enum {
  MCFL_TOT_LAUNCHED_W = 0, /* sum of launch weights */
  MCFL_TOT_R = 1,          /* weight leaving through the top surface (whole plane) */
  MCFL_TOT_T = 2,          /* weight leaving through the bottom surface */
  MCFL_TOT_A = 3,          /* absorbed weight */
  MCFL_TOT_ROUL_GAIN = 4,  /* weight added to roulette survivors */
  MCFL_TOT_ROUL_LOSS = 5,  /* weight of roulette victims */
  MCFL_TOT_ZERO_KILL = 6,  /* weight dropped when the absorbed quantum rounds to 0 */
  MCFL_TOT_PHOTONS = 7,    /* photons launched */
  MCFL_TOT_STEPS = 8,      /* photon steps (transport loop iterations) */
  MCFL_TOT_ABORTED = 9     /* weight of photons stopped by the per-photon step cap (2^28 steps;
                              env MCFL_MAX_STEPS at context creation lowers it) */
};

typedef struct mcfl_tallies {
  uint64_t* Rd_xy;    /* det.nx*(det.ny+1), required (CUDAMCFLmem.cu:113) */
  uint64_t* Tt_xy;    /* det.nx*(det.ny+1), required */
  uint64_t* fhd;      /* nx*ny*nz photon hitting density; filled when fhd_activated and non-NULL */
  uint64_t* time_xyt; /* x_temp_numdets*y_temp_numdets*temp_bins; filled when do_temp_sim and non-NULL */
  uint64_t* A_xyz;    /* extension: absorbed weight per voxel; NULL = off */
  uint64_t* Rd2_xy;   /* extension: per-pixel sum of (w*w)>>32 for variance (exact below 2^32 exits per pixel); NULL = off */
  uint64_t* Tt2_xy;
  uint64_t totals[MCFL_NTOT];
  uint64_t num_terminated_photons; /* photons finished (the reference's counter) */
  double kernel_ms;                /* device time of the transport kernel (HIP events) */
} mcfl_tallies;

/* The device part of DoOneSimulation (CUDAMCFLmain.cu:60-167): runs one
 * excitation simulation to completion in a single persistent launch and
 * returns the integer tallies.  Outputs are overwritten.  number_of_photons
 * = 0 is MCFL_ERR_ARG (the reference's parser never produces it).
 * A voxel of fhd / A_xyz whose summed weight exceeds 2^64 - 1 returns
 * 2^64 - 1 (the reference stops adding at LLONG_MAX, CUDAMCFLtransport.cu:364).
 * On photon streams a run above 2^27 photons (env MCFL_RUN_PHOTONS at context
 * creation) is split into launches of photons [k0, k0 + n) with the same
 * results; the weight totals then saturate the same way.  Each launch keeps
 * its volume tallies in 8 per-XCD u64 copies: a launch of n photons is exact
 * while n * F < 8 * 2^32, F being the full weights one photon deposits on
 * average in the hottest voxel (6 for the reference's sample2.mci), which
 * the 2^27-photon split guarantees for F < 256.  With MCFL_TALLY_COPIES=1 in
 * the environment (one copy) the split is 2^24 photons, for the same bound.
 * Runs on MCFL_RNG_LANES are one launch whatever N: they are exact while
 * N * F < 8 * 2^32 (N * F < 2^32 with one copy). */
mcfl_status mcfl_run_excitation(mcfl_ctx* ctx, const mcfl_sim* sim, const mcfl_rng* rng, mcfl_tallies* out);

/* No counterpart in the reference.  mcfl_run_excitation for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid allocation): the same arguments, results and
 * errors, but every device buffer of the run is a slice of one fixed scratch
 * array per GPU (69.9 MiB, held from the library's load until the process
 * ends), so the call allocates no device memory.  Limits, refused with
 * MCFL_ERR_CAPACITY: at most 2^20 voxels (2^19 with both fhd and A_xyz),
 * det.nx*(det.ny+1) <= 2^17, temporal tally <= 2^16 entries and <= 2048
 * detectors, <= 32768 reference lanes, and no split run (photon streams of
 * at most 2^27 photons, env MCFL_RUN_PHOTONS).  All such calls on one GPU
 * share its scratch: calls that overlap on one GPU (two contexts or two
 * threads) race and must be serialised by the caller. */
mcfl_status mcfl_run_excitation_noalloc(mcfl_ctx* ctx, const mcfl_sim* sim, const mcfl_rng* rng, mcfl_tallies* out);

/* DoOneSimulation, CUDAMCFLmain.cu:60-193: reference lane streams (x, a of
 * MCFL_NUM_THREADS lanes, advanced in place), writes _param/_Refl/_Trans
 * [/_time] files, fills Fx = fhd/((2^32-1) N_finished) when fhd_activated,
 * returns the number of photons finished (0 on error). */
uint64_t mcfl_DoOneSimulation(mcfl_ctx* ctx, mcfl_sim* simulation, uint64_t* x, uint32_t* a, double* Fx);

/* ---- fluorescence (emission) stage ---- */

/* DoOneSimulationFl, CUDAMCFLmain.cu:196-288: one emission simulation of the
 * reference's voxel loop, for callers that keep that loop (:471-594) and
 * call this once per voxel.  The simulation is used as the caller left it
 * (the loop sets mua <- muaf, dir = 0, number_of_photons = photons per voxel
 * and the source at the voxel corner, :453-487); no fhd is tallied.  Runs on
 * the reference lanes x, a (MCFL_NUM_THREADS of them, advanced in place, the
 * first photon of each lane replaying its launch draws as
 * LaunchPhoton_Global does), fills R_out and T_out (det.nx*(det.ny+1) raw
 * sums each, the reference's tempvoxelR/T) and returns the photons finished,
 * 0 on failure (fewer photons than lanes, as the reference, or a device
 * error).  mcfl_run_fluorescence_lanes runs the whole loop in one launch. */
uint64_t mcfl_DoOneSimulationFl(mcfl_ctx* ctx, mcfl_sim* simulation, uint64_t* x, uint32_t* a, uint64_t* R_out,
                                uint64_t* T_out);


typedef struct mcfl_fl_result {
  uint64_t* FlR_raw; /* optional det.nx*(det.ny+1) fixed-point tallies */
  uint64_t* FlT_raw;
  double unit;       /* image value of one raw unit times dx*dy */
  uint64_t voxels;   /* emitting voxels simulated */
  uint64_t photons;  /* emission photons simulated */
  uint64_t totals[MCFL_NTOT];
  double kernel_ms;
} mcfl_fl_result;

/* Emission plan of the reference's voxel loop (CUDAMCFLmain.cu:471-562):
 * every voxel whose material has albedof < 1 emits n_per_voxel photons with
 * weight voxelw = eY (1-albedof) Fx / (n_per_voxel (2^32-1)).  Fills vox[]
 * (voxel indices, ascending) and scale[] (voxelw / unit); *n_vox in: capacity,
 * out: count. */
mcfl_status mcfl_fl_plan(const mcfl_sim* sim, const double* Fx, uint64_t n_per_voxel, uint32_t* vox, double* scale,
                         uint64_t* n_vox, double* unit);

/* The fluorescence voxel loop (CUDAMCFLmain.cu:409-647 minus file output) in
 * one persistent launch over all emitting voxels: mua <- muaf for every
 * material (CUDAMCFLmain.cu:458-461), isotropic sources at voxel corners,
 * MCFL_RNG_PHOTON streams keyed by (seed, voxel).  FlR/FlT (det.nx*det.ny,
 * may be NULL) receive the reference's images. */
mcfl_status mcfl_run_fluorescence(mcfl_ctx* ctx, const mcfl_sim* sim, const double* Fx, uint64_t n_per_voxel,
                                  uint64_t seed, double* FlR, double* FlT, mcfl_fl_result* res);

/* The same voxel loop on the reference's lane streams (DoOneSimulationFl per
 * voxel, CUDAMCFLmain.cu:196-288 inside the loop :471-594): voxels in the
 * reference's order (x outermost, z innermost), each one simulation of
 * n_per_voxel photons on rng's MCFL_RNG_LANES lanes, which continue from voxel
 * to voxel (rng->x advanced in place, CUDAMCFLmem.cu:42); in every voxel a
 * lane's first photon replays its launch draws (CUDAMCFLtransport.cu:574-581).
 * That replay correlates the first photon's launch with its first steps, and
 * with n_per_voxel close to the lane count (as the reference's >= 32769
 * photons on 32768 lanes) most emission photons are first photons: the images
 * then carry the reference's own bias (percent-level), which the photon-stream
 * loop above does not.  n_per_voxel < n_lanes is MCFL_ERR_PHOTONS (the
 * reference's DoOneSimulationFl fails, CUDAMCFLmain.cu:268-272). */
mcfl_status mcfl_run_fluorescence_lanes(mcfl_ctx* ctx, const mcfl_sim* sim, const double* Fx, uint64_t n_per_voxel,
                                        const mcfl_rng* rng, double* FlR, double* FlT, mcfl_fl_result* res);

/* ---- batched launches ----
 * No counterpart in the reference, whose multi-source scans run one process
 * per .mci file (ScriptGenerator-CeCAR-Translation.ipynb:48-55,
 * cudamcfl-array.sh:15).  Many simulations -- different sources, budgets and
 * media -- share one persistent launch and one photon queue; each job's
 * photons draw from streams keyed by its own seed, so every job's tallies are
 * identical to a run of that job alone.  Media are voxel media (bulk method
 * 2) or one layered slab (bulk method 1) per batch, which emits no
 * fluorescence (MCFL_ERR_UNSUPPORTED for a layered medium beside another
 * medium or under an emission job).  Jobs take at most 2^32 - 1 photons
 * (MCFL_ERR_CAPACITY).  A batch with a job above the photon split of
 * mcfl_run_excitation (2^27, env MCFL_RUN_PHOTONS) runs in slices of at most
 * that many photons per job, its fhd / A_xyz summed between slices into
 * accumulators that saturate at 2^64 - 1 and its totals summed the same way,
 * exactly as mcfl_run_excitation splits a long run: the results equal that
 * call's.  Served requests (below) tally fhd only internally, for their
 * emission pass, and sum it into an accumulator after every 2^27 photons the
 * same way. */

enum { MCFL_TALLY_FHD = 1, MCFL_TALLY_SQ = 2, MCFL_TALLY_TEMP = 4, MCFL_TALLY_AXYZ = 8 };
#define MCFL_MAX_BATCH_JOBS 64

typedef struct mcfl_job {
  int32_t medium;          /* index into the media list */
  int32_t src_kind;        /* MCFL_SOURCE_*; EMISSION needs Fx */
  float xi, yi, zi, src_r; /* source position and disk radius [cm] */
  uint64_t photons;        /* photon budget */
  uint64_t seed;           /* photon-stream seed */
  uint32_t tallies;        /* MCFL_TALLY_* */
  const double* Fx;        /* EMISSION: normalised excitation fhd of the medium, nx*ny*nz */
  /* outputs: caller-owned host buffers, Rd_xy/Tt_xy required, others as enabled */
  uint64_t* Rd_xy;
  uint64_t* Tt_xy;
  uint64_t* fhd;
  uint64_t* A_xyz;
  uint64_t* Rd2_xy;
  uint64_t* Tt2_xy;
  uint64_t* time_xyt;
  uint64_t totals[MCFL_NTOT];
  double fl_unit; /* EMISSION: image value (times dx*dy) of one raw unit, sum(eY (1-albedof) Fx) / (photons (2^32-1)) */
} mcfl_job;

typedef struct mcfl_batch mcfl_batch;

/* Uploads the media, allocates every job's device tallies and captures the
 * batch (tally clearing + one transport launch) in a HIP graph. */
mcfl_status mcfl_batch_create(mcfl_ctx* ctx, const mcfl_sim* const* media, int n_media, const mcfl_job* jobs,
                              int n_jobs, mcfl_batch** out);
/* Replays the graph with jobs[i].seed and copies each job's outputs into jobs[i]. */
mcfl_status mcfl_batch_run(mcfl_batch* b, mcfl_job* jobs, double* kernel_ms);
void mcfl_batch_destroy(mcfl_batch* b);

/* Launch settings of a batch picked by mcfl_batch_autotune. */
typedef struct mcfl_tune {
  int32_t tile_side;    /* lateral size of the fhd LDS tile in voxels (0: no fhd tally) */
  uint32_t tile_voxels; /* entries of the tile */
  uint32_t refill_min;  /* idle lanes of a wave that trigger its refill */
  uint32_t chunk_max;   /* largest photon chunk a wave takes from the work counter */
  double trial_ms;      /* device time of the chosen settings on the scaled-down batch */
  double default_ms;    /* the same for the default settings */
  int32_t trials;       /* trial launches made */
  uint32_t wg;          /* threads per workgroup (512, 768 or 1024) */
} mcfl_tune;

/* No counterpart in the reference.  Times candidate launch settings on a copy
 * of the batch with every budget scaled by `fraction` (at least 4096 photons
 * per job): first the workgroup size (512, 768, 1024 threads), then, at that
 * size, fhd tile side x refill threshold x chunk size; keeps the fastest and
 * re-captures the graph with it.  Results do not depend on these settings. */
mcfl_status mcfl_batch_autotune(mcfl_batch* b, double fraction, mcfl_tune* out);
/* No counterpart in the reference.  Applies launch settings (as returned by
 * mcfl_batch_autotune, e.g. tuned once on another batch of the same shape):
 * wg (0 or 768, 512, 1024), tile_side (0: the LDS budget rule), refill_min
 * (1..64), chunk_max (64..2^20); the other fields are ignored.  Re-captures
 * the graph.  Results do not depend on these settings. */
mcfl_status mcfl_batch_configure(mcfl_batch* b, const mcfl_tune* in);
/* create + run + destroy */
mcfl_status mcfl_run_batch(mcfl_ctx* ctx, const mcfl_sim* const* media, int n_media, mcfl_job* jobs, int n_jobs,
                           double* kernel_ms);

/* ---- serving: continuous batching of forward-simulation requests ----
 * Requests are admitted while their device footprint fits the memory budget
 * and a job slot is free; each step runs one epoch kernel over a slice of
 * every admitted request's photons, tallies stay resident in HBM between
 * epochs, finished requests release their memory.  A fluorescence request
 * runs its excitation, then an importance-sampled emission pass of the same
 * budget.  A request's result does not depend on what it shared epochs with. */

typedef struct mcfl_server mcfl_server;

typedef struct mcfl_request {
  int32_t medium; /* id returned by mcfl_server_add_medium */
  int32_t src_kind;
  float xi, yi, zi, src_r;
  uint64_t photons;
  uint64_t seed;
  int32_t fluorescence; /* 1: excitation + emission, 0: excitation only */
} mcfl_request;

typedef struct mcfl_request_result {
  uint64_t* Rd_xy; /* det.nx*(det.ny+1) each; caller-owned */
  uint64_t* Tt_xy;
  uint64_t* FlR_raw; /* fluorescence requests */
  uint64_t* FlT_raw;
  double fl_unit;
  uint64_t totals[MCFL_NTOT];
  uint64_t fl_totals[MCFL_NTOT];
  uint64_t epochs; /* epochs the request took part in */
} mcfl_request_result;

/* Each epoch is one launch of about 2e7 photons (env MCFL_EPOCH_TARGET): every
   active request gets at least epoch_photons of its phase; of the rest of the
   target the oldest request takes 3/4 (env MCFL_OLDEST_SHARE) and the others
   share what remains, least work left first.  Admission is
   FIFO within mem_budget_bytes.  Two epochs are in flight at a time on two
   streams; an epoch leaves 8 workgroup slots free (env MCFL_SERVER_RESERVE)
   so that finished requests' tallies are copied out beside it. */
mcfl_status mcfl_server_create(mcfl_ctx* ctx, uint64_t mem_budget_bytes, uint64_t epoch_photons, mcfl_server** out);
/* Makes a voxel medium resident (a layered one is MCFL_ERR_UNSUPPORTED); the
   simulation's bulks/labels/detector are copied. */
mcfl_status mcfl_server_add_medium(mcfl_server* s, const mcfl_sim* medium, int32_t* medium_id);
/* Queues a request (admitted at the next step that has room). */
mcfl_status mcfl_server_submit(mcfl_server* s, const mcfl_request* req, int64_t* request_id);
/* Admits, launches the next epoch, completes the oldest one in flight and
   retires finished requests (their ids in done_ids). */
mcfl_status mcfl_server_step(mcfl_server* s, int64_t* done_ids, int32_t max_done, int32_t* n_done, double* kernel_ms);
/* Copies a finished request's results and forgets it. */
mcfl_status mcfl_server_collect(mcfl_server* s, int64_t request_id, mcfl_request_result* out);
mcfl_status mcfl_server_stats(mcfl_server* s, int32_t* queued, int32_t* active, uint64_t* bytes_in_use);
/* Cumulative photons launched, epochs, device time during which an epoch ran
   (two epochs overlapping in time count once) and host time spent in steps
   outside the waits for the device. */
mcfl_status mcfl_server_counters(mcfl_server* s, uint64_t* photons, uint64_t* epochs, double* kernel_ms, double* host_ms);
void mcfl_server_destroy(mcfl_server* s);

/* ---- writers (byte layouts of the reference) ---- */

/* Write_Simulation_Results, CUDAMCFLio.cu:57-191: _param.dat, _Trans.dat,
 * _Refl.dat and (do_temp_sim) _time.dat.  0 = OK. */
int mcfl_write_simulation_results(const mcfl_sim* sim, const mcfl_tallies* t, double simulation_time_s);
/* _PHD-Ascii.dat, CUDAMCFLmain.cu:368-392. */
int mcfl_write_phd(const mcfl_sim* sim, const double* Fx);
/* _FlR.dat and _FlT.dat, CUDAMCFLmain.cu:604-647 (images det.nx*det.ny). */
int mcfl_write_fl(const mcfl_sim* sim, const double* FlR, const double* FlT);

#ifdef __cplusplus
}
#endif
#endif /* MCFL_H */
