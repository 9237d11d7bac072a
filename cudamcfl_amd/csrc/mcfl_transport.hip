/*
 * mcfl_transport.hip -- persistent gfx950 kernel for CUDAMCFL's voxelized
 * photon transport (bulk method 2).
 *
 * Replaces the reference's MCd3D / LaunchPhoton_Global kernels and the host
 * relaunch loop around them (CUDAMCFLtransport.cu:242-456,460-582;
 * CUDAMCFLmain.cu:117-165,229-273).  One launch runs a whole simulation, a
 * batch of simulations, or the fluorescence stage of every emitting voxel:
 *
 *  - Work: a flattened photon index space.  Each wavefront keeps a chunk of
 *    indices; lanes whose photon died are found with a ballot, ranked with
 *    mbcnt, and given the next indices of the chunk; an empty chunk is
 *    refilled with one 64-bit atomic on a global counter by the wave's first
 *    lane (guided sizes: large early, small near the end).  No host round
 *    trips, no relaunches.
 *  - RNG: each photon (or each reference lane) owns an MWC stream
 *    (mcfl_mwc.h), so results do not depend on which lane runs what.
 *  - Medium: 8-bit labels in HBM (L2/MALL resident), per-material constants
 *    in LDS.
 *  - Tallies: 64-bit integer atomics (order independent); per-job scalar
 *    totals are privatised in LDS and flushed once per workgroup.  The
 *    per-step volume tallies (photon hitting density, absorbed weight) go to
 *    one copy per XCD, chosen by the hardware XCC_ID, so their cache lines
 *    never move between the XCDs' L2 caches; a reduce kernel sums the copies.
 *  - Variants: the launch's shape is compiled in (one job or several, one
 *    medium or several, time-of-flight tally or not, the volume tallies its
 *    jobs carry), so a launch holds only the registers and code it uses.
 *
 * The step follows CUDAMCFLtransport.cu:289-447 quirk for quirk (end-point
 * label lookup, z-normal Fresnel, absorption and spin at a surface point after
 * internal reflection, sin/cos swap in Spin, integer weights, roulette).
 */
#include <hip/hip_runtime.h>

#include <mutex>

#include "mcfl_device.h"
#include "mcfl_mwc.h"


namespace {

constexpr uint32_t kWeightI = 429497u;   /* CUDAMCFL.h:33 */
constexpr float kChance = 0.1f;          /* CUDAMCFL.h:34 */
constexpr float kPi = 3.141592654f;      /* CUDAMCFL.h:22 */
constexpr float kFltMax = 3.402823466e+38f;

struct Ph {
  float x, y, z, dx, dy, dz;
  uint32_t w;   /* weight, 2^-32 fixed point */
  uint32_t lab; /* current medium label (bulkpos) */
  uint64_t tof; /* time of flight [fs] */
};

__device__ __forceinline__ bool in_grid(const MediumDev& m, float x, float y, float z) {
  return mcfl_fabs(x) < m.halfx && mcfl_fabs(y) < m.halfy && z < m.esp;
}

/* CUDAMCFLtransport.cu:323-325 with each component clamped into the grid. */
__device__ __forceinline__ uint32_t vox_of(const MediumDev& m, float x, float y, float z) {
  uint32_t ix = min(mcfl_f2u_rz((x + m.halfx) * m.gs), (uint32_t)m.nx - 1u);
  uint32_t iy = min(mcfl_f2u_rz((y + m.halfy) * m.gs), (uint32_t)m.ny - 1u);
  uint32_t iz = min(mcfl_f2u_rz(z * m.gs), (uint32_t)m.nz - 1u);
  return ix + (uint32_t)m.nx * (iy + (uint32_t)m.ny * iz);
}

/* Cell of an in-grid point (|x| < halfx, |y| < halfy, z < esp), the same
   integers as vox_of: x + halfx and y + halfy are >= 0, so the truncating
   conversion needs no saturation; z can be a rounding below 0 on the top
   surface, which the saturating conversion takes to 0. */
struct Cell {
  uint32_t ix, iy, iz;
};
__device__ __forceinline__ Cell cell_in(const MediumDev& m, float x, float y, float z) {
  Cell c;
  c.ix = min((uint32_t)((x + m.halfx) * m.gs), (uint32_t)m.nx - 1u);
  c.iy = min((uint32_t)((y + m.halfy) * m.gs), (uint32_t)m.ny - 1u);
  c.iz = min(mcfl_f2u_rz(z * m.gs), (uint32_t)m.nz - 1u);
  return c;
}
/* nx, ny < 2^24 and nx*ny*nz < 2^32 (host validation) keep the products exact */
__device__ __forceinline__ uint32_t cell_index(const MediumDev& m, Cell c) {
  return c.ix + __umul24(c.iy, (uint32_t)m.nx) + c.iz * m.nxy;
}

/* Inside the layered medium's spherical inclusion (CUDAMCFLtransport.cu:83-86, 201-204). */
__device__ __forceinline__ bool in_sphere(const MediumDev& m, float x, float y, float z) {
  const float dx = x - m.inc_x, dy = y - m.inc_y, dz = z - m.inc_z;
  return (dx * dx + dy * dy) + dz * dz <= m.inc_r2;
}

typedef __attribute__((address_space(1))) unsigned long long gu64;
typedef __attribute__((address_space(1))) const uint32_t gcu32;

/* p[i] through a global (not generic) pointer: a global load instead of a flat one */
__device__ __forceinline__ uint32_t gload_u32(const uint32_t* p, uint32_t i) { return ((gcu32*)p)[i]; }
/* The job table does not change during a launch: read through the constant
   address space, its wave-uniform reads are scalar loads, which wait neither
   for nor behind the wave's outstanding vector memory operations (the
   compiler cannot prove that the tally atomics leave generic job-table
   memory untouched, and would issue vector loads). */
typedef __attribute__((address_space(4))) const JobDev cjob;
__device__ __forceinline__ cjob* ctab(const JobDev* p) { return (cjob*)p; }

__device__ __forceinline__ void atomic_add_u64(unsigned long long* p, unsigned long long v) {
  __hip_atomic_fetch_add((gu64*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void lds_add_u64(unsigned long long* p, unsigned long long v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

/* LaunchPhoton, CUDAMCFLtransport.cu:460-562 (method 2). */
__device__ void launch(const MediumDev& m, int kind, float xi, float yi, float zi, float src_r, uint32_t sw, Ph& p,
                       uint64_t& x, uint32_t a, bool layered, uint32_t start_layer) {
  if (kind == MCFL_SRC_ISOTROPIC) {
    const float step = 0.999999f / m.gs;
    p.x = xi + step * mcfl_rand_oc(&x, a);
    p.y = yi + step * mcfl_rand_oc(&x, a);
    p.z = zi + step * mcfl_rand_oc(&x, a);
    float ct = (float)(1.0 - 2.0 * (double)mcfl_rand_oc(&x, a));
    float st = (float)sqrt(1.0 - (double)(ct * ct));
    float u = mcfl_rand_oc(&x, a);
    float psi = (float)(2.0 * (double)kPi * (double)u);
    float sn, cp;
    mcfl_sincos_turn(u, &sn, &cp);
    float sp = mcfl_sqrt((float)(1.0 - (double)(cp * cp)));
    if (!(psi < kPi)) sp = -sp;
    p.dx = st * cp;
    p.dy = st * sp;
    p.dz = ct;
    p.w = 0xFFFFFFFFu;
  } else {
    float u1 = mcfl_rand_oc(&x, a);
    float u2 = mcfl_rand_oc(&x, a);
    float rad = (kind == MCFL_SRC_PENCIL) ? 0.03f * mcfl_sqrt(-mcfl_log(u1)) : src_r * mcfl_sqrt(u1);
    float sn, cs;
    mcfl_sincos_turn(u2, &sn, &cs);
    p.x = xi + rad * cs;
    p.y = yi + rad * sn;
    p.z = zi;
    p.dx = 0.0f;
    p.dy = 0.0f;
    p.dz = 1.0f;
    p.w = sw;
  }
  if (layered) /* the layer holding the source depth (:547-560) */
    p.lab = start_layer;
  else if (p.z < m.esp)
    p.lab = (mcfl_fabs(p.x) < m.halfx && mcfl_fabs(p.y) < m.halfy) ? (uint32_t)m.labels[vox_of(m, p.x, p.y, p.z)] : 1u;
  else
    p.lab = m.last;
  p.tof = 0;
}

/* An emission photon from the corner of voxel v (CUDAMCFLmain.cu:481-487). */
__device__ __forceinline__ void launch_voxel(const MediumDev& m, uint32_t v, Ph& p, uint64_t& x, uint32_t a) {
  const uint32_t nxy = (uint32_t)m.nx * (uint32_t)m.ny;
  const uint32_t iz = v / nxy, rem = v - iz * nxy, iy = rem / (uint32_t)m.nx, ix = rem - iy * (uint32_t)m.nx;
  launch(m, MCFL_SRC_ISOTROPIC, ((float)ix / m.gs) - m.halfx, ((float)iy / m.gs) - m.halfy, ((float)iz / m.gs), 0.0f,
         0u, p, x, a, false, 0u);
}

/* a / b, the correctly rounded quotient, for a and b whose quotient needs
   none of the range scaling of the IEEE division sequence (|a| and |b| in
   [2^-100, 2^100], normal quotient): the same reciprocal, Newton step and two
   FMA corrections, without the scale and special-case fix-up steps that leave
   such operands unchanged.  Equal to a / b for Spin's quotients (1 - g^2) /
   (1 - g + 2 g xi), checked on the GPU against the division for every
   denominator of several g. */
__device__ __forceinline__ float div_unscaled(float a, float b) {
  float r = __builtin_amdgcn_rcpf(b);
  r = mcfl_fma(mcfl_fma(-b, r, 1.0f), r, r);
  float q = a * r;
  q = mcfl_fma(mcfl_fma(-b, q, a), r, q);
  return mcfl_fma(mcfl_fma(-b, q, a), r, q);
}

/* A material from the LDS table: its first 32 bytes (the fields of the step,
   the absorption and Spin) as two 16-byte loads; the table is 16-byte
   aligned. */
__device__ __forceinline__ MatDev load_mat(const MatDev* t) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  const f4* q = reinterpret_cast<const f4*>(t);
  const f4 a = q[0], b = q[1];
  MatDev r;
  r.mutr = a.x;
  r.mua_mutr = a.y;
  r.n = a.z;
  r.onemg = a.w;
  r.onemg2 = b.x;
  r.onepg2 = b.y;
  r.twog32 = b.z;
  r.inv2g = b.w;
  r.q_top = t->q_top;
  r.r_top = t->r_top;
  r.q_bot = t->q_bot;
  r.r_bot = t->r_bot;
  return r;
}

/* Spin, CUDAMCFLtransport.cu:585-625. */
__device__ __forceinline__ void spin(const MatDev& mt, Ph& p, uint64_t& x, uint32_t a) {
  /* 1 - g + 2g xi with xi = U 2^-32 (U the draw converted toward zero) as
     1 - g + (2g 2^-32) U: the same real product, so the same FMA, wherever
     2g 2^-32 is exact (|2g| >= 2^-94); below that both round to 1 - g = 1 */
  const float u = mcfl_u2f_rz(mcfl_mwc_next(&x, a));
  float t = div_unscaled(mt.onemg2, mcfl_fma(mt.twog32, u, mt.onemg));
  float cost = mcfl_fma(-t, t, mt.onepg2) * mt.inv2g;
  if (mt.inv2g == 0.0f) cost = 2.0f * mcfl_rand_co(&x, a) - 1.0f; /* g == 0 */
  const float s2 = mcfl_clamp0(mcfl_fma(-cost, cost, 1.0f)); /* sin^2 theta: 0 or >= 2^-23 */
  float sps, cps;
  mcfl_rand_sincos(&x, a, &sps, &cps);
  const float cosp = sps, sinp = cps; /* the reference's __sincosf argument order */
  const float dx = p.dx, dy = p.dy, dz = p.dz;
  const float q = mcfl_clamp0(mcfl_fma(-dz, dz, 1.0f)); /* temp^2 = 1 - dz^2: 0 or >= 6e-8 */
  float nx, ny, nz;
  if (q == 0.0f) { /* the reference's temp == 0 test */
    const float sint = mcfl_sqrt_rs(s2);
    nx = sint * cosp;
    ny = sint * sinp;
    nz = mcfl_copysign(cost, dz * cost);
  } else {
    /* sint/temp and sint*temp from one reciprocal root of sint^2 temp^2;
       nx = dx (cost + stt cosp dz) - dy stt sinp, ny = dy (...) + dx stt sinp */
    const float pq = s2 * q;
    const float r = pq > 0.0f ? mcfl_rsqrt(pq) : 0.0f;
    const float stt = s2 * r; /* sint / temp */
    const float stq = pq * r; /* sint * temp */
    const float c = cost + (stt * cosp) * dz, w = stt * sinp;
    nx = mcfl_fma(dx, c, -(dy * w));
    ny = mcfl_fma(dy, c, dx * w);
    nz = mcfl_fma(-stq, cosp, dz * cost);
  }
  const float r = mcfl_renorm(mcfl_fma(nz, nz, mcfl_fma(ny, ny, nx * nx)));
  p.dx = nx * r;
  p.dy = ny * r;
  p.dz = nz * r;
}

/* Reflect, CUDAMCFLtransport.cu:628-700 (bulk method 2).  With PRE, against
   the outer media (nb 0 or `last`) the two quotients come precomputed with
   the same rounding (the variant with the most live registers divides). */
template <bool PRE>
__device__ __forceinline__ void reflect(const MatDev* mats, Ph& p, uint32_t nb, uint32_t last, uint64_t& x,
                                        uint32_t a) {
  const MatDev& m1 = mats[p.lab];
  const float n1 = m1.n, n2 = mats[nb].n;
  const float ci = mcfl_fabs(p.dz);
  if (n1 == n2) {
    p.lab = nb;
    return;
  }
  const float ci2 = ci * ci;
  if (n2 * n2 < n1 * n1 * (1.0f - ci2)) {
    p.dz = -p.dz;
    return;
  }
  if (ci == 1.0f) {
    float r = (n1 - n2) / (n1 + n2);
    if (mcfl_rand_co(&x, a) <= r * r)
      p.dz = -p.dz;
    else
      p.lab = nb;
    return;
  }
  const float q = PRE && nb == 0u ? m1.q_top : (PRE && nb == last ? m1.q_bot : (n1 * n1) / (n2 * n2));
  float e = q * (1.0f - ci2);
  float r = 2.0f * mcfl_sqrt(mcfl_fmax(0.0f, (1.0f - ci2) * (1.0f - e) * e * ci * ci));
  e = e + ci2 * (1.0f - 2.0f * e);
  r = e * ((1.0f - e - r) / ((1.0f - e + r) * (e + r)));
  if (mcfl_rand_co(&x, a) <= r) {
    p.dz = -p.dz;
    return;
  }
  r = PRE && nb == 0u ? m1.r_top : (PRE && nb == last ? m1.r_bot : n1 / n2);
  e = r * r * (1.0f - ci2);
  p.dx *= r;
  p.dy *= r;
  p.dz = mcfl_copysign(mcfl_sqrt(mcfl_fmax(0.0f, 1.0f - e)), p.dz);
  p.lab = nb;
}

/* MoveToFirstBoundary, CUDAMCFLtransport.cu:702-847: half-voxel march through
 * a glass (mus = 0) region, then back off to the sphere of radius 0.707107/gs
 * around the first voxel of a different label. */
template <bool TEMPORAL>
__device__ uint32_t glass_march(const MediumDev& m, Ph& p, uint32_t old_lab, float max_s) {
  float step = 1.0f / (2.0f * m.gs);
  float moved = 0.0f;
  int64_t index = 1;
  uint32_t present = old_lab;
  /* A march inside a region whose label is an exit (0 or n_bulks+1, e.g. a
     photon launched there) never sees its label change: at the surface its
     step shrinks to nothing and the reference (CUDAMCFLtransport.cu:717)
     loops for ever.  Every other march advances 1/(2 gs) per pass and ends
     within about max_s * 2 gs passes (the float sum of the steps drifts by
     far less than that), so twice that bound only stops the endless ones. */
  const uint32_t passes = (uint32_t)(max_s * 4.0f * m.gs) + 64u;
  uint32_t pass = 0;
  while (present == old_lab && moved < max_s && pass++ < passes) {
    p.x = mcfl_fma(p.dx, step, p.x);
    p.y = mcfl_fma(p.dy, step, p.y);
    p.z = mcfl_fma(p.dz, step, p.z);
    if (TEMPORAL) p.tof += mcfl_tof_inc(step);
    if (mcfl_fma(p.dz, step, p.z) < 0.0f) {
      present = 0;
      index = -1;
      step = -p.z / p.dz;
    } else if (mcfl_fma(p.dz, step, p.z) > m.esp) {
      present = m.last;
      index = -1;
      step = (m.esp - p.z) / p.dz;
    } else if (mcfl_fabs(mcfl_fma(p.dx, step, p.x)) >= m.halfx || mcfl_fabs(mcfl_fma(p.dy, step, p.y)) >= m.halfy) {
      present = 1;
      index = -1;
    } else {
      index = vox_of(m, p.x, p.y, p.z);
      present = m.labels[index];
    }
    moved += step;
  }
  if (index >= 0 && present != old_lab) {
    const uint32_t nxy = (uint32_t)m.nx * (uint32_t)m.ny;
    const uint32_t iz = (uint32_t)index / nxy;
    const uint32_t rem = (uint32_t)index - iz * nxy;
    const uint32_t iy = rem / (uint32_t)m.nx;
    const uint32_t ix = rem - iy * (uint32_t)m.nx;
    const double half = 1.0 / (2.0 * (double)m.gs);
    const float cx = (float)((double)((float)ix / m.gs - m.halfx) + half);
    const float cy = (float)((double)((float)iy / m.gs - m.halfy) + half);
    const float cz = (float)((double)((float)iz / m.gs) + half);
    const double rv = 0.707107 * (double)(1.0f / m.gs);
    const float r2 = (float)(rv * rv);
    /* roots of |p - t d - c|^2 = r2 */
    const float qx = p.x - cx, qy = p.y - cy, qz = p.z - cz;
    const float dd = p.dx * p.dx + p.dy * p.dy + p.dz * p.dz;
    const float b = p.dx * qx + p.dy * qy + p.dz * qz;
    const float cc = (qx * qx + qy * qy + qz * qz) - r2;
    const float sq = mcfl_sqrt(mcfl_fmax(0.0f, b * b - dd * cc));
    const float t1 = (b - sq) / dd, t2 = (b + sq) / dd;
    const float x1 = p.x - p.dx * t1, y1 = p.y - p.dy * t1, z1 = p.z - p.dz * t1;
    const float x2 = p.x - p.dx * t2, y2 = p.y - p.dy * t2, z2 = p.z - p.dz * t2;
    const float d1 = (x1 - cx) * (x1 - cx) + (y1 - cy) * (y1 - cy) + (z1 - cz) * (z1 - cz);
    const float d2 = (x2 - cx) * (x2 - cx) + (y2 - cy) * (y2 - cy) + (z2 - cz) * (z2 - cz);
    const bool first = d1 < d2;
    p.x = first ? x1 : x2;
    p.y = first ? y1 : y2;
    p.z = first ? z1 : z2;
    if (TEMPORAL) p.tof -= mcfl_tof_inc(first ? t1 : t2);
  }
  return present;
}

/* Reflectance / transmittance exit tally, CUDAMCFLtransport.cu:369-421: the
   image from the job's exit constants in LDS (E), the temporal detectors
   from its job table entry (J). */
template <bool TEMPORAL>
__device__ void exit_tally(const JobDev& J, const ExitDev& E, const Ph& p, bool top, bool scaled, double escale,
                           unsigned long long* st) {
  lds_add_u64(&st[top ? MCFL_DT_R : MCFL_DT_T], p.w);
  if (mcfl_fabs(p.x - E.x0) < E.size_x && mcfl_fabs(p.y - E.y0) < E.size_y) {
    /* the float quotients (p - origin + half size) / pixel size as the double
       product with the pixel size's reciprocal, rounded to float: within
       2^-52 of the quotient, which is 2^-49 or farther from every rounding
       boundary of a float it is not equal to, so the same float */
    const uint32_t iy = mcfl_f2u_rz((float)((double)(p.y - E.y0 + E.size_y) * E.rdy));
    const uint32_t ix = mcfl_f2u_rz((float)((double)(p.x - E.x0 + E.size_x) * E.rdx));
    const uint64_t idx = (uint64_t)iy * (uint64_t)E.nx + ix;
    if (idx < E.npix) {
      unsigned long long inc = p.w;
      if (__ballot(scaled) != 0ull && scaled) inc = (unsigned long long)((double)p.w * escale + 0.5);
      atomic_add_u64((top ? E.Rd : E.Tt) + idx, inc);
      if (E.flags & MCFL_F_SQ)
        atomic_add_u64((top ? E.Rd2 : E.Tt2) + idx, ((unsigned long long)p.w * (unsigned long long)p.w) >> 32);
    }
  }
  const DetDev& d = J.det;
  if (TEMPORAL && (J.flags & MCFL_F_TEMP) && d.t_rort == (top ? 0u : 1u) && p.tof < d.t_max) {
    for (uint32_t xp = 0; xp < d.t_ndx; ++xp)
      for (uint32_t yp = 0; yp < d.t_ndy; ++yp) {
        const float ddx = p.x - d.t_posx[xp], ddy = p.y - d.t_posy[yp];
        if (ddx * ddx + ddy * ddy < d.t_r * d.t_r) {
          const uint32_t bin = mcfl_f2u_rz(((float)p.tof / (float)d.t_max) * (float)d.t_bins);
          if (bin < d.t_bins) atomic_add_u64(J.time + (xp + d.t_ndx * (yp + d.t_ndy * (uint64_t)bin)), p.w);
        }
      }
  }
}

__device__ __forceinline__ uint32_t lane_rank(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

/* The job whose flattened index range holds g (jobs in begin order). */
__device__ __forceinline__ uint32_t job_at(const LaunchArgs& A, uint64_t g) {
  const JobDev* J = A.jobs;
  uint32_t lo = 0, hi = A.njobs; /* last job with begin <= g */
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (J[mid].begin <= g)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

} /* namespace */

/* resident waves per SIMD the register allocation is sized for: a 1024-thread
   workgroup (4 waves per SIMD) is alone on its CU */
constexpr int waves_per_eu_of(int wg) { return wg > 768 ? 4 : MCFL_WAVES_PER_EU; }

template <int WORK, bool TEMPORAL, bool PERJOB, bool LAYERED, int WGS, int VM>
__global__ __launch_bounds__(WGS) __attribute__((amdgpu_waves_per_eu(waves_per_eu_of(WGS)))) void mcfl_transport_kernel(LaunchArgs A) {
  /* dynamic LDS: [per-job totals | fhd tile | materials | layer bounds | per-job exit constants] */
  extern __shared__ __align__(16) unsigned long long lds[];

  /* JOBS: photons of the job table (ONE: a table of one job) */
  constexpr bool JOBS = WORK == MCFL_WORK_JOBS || WORK == MCFL_WORK_ONE;
  constexpr bool ONE = WORK == MCFL_WORK_ONE;
  /* volume tallies: fhd and A_xyz (VM = MCFL_VOL_ALL), fhd alone
     (MCFL_VOL_FHD) or none (MCFL_VOL_NONE) compiled in, as the launch's jobs
     need them; the voxel-emitter launches of the fluorescence stage never
     carry them */
  constexpr bool VOL = VM != MCFL_VOL_NONE && WORK != MCFL_WORK_VOXELS && WORK != MCFL_WORK_VLANES;
  constexpr bool AX = VOL && VM == MCFL_VOL_ALL;
  const uint32_t tid = threadIdx.x;
  const uint32_t nslots = (WORK == MCFL_WORK_JOBS) ? A.njobs : 1u;
  unsigned long long* stot = lds;
  unsigned long long* stile = lds + nslots * MCFL_NTOT;
  MatDev* smat = static_cast<MatDev*>(__builtin_assume_aligned(stile + MCFL_TILE_SLOTS(A.tile_n), 16));
  float* szb = reinterpret_cast<float*>(smat + A.n_mats);
  ExitDev* sexit = reinterpret_cast<ExitDev*>(szb + (LAYERED ? 2u * (A.med.n_layers + 2u) : 0u));
  for (uint32_t i = tid; i < A.n_mats; i += WGS) smat[i] = A.mats[i];
  for (uint32_t i = tid; i < nslots; i += WGS) {
    const JobDev& J = A.jobs[i];
    ExitDev e;
    e.x0 = J.det.x0;
    e.y0 = J.det.y0;
    e.size_x = J.det.size_x;
    e.size_y = J.det.size_y;
    e.dx = J.det.dx;
    e.dy = J.det.dy;
    e.nx = (uint32_t)J.det.nx;
    e.npix = J.det.npix;
    e.flags = J.flags;
    e.pad = 0u;
    e.rdx = 1.0 / (double)J.det.dx;
    e.rdy = 1.0 / (double)J.det.dy;
    e.exit_scale = J.exit_scale;
    e.Rd = J.Rd;
    e.Tt = J.Tt;
    e.Rd2 = J.Rd2;
    e.Tt2 = J.Tt2;
    sexit[i] = e;
  }
  if (LAYERED)
    for (uint32_t i = tid; i < 2u * (A.med.n_layers + 2u); i += WGS) szb[i] = A.med.zb[i];
  for (uint32_t i = tid; i < nslots * MCFL_NTOT + A.tile_n; i += WGS) lds[i] = 0ull;
  /* Work: each job has its own counter over its photons [0, count).  A
     workgroup starts on its home job (jobs get workgroups in proportion to
     their share of the photons), whose fhd it privatises in its LDS tile,
     and moves on to the other jobs when its home job is exhausted.  The
     workgroup's constants live in LDS (read on voxel changes and refills, so
     they hold no registers): home job, tile job (~0: no tile), tile origin. */
  __shared__ __align__(16) uint32_t swg[8];
  if (tid == 0) {
    uint32_t home = 0;
    if (WORK == MCFL_WORK_JOBS && A.njobs > 1) {
      const uint64_t gpos = (uint64_t)(((double)blockIdx.x + 0.5) / (double)gridDim.x * (double)A.total);
      home = job_at(A, gpos < A.total ? gpos : A.total - 1ull);
    }
    const JobDev& HJ = A.jobs[home];
    const bool tiled = A.tile_n != 0u && HJ.tile != 0u;
    swg[0] = home;
    swg[1] = tiled ? home : 0xFFFFFFFFu;
    swg[2] = tiled ? HJ.tx0 : 0u;
    swg[3] = tiled ? HJ.ty0 : 0u;
    /* the tile box, beside its origin (read on voxel changes by single-job launches) */
    swg[4] = A.tnx;
    swg[5] = A.tny;
    swg[6] = A.tnz;
    swg[7] = 0u;
  }
  __syncthreads();

  /* the medium: launch-uniform, or the wave's current job's (PERJOB) */
  MediumDev m = A.med;
  const uint32_t lane = __lane_id();
  /* XCD of this workgroup (HW_REG_XCC_ID[3:0]); selects the volume-tally copy */
  const uint32_t xcc = __builtin_amdgcn_s_getreg((3 << 11) | 20) & (MCFL_TALLY_COPIES - 1);

  Ph p = {};
  uint64_t x = 0;
  uint32_t a = 0;
  uint32_t job = 0;
  uint32_t mbase = 0;
  uint32_t flags = 0;
  unsigned long long* fhd = nullptr;
  unsigned long long* axyz = nullptr;
  double escale = 0.0;
  uint64_t accA = 0;
  /* Run-length aggregation of the volume tallies: a photon takes many steps
     inside one voxel (voxel size / mean free path ~ 6 and more), so each lane
     sums its contributions while it stays in a voxel and issues one atomic
     when it leaves the voxel or dies.  Integer sums are unchanged. */
  uint32_t fv = 0xFFFFFFFFu, av = 0xFFFFFFFFu, ft = 0xFFFFFFFFu;
  unsigned long long facc = 0ull, aacc = 0ull;
  /* A flush of the hitting density to global memory waits in (pv, pa) for the
     start of the lane's next step: the step's label gather waits for all of
     the wave's earlier memory operations, so an atomic issued just before
     that wait would stall the wave for its whole latency. */
  uint32_t pv = 0u;
  unsigned long long pa = 0ull;
  /* the lane's last label lookup (voxel, label); cleared for each photon, whose
     medium may differ from the last one's */
  uint32_t lv = 0xFFFFFFFFu, ll = 0u;
  /* the material of the photon's label, kept from the step (or launch) that
     set the label instead of read again at every step: voxel medium, no
     time-of-flight tally, not both volume tallies (the variants with the
     most live registers read it every step; measured 5% slower kept) */
  constexpr bool KEEPMAT = !LAYERED && !TEMPORAL && (!VOL || VM != MCFL_VOL_ALL);
  MatDev mcur{};
  uint32_t iters = 0;
  bool alive = false, retired = false;
  uint32_t lane_item = 0;
  uint64_t lane_left = 0;
  uint32_t vj = 0; /* MCFL_WORK_VLANES: the lane's current voxel */
  /* bounded work per photon: a photon still alive after this many steps is
     stopped and its weight tallied as aborted */
  const uint32_t max_steps = A.max_steps ? A.max_steps : MCFL_MAX_STEPS_PER_PHOTON;

  /* the tile's job and origin as scalars (no LDS read on voxel changes) */
  const uint32_t tile_job = __builtin_amdgcn_readfirstlane(swg[1]);
  /* the tile origin in scalar registers (jobs launches), or read from LDS
     where it is used (single-job launches, whose scalar registers are full
     with the job's own values) */
  const uint32_t tx0 = ONE ? 0u : __builtin_amdgcn_readfirstlane(swg[2]);
  const uint32_t ty0 = ONE ? 0u : __builtin_amdgcn_readfirstlane(swg[3]);
  /* wave-uniform chunk of the current job's photons: [pool_next, pool_next + pool_left) */
  uint32_t cur = ONE ? 0u : __builtin_amdgcn_readfirstlane(swg[0]), pool_left = 0;
  /* PERJOB: the wave's photons share one medium (m stays wave-uniform), that
     of job wmed; a wave moving on to a job in another medium lets its photons
     finish first */
  uint32_t wmed = cur;
  if (PERJOB) m = A.jobs[cur].med;
  uint64_t pool_next = 0;
  bool pool_dead = false;

  if (ONE || WORK == MCFL_WORK_VOXELS || WORK == MCFL_WORK_VLANES) {
    mbase = A.jobs[0].mat_base;
    flags = A.jobs[0].flags;
    fhd = A.jobs[0].fhd ? A.jobs[0].fhd + xcc * A.jobs[0].copy_stride : nullptr;
    axyz = A.jobs[0].axyz ? A.jobs[0].axyz + xcc * A.jobs[0].copy_stride : nullptr;
  }

  for (;;) {
    /* ---- refill: ballot the lanes without a photon, hand out indices ---- */
    bool need = !alive && !retired;
    uint64_t needmask = __ballot(need);
    /* Refilling costs a wave the same whether one lane or many take a photon,
       so idle lanes wait until refill_min of them are idle (or none is busy). */
    if (needmask != 0ull && (uint32_t)__popcll(needmask) < A.refill_min && __ballot(alive) != 0ull && !pool_dead)
      needmask = 0ull;
    while (needmask != 0ull) {
      if (pool_left == 0u) {
        if (pool_dead) break;
        const uint64_t cnt = ONE ? ctab(A.jobs)[0].count : (JOBS ? A.jobs[cur].count : A.total);
        /* guided: the job's remainder (beyond the wave's last chunk) over
           four times the launch's waves, within [64, chunk_max] */
        const uint64_t rem = cnt > pool_next ? cnt - pool_next : 0ull;
        uint64_t chunk = rem / ((uint64_t)A.n_waves * 4ull);
        chunk = chunk < 64ull ? 64ull : (chunk > A.chunk_max ? (uint64_t)A.chunk_max : chunk);
        unsigned long long base = 0ull;
        if (lane == 0) base = atomicAdd(A.counter + cur, (unsigned long long)chunk);
        /* lane 0's result, as a wave-uniform (scalar) value */
        base = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(base >> 32)) << 32) |
               (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)base);
        if (base >= cnt) { /* this job is exhausted: the next one, until the home job comes round again */
          cur = (cur + 1u == nslots) ? 0u : cur + 1u;
          pool_next = 0;
          if (cur == swg[0]) {
            pool_dead = true;
            break;
          }
          continue;
        }
        pool_next = base;
        pool_left = (uint32_t)(base + chunk < cnt ? chunk : cnt - base);
      }
      if (PERJOB && cur != wmed) {
        if (A.jobs[cur].med.labels != m.labels && __ballot(alive) != 0ull) break;
        wmed = cur;
        m = A.jobs[cur].med;
      }
      const uint32_t n = (uint32_t)__popcll(needmask);
      const uint32_t take = pool_left < n ? pool_left : n;
      const uint32_t rank = lane_rank(needmask);
      if (need && rank < take) {
        const uint64_t g = pool_next + rank;
        need = false;
        alive = true;
        accA = 0;
        iters = 0;
        if (JOBS) {
          /* the job's photon k: its stream, then its launch */
          auto start = [&](auto& J) {
            const uint64_t k = g + J.k0;
            a = A.a_tab[k & A.a_mask];
            x = mcfl_photon_state(J.key, k, a);
            if (!ONE) {
              job = cur;
              mbase = J.mat_base;
              flags = J.flags;
              fhd = J.fhd ? J.fhd + xcc * J.copy_stride : nullptr;
              axyz = J.axyz ? J.axyz + xcc * J.copy_stride : nullptr;
            }
            if (J.src_kind == MCFL_SRC_EMISSION) {
              /* importance-sampled emission voxel (Walker alias table), then an
                 isotropic photon from its corner as in CUDAMCFLmain.cu:481-487 */
              const uint32_t col = (uint32_t)(((uint64_t)mcfl_mwc_next(&x, a) * (uint64_t)J.al_n) >> 32);
              const uint32_t u = mcfl_mwc_next(&x, a);
              const uint32_t v =
                  gload_u32(J.al_vox, u < gload_u32(J.al_prob, col) ? col : gload_u32(J.al_alias, col));
              const uint32_t nxy = (uint32_t)m.nx * (uint32_t)m.ny;
              const uint32_t iz = v / nxy, rem = v - iz * nxy, iy = rem / (uint32_t)m.nx, ix = rem - iy * (uint32_t)m.nx;
              launch(m, MCFL_SRC_ISOTROPIC, ((float)ix / m.gs) - m.halfx, ((float)iy / m.gs) - m.halfy,
                     ((float)iz / m.gs), 0.0f, 0u, p, x, a, false, 0u);
            } else {
              launch(m, J.src_kind, J.xi, J.yi, J.zi, J.src_r, J.start_weight, p, x, a, LAYERED, J.start_lab);
            }
          };
          /* one job: its entry through scalar loads, which do not wait for
             the wave's outstanding vector memory operations (tally atomics);
             several: the generic entry of the wave-uniform cur (the scalar
             loads' registers spill in that variant, measured slower) */
          if (ONE)
            start(ctab(A.jobs)[0]);
          else
            start(A.jobs[cur]);
        } else if (WORK == MCFL_WORK_VOXELS) {
          const uint64_t j = g / A.per_job;
          const uint64_t k = g - j * A.per_job;
          const uint32_t v = A.vox[j];
          escale = A.vscale[j];
          a = A.a_tab[k & A.a_mask];
          x = mcfl_photon_state(mcfl_key(A.vseed, v), k, a);
          const uint32_t nxy = (uint32_t)m.nx * (uint32_t)m.ny;
          const uint32_t iz = v / nxy, rem = v - iz * nxy, iy = rem / (uint32_t)m.nx, ix = rem - iy * (uint32_t)m.nx;
          const float xi = ((float)ix / m.gs) - m.halfx;
          const float yi = ((float)iy / m.gs) - m.halfy;
          const float zi = ((float)iz / m.gs);
          launch(m, MCFL_SRC_ISOTROPIC, xi, yi, zi, 0.0f, 0u, p, x, a, false, 0u);
        } else if (WORK == MCFL_WORK_VLANES) { /* the reference's emission loop on its lanes */
          lane_item = (uint32_t)g;
          x = A.xs[lane_item];
          a = A.as[lane_item];
          vj = 0;
          lane_left = A.lane_q + ((uint64_t)lane_item < A.lane_rem ? 1ull : 0ull) - 1ull;
          escale = A.vscale[0];
          uint64_t xl = x; /* LaunchPhoton_Global does not store the advanced state */
          launch_voxel(m, A.vox[0], p, xl, a);
        } else { /* reference lane streams */
          const JobDev& J = A.jobs[0];
          lane_item = (uint32_t)g;
          x = A.xs[lane_item];
          a = A.as[lane_item];
          lane_left = A.lane_q + ((uint64_t)lane_item < A.lane_rem ? 1ull : 0ull) - 1ull;
          mbase = J.mat_base;
          flags = J.flags;
          fhd = J.fhd ? J.fhd + xcc * J.copy_stride : nullptr;
          axyz = J.axyz ? J.axyz + xcc * J.copy_stride : nullptr;
          uint64_t xl = x; /* LaunchPhoton_Global does not store the advanced state */
          launch(m, J.src_kind, J.xi, J.yi, J.zi, J.src_r, J.start_weight, p, xl, a, LAYERED, J.start_lab);
        }
        lv = 0xFFFFFFFFu;
        if (KEEPMAT) mcur = load_mat(smat + mbase + p.lab);
        unsigned long long* st = stot + job * MCFL_NTOT;
        lds_add_u64(&st[MCFL_DT_PHOTONS], 1ull);
        lds_add_u64(&st[MCFL_DT_LAUNCHED_W], p.w);
      }
      pool_next += take;
      pool_left -= take;
      needmask = __ballot(need);
    }
    if (need && pool_dead) retired = true;
    if (__ballot(alive) == 0ull) break;
    if (!alive) continue;

    /* ---- one transport step: MCd3D CUDAMCFLtransport.cu:289-435 (voxel
       medium) or MCd :79-226 (layered slab with one spherical inclusion) ---- */
    const MatDev* mats = smat + mbase;
    unsigned long long* st = stot + job * MCFL_NTOT;
    /* the previous step's flush, issued once the step's material is read (a
       relaunched photon's entry label may still be loading before that) */
    auto flush_pending = [&]() {
      if (VOL && pa != 0ull) {
        atomic_add_u64(fhd + pv, pa);
        pa = 0ull;
      }
    };
    float s;
    uint32_t nb = p.lab;
    MatDev mm; /* material of the absorption and Spin */
    ++iters;
    bool in;
    Cell cl;
    uint32_t vi;
    /* hitting density: the pre-absorption weight at the step's end point,
       summed per lane while the photon stays in one voxel (CUDAMCFLtransport.cu:358-366) */
    auto fhd_visit = [&]() {
      if (VOL && (flags & MCFL_F_FHD) && in) {
        if (vi != fv) {
          /* the sum of the voxel left: into the tile, or pending (pa is 0
             here: the step's start flushed it) */
          const bool tiled = ft != 0xFFFFFFFFu;
          if (tiled && facc != 0ull) lds_add_u64(stile + ft, facc);
          pv = fv;
          pa = tiled ? 0ull : facc;
          fv = vi;
          facc = 0ull;
          ft = 0xFFFFFFFFu;
          if (job == tile_job) { /* inside the LDS tile of the workgroup's home job? */
            typedef uint32_t u2 __attribute__((ext_vector_type(2)));
            typedef uint32_t u4 __attribute__((ext_vector_type(4)));
            const u2 o = ONE ? *reinterpret_cast<const u2*>(swg + 2) : u2{tx0, ty0}; /* [tx0, ty0] */
            const u4 b = ONE ? *reinterpret_cast<const u4*>(swg + 4) : u4{A.tnx, A.tny, A.tnz, 0u}; /* [tnx, tny, tnz, 0] */
            const uint32_t lx = cl.ix - o.x, ly = cl.iy - o.y, lz = cl.iz;
            if (lx < b.x && ly < b.y && lz < b.z) ft = lx + __umul24(b.x, ly + __umul24(b.y, lz));
          }
        }
        facc += p.w;
      }
    };
    if constexpr (LAYERED) {
      const uint32_t L = p.lab;
      const float zlo = szb[2u * L], zhi = szb[2u * L + 1u];
      const MatDev ml = load_mat(mats + L);
      flush_pending();
      if (in_sphere(m, p.x, p.y, p.z)) { /* :83-92 */
        const float mutr = mats[m.n_layers + 2u].mutr;
        s = (mutr != kFltMax) ? -mcfl_log(mcfl_rand_oc(&x, a)) * mutr : 2.0f;
      } else { /* :94-100: deterministic step outside the inclusion */
        s = (ml.mutr != kFltMax) ? ml.mutr : 2.0f;
      }
      int nl = (int)L;
      if (mcfl_fma(s, p.dz, p.z) < zlo) { /* :106-115 */
        --nl;
        s = (zlo - p.z) / p.dz;
      }
      if (mcfl_fma(s, p.dz, p.z) >= zhi) {
        ++nl;
        s = (zhi - p.z) / p.dz;
      }
      p.x = mcfl_fma(p.dx, s, p.x);
      p.y = mcfl_fma(p.dy, s, p.y);
      p.z = mcfl_fma(p.dz, s, p.z);
      if (TEMPORAL) p.tof += mcfl_tof_inc(s * ml.n);
      if (p.z > zhi) p.z = zhi; /* :125-126 */
      if (p.z < zlo) p.z = zlo;
      in = in_grid(m, p.x, p.y, p.z);
      cl = cell_in(m, p.x, p.y, p.z);
      vi = in ? cell_index(m, cl) : 0u;
      nb = (uint32_t)nl;
      if (nb != L) { /* :141-145: no absorption or scattering after a layer change */
        s = 0.0f;
        reflect<!(TEMPORAL && PERJOB)>(mats, p, nb, m.last, x, a);
      }
      mm = load_mat(mats + (in_sphere(m, p.x, p.y, p.z) ? m.n_layers + 2u : p.lab)); /* :200-224 */
    } else {
      const MatDev mc = KEEPMAT ? mcur : load_mat(mats + p.lab);
      flush_pending();
      if (mc.mutr != kFltMax) {
        s = -mcfl_log(mcfl_rand_oc(&x, a)) * mc.mutr;
        if (mcfl_fma(s, p.dz, p.z) < 0.0f) {
          nb = 0;
          s = -p.z / p.dz;
        }
        if (mcfl_fma(s, p.dz, p.z) > m.esp) {
          nb = m.last;
          s = (m.esp - p.z) / p.dz;
        }
        p.x = mcfl_fma(p.dx, s, p.x);
        p.y = mcfl_fma(p.dy, s, p.y);
        p.z = mcfl_fma(p.dz, s, p.z);
        if (TEMPORAL) p.tof += mcfl_tof_inc(s * mc.n);
      } else {
        s = 100.0f;
        nb = glass_march<TEMPORAL>(m, p, p.lab, s);
      }
      in = in_grid(m, p.x, p.y, p.z);
      cl = cell_in(m, p.x, p.y, p.z);
      vi = in ? cell_index(m, cl) : 0u;
      /* the end point's label (:319-333): gathered unless it is the voxel of
         the lane's last lookup (the temporal mixed-media variant, short of
         registers, keeps no such cache); the hitting-density update, which
         does not depend on it, runs while the gather is in flight */
      const bool look = nb == p.lab && mc.mutr != kFltMax;
      const bool gather = look && in && (TEMPORAL && PERJOB ? true : vi != lv);
      uint32_t got = 0u;
      if (gather) got = (uint32_t)m.labels[vi];
      fhd_visit();
      if (look) {
        if (!in) {
          nb = 1u;
        } else if (gather) {
          nb = got;
          if (!(TEMPORAL && PERJOB)) {
            lv = vi;
            ll = got;
          }
        } else {
          nb = ll;
        }
      }
      mm = mc; /* material after a possible interface crossing */
      if (nb != p.lab) {
        reflect<!(TEMPORAL && PERJOB)>(mats, p, nb, m.last, x, a);
        mm = load_mat(mats + p.lab);
        if (KEEPMAT) mcur = mm;
      }
    }

    if constexpr (LAYERED) fhd_visit();

    if (p.lab == 0u || p.lab == m.last) {
      /* a job's image scale from its exit constants in LDS; a voxel emission job's per voxel */
      exit_tally<TEMPORAL>(A.jobs[job], sexit[job], p, p.lab == 0u, (flags & MCFL_F_SCALED) != 0u,
                           JOBS ? sexit[job].exit_scale : escale, st);
      p.w = 0u;
      s = 0.0f;
    }
    if (s > 0.0f) {
      /* voxel medium: w = ceil(mua mutr weight), killed at w = 0 (:425-431);
         layered: rounded to nearest, no kill (:205-223) */
      const uint32_t w = LAYERED ? mcfl_f2u_rn(mm.mua_mutr * (float)p.w) : mcfl_f2u_ru(mm.mua_mutr * (float)p.w);
      if (LAYERED || w > 0u) {
        p.w -= w;
        accA += w;
        if (AX && (flags & MCFL_F_AXYZ) && in) {
          if (vi != av) {
            if (aacc != 0ull) atomic_add_u64(axyz + av, aacc);
            av = vi;
            aacc = 0ull;
          }
          aacc += w;
        }
      } else {
        lds_add_u64(&st[MCFL_DT_ZERO_KILL], p.w);
        p.w = 0u;
      }
      spin(mm, p, x, a);
    }
    /* PhotonSurvive, CUDAMCFLtransport.cu:850-864 */
    bool lives;
    if (p.w > kWeightI) {
      lives = true;
    } else if (p.w == 0u) {
      lives = false;
    } else if (mcfl_rand_co(&x, a) < kChance) {
      const uint32_t nw = mcfl_f2u_rn((float)p.w / kChance);
      lds_add_u64(&st[MCFL_DT_ROUL_GAIN], (unsigned long long)(nw - p.w));
      p.w = nw;
      lives = true;
    } else {
      lds_add_u64(&st[MCFL_DT_ROUL_LOSS], p.w);
      p.w = 0u;
      lives = false;
    }
    if (lives && iters >= max_steps) {
      lds_add_u64(&st[MCFL_DT_ABORTED], p.w);
      p.w = 0u;
      lives = false;
    }
    if (!lives) {
      if (VOL && facc != 0ull) {
        if (ft != 0xFFFFFFFFu)
          lds_add_u64(stile + ft, facc);
        else
          atomic_add_u64(fhd + fv, facc);
      }
      ft = 0xFFFFFFFFu;
      if (VOL && pa != 0ull) {
        atomic_add_u64(fhd + pv, pa);
        pa = 0ull;
      }
      if (AX && aacc != 0ull) atomic_add_u64(axyz + av, aacc);
      fv = av = 0xFFFFFFFFu;
      facc = aacc = 0ull;
      lds_add_u64(&st[MCFL_DT_A], accA);
      lds_add_u64(&st[MCFL_DT_STEPS], iters);
      alive = false;
      if (WORK == MCFL_WORK_LANES) {
        if (lane_left > 0ull) { /* next photon of the same lane, CUDAMCFLtransport.cu:437-440 */
          cjob& J = ctab(A.jobs)[0];
          --lane_left;
          accA = 0;
          iters = 0;
          launch(m, J.src_kind, J.xi, J.yi, J.zi, J.src_r, J.start_weight, p, x, a, LAYERED, J.start_lab);
          if (KEEPMAT) mcur = load_mat(mats + p.lab);
          lds_add_u64(&st[MCFL_DT_PHOTONS], 1ull);
          lds_add_u64(&st[MCFL_DT_LAUNCHED_W], p.w);
          alive = true;
        } else {
          A.xs[lane_item] = x; /* CUDAMCFLmem.cu:42: advanced states go back */
        }
      } else if (WORK == MCFL_WORK_VLANES) {
        /* next photon of this voxel's simulation, or the next voxel's
           simulation (a new LaunchPhoton_Global: the first photon replays) */
        bool replay = false, more = true;
        if (lane_left == 0ull) {
          if (++vj < A.n_vox) {
            lane_left = A.lane_q + ((uint64_t)lane_item < A.lane_rem ? 1ull : 0ull);
            escale = A.vscale[vj];
            replay = true;
          } else {
            more = false;
          }
        }
        if (more) {
          --lane_left;
          accA = 0;
          iters = 0;
          if (replay) {
            uint64_t xl = x;
            launch_voxel(m, A.vox[vj], p, xl, a);
          } else {
            launch_voxel(m, A.vox[vj], p, x, a);
          }
          if (KEEPMAT) mcur = load_mat(mats + p.lab);
          lds_add_u64(&st[MCFL_DT_PHOTONS], 1ull);
          lds_add_u64(&st[MCFL_DT_LAUNCHED_W], p.w);
          alive = true;
        } else {
          A.xs[lane_item] = x;
        }
      }
    }
  }

  __syncthreads();
  for (uint32_t i = tid; i < nslots * MCFL_NTOT; i += WGS) {
    const unsigned long long v = stot[i];
    if (v != 0ull) atomic_add_u64(A.jobs[i / MCFL_NTOT].tot + (i % MCFL_NTOT), v);
  }
  if (swg[1] != 0xFFFFFFFFu) {
    const JobDev& HJ = A.jobs[swg[1]];
    unsigned long long* dst = HJ.fhd + xcc * HJ.copy_stride;
    const MediumDev& tm = PERJOB ? HJ.med : A.med; /* the home job's medium, not the lane's last one */
    const uint32_t tx0 = swg[2], ty0 = swg[3];
    for (uint32_t i = tid; i < A.tile_n; i += WGS) {
      const unsigned long long v = stile[i];
      if (v == 0ull) continue; /* entries outside the medium stay 0 */
      const uint32_t lx = i % A.tnx, r2 = i / A.tnx, ly = r2 % A.tny, lz = r2 / A.tny;
      const uint32_t v_idx = tx0 + lx + (uint32_t)tm.nx * (ty0 + ly + (uint32_t)tm.ny * lz);
      atomic_add_u64(dst + v_idx, v);
    }
  }
}

/* Zero-fill used inside captured graphs (no memset nodes in graphs). */
__global__ void mcfl_fill_u64_kernel(unsigned long long* p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

/* Sums copies 1..ncopies-1 of a per-XCD volume tally into copy 0. */
/* Sum of the ncopies copies of element i, saturated at 2^64-1 (a voxel whose
   total weight exceeds the u64 range keeps the largest value instead of
   wrapping; the reference stops adding at LLONG_MAX, CUDAMCFLtransport.cu:358-366). */
__device__ __forceinline__ unsigned long long sat_sum_copies(const unsigned long long* p, uint64_t n, uint32_t ncopies,
                                                             uint64_t i) {
  unsigned long long s = p[i];
  bool over = false;
  for (uint32_t c = 1; c < ncopies; ++c) {
    const unsigned long long v = p[(uint64_t)c * n + i];
    s += v;
    over |= s < v;
  }
  return over ? ~0ull : s;
}

__global__ void mcfl_reduce_copies_kernel(unsigned long long* p, uint64_t n, uint32_t ncopies) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = sat_sum_copies(p, n, ncopies, i);
}

__global__ void mcfl_accumulate_copies_kernel(unsigned long long* p, uint64_t n, uint32_t ncopies,
                                              unsigned long long* acc) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned long long s = sat_sum_copies(p, n, ncopies, i), a = acc[i] + s;
    acc[i] = a < s ? ~0ull : a;
    for (uint32_t c = 0; c < ncopies; ++c) p[(uint64_t)c * n + i] = 0ull;
  }
}

/* Scratch of the allocation-free excitation: the intermediate device buffers
   of one run (work counter, totals, job and material tables, labels, images,
   per-XCD volume-tally copies), laid out by the MCFL_NA_OFF_* offsets of
   mcfl_device.h.  Each GPU the library runs on holds its own copy until the
   process ends; every allocation-free call on a GPU uses that GPU's copy, so
   such calls on one GPU must not overlap. */
static __device__ __align__(256) unsigned long long scratch[MCFL_NA_BYTES / 8];

/* ---- self-checks of the per-step arithmetic (test support) ----
   Each compares a form the transport kernel uses with an independent one on
   the GPU itself, over a whole range of inputs, counting mismatches. */

/* float -> uint32 toward zero from the bits alone (no conversion instruction) */
__device__ __forceinline__ uint32_t f2u_rz_bits(uint32_t b) {
  if (b >> 31) return 0u; /* negative or -0 (and NaN with the sign bit) */
  const uint32_t e = b >> 23, m = (b & 0x007FFFFFu) | 0x00800000u;
  if (e == 255u) return (b & 0x007FFFFFu) ? 0u : 0xFFFFFFFFu; /* NaN, +inf */
  if (e >= 159u) return 0xFFFFFFFFu;                           /* >= 2^32 */
  if (e < 127u) return 0u;                                     /* < 1 */
  return e >= 150u ? m << (e - 150u) : m >> (150u - e);
}

__global__ void mcfl_selfcheck_kernel(int which, uint32_t p, uint64_t lo, uint64_t hi, unsigned long long* bad,
                                      unsigned long long* first) {
  for (uint64_t i = lo + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < hi;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t u = (uint32_t)i;
    bool ok = true;
    if (which == MCFL_SC_U2F_RZ) { /* instruction under toward-zero MODE vs the exact conversion of u's top 24 bits */
      ok = mcfl_fbits(mcfl_u2f_rz(u)) == mcfl_fbits((float)(u & (0xFFFFFF00u >> __builtin_clz(u | 1u))));
    } else if (which == MCFL_SC_F2U_RZ) { /* saturating conversion instruction vs the bits */
      ok = mcfl_f2u_rz(mcfl_bitsf(u)) == f2u_rz_bits(u);
    } else if (which == MCFL_SC_DIV) { /* Spin's quotient: numerator bits p, denominator bits u */
      const float a = mcfl_bitsf(p), b = mcfl_bitsf(u);
      ok = mcfl_fbits(div_unscaled(a, b)) == mcfl_fbits(a / b);
    } else if (which == MCFL_SC_PIXEL) { /* exit-tally quotient: pixel size bits p, offset bits u */
      const float d = mcfl_bitsf(p), q = mcfl_bitsf(u);
      const double rd = 1.0 / (double)d;
      ok = mcfl_fbits((float)((double)q * rd)) == mcfl_fbits(q / d);
    }
    if (!ok) {
      atomicAdd(bad, 1ull);
      atomicMin(first, (unsigned long long)i);
    }
  }
}

/* The shared FP32 functions on the GPU, for comparison with the host build. */
__global__ void mcfl_math_eval_kernel(int fn, const uint32_t* in, uint32_t* out, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t u = in[i];
    if (fn == MCFL_ME_LOG) {
      out[i] = mcfl_fbits(mcfl_log(mcfl_bitsf(u)));
    } else if (fn == MCFL_ME_RSQRT) {
      out[i] = mcfl_fbits(mcfl_rsqrt(mcfl_bitsf(u)));
    } else if (fn == MCFL_ME_RENORM) {
      out[i] = mcfl_fbits(mcfl_renorm(mcfl_bitsf(u)));
    } else if (fn == MCFL_ME_SINCOS) { /* of 2 pi times the [0,1) draw whose 32-bit value is u */
      float s, c;
      mcfl_sincos_quarter(mcfl_u2f_rz(u) * 0x1p-30f, &s, &c);
      out[2 * i] = mcfl_fbits(s);
      out[2 * i + 1] = mcfl_fbits(c);
    } else if (fn == MCFL_ME_DRAW_OC) {
      out[i] = mcfl_fbits(mcfl_fma(mcfl_u2f_rz(u), -0x1p-32f, 1.0f));
    }
  }
}

/* ------------------------------------------------------------------ host */

namespace {
template <int W, bool T, bool P, bool L, int G = MCFL_WG, int V = MCFL_VOL_ALL>
const void* kfn() {
  return (const void*)mcfl_transport_kernel<W, T, P, L, G, V>;
}
/* job launches of workgroup size G: mixed media (P) or not, one job (ONE)
   or several, with the volume tallies V */
template <int G, int V>
const void* job_kernel(int temporal, int perjob, bool one) {
  if constexpr (G == MCFL_WG) /* single-job launches run at the default size only */
    if (one && !perjob)
      return temporal ? kfn<MCFL_WORK_ONE, true, false, false, G, V>() : kfn<MCFL_WORK_ONE, false, false, false, G, V>();
  if (perjob) return temporal ? kfn<MCFL_WORK_JOBS, true, true, false, G, V>() : kfn<MCFL_WORK_JOBS, false, true, false, G, V>();
  return temporal ? kfn<MCFL_WORK_JOBS, true, false, false, G, V>() : kfn<MCFL_WORK_JOBS, false, false, false, G, V>();
}
template <int G>
const void* job_kernel(int temporal, int perjob, bool one, uint32_t vol) {
  if (vol == MCFL_VOL_NONE) return job_kernel<G, MCFL_VOL_NONE>(temporal, perjob, one);
  if (vol == MCFL_VOL_FHD) return job_kernel<G, MCFL_VOL_FHD>(temporal, perjob, one);
  return job_kernel<G, MCFL_VOL_ALL>(temporal, perjob, one);
}
/* perjob selects mixed-media launches (jobs only); layered the bulk-method-1
   medium (jobs and reference lanes, one medium per launch); wg the workgroup
   size (MCFL_WG, or MCFL_WG_ALT1/2 for non-layered job launches); one a job
   launch of a single job; vol the volume tallies of a job launch (MCFL_VOL_*) */
const void* kernel_of(int work, int temporal, int perjob, int layered, uint32_t wg, bool one = false,
                      uint32_t vol = MCFL_VOL_ALL) {
  if (layered) {
    if (wg != MCFL_WG) return nullptr;
    if (work == MCFL_WORK_JOBS)
      return temporal ? kfn<MCFL_WORK_JOBS, true, false, true>() : kfn<MCFL_WORK_JOBS, false, false, true>();
    if (work == MCFL_WORK_LANES)
      return temporal ? kfn<MCFL_WORK_LANES, true, false, true>() : kfn<MCFL_WORK_LANES, false, false, true>();
    return nullptr;
  }
  if (work == MCFL_WORK_JOBS) {
    if (wg == MCFL_WG) return job_kernel<MCFL_WG>(temporal, perjob, one, vol);
    if (wg == MCFL_WG_ALT1) return job_kernel<MCFL_WG_ALT1>(temporal, perjob, false, vol);
    if (wg == MCFL_WG_ALT2) return job_kernel<MCFL_WG_ALT2>(temporal, perjob, false, vol);
    return nullptr;
  }
  if (wg != MCFL_WG) return nullptr;
  if (work == MCFL_WORK_VOXELS)
    return temporal ? kfn<MCFL_WORK_VOXELS, true, false, false>() : kfn<MCFL_WORK_VOXELS, false, false, false>();
  if (work == MCFL_WORK_VLANES) return temporal ? nullptr : kfn<MCFL_WORK_VLANES, false, false, false>();
  return temporal ? kfn<MCFL_WORK_LANES, true, false, false>() : kfn<MCFL_WORK_LANES, false, false, false>();
}
}  // namespace

extern "C" int mcfl_transport_occupancy(int work, int temporal, int perjob, int layered, uint32_t wg, uint32_t lds_bytes,
                                        int* blocks_per_cu) {
  const void* k = kernel_of(work, temporal, perjob, layered, wg ? wg : MCFL_WG);
  if (!k) return -1;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k, (int)(wg ? wg : MCFL_WG), lds_bytes);
  return e == hipSuccess ? 0 : (int)e;
}

extern "C" int mcfl_transport_launch(const LaunchArgs* A, int temporal, int perjob, int nblocks, hipStream_t stream) {
  if (A->work < 0 || A->work > 3 || nblocks <= 0) return -1;
  /* every launch reads the job table, the work counters and the multiplier
     table; the other pointers as its work kind needs them */
  if (!A->jobs || !A->counter || !A->a_tab || (A->n_mats && !A->mats) || A->njobs == 0 || A->njobs > MCFL_MAX_JOBS)
    return -1;
  if ((A->work == MCFL_WORK_LANES || A->work == MCFL_WORK_VLANES) && (!A->xs || !A->as)) return -1;
  if (A->work == MCFL_WORK_VLANES && (!A->vox || !A->vscale || A->n_vox == 0 || A->lane_q == 0)) return -1;
  if (A->work == MCFL_WORK_VOXELS && (!A->vox || !A->vscale || A->per_job == 0)) return -1;
  if (!(A->work == MCFL_WORK_JOBS && perjob) && !A->med.labels && A->med.n_layers == 0) return -1;
  if (A->med.n_layers > 0 && !A->med.zb) return -1;
  const int layered = A->med.n_layers > 0u ? 1 : 0;
  const uint32_t wg = mcfl_wg_of(A);
  const void* k = kernel_of(A->work, temporal, A->work == MCFL_WORK_JOBS ? perjob : 0, layered, wg, A->njobs == 1u,
                            A->vol);
  if (!k || (layered && perjob)) return -1;
  if (A->n_waves != (uint32_t)nblocks * (wg / 64u)) return -1;
  void* args[] = {const_cast<LaunchArgs*>(A)};
  const uint32_t nslots = A->work == MCFL_WORK_JOBS ? A->njobs : 1u;
  hipError_t e = hipLaunchKernel(k, dim3((unsigned)nblocks), dim3(wg), args,
                                 mcfl_lds_bytes(nslots, A->tile_n, A->n_mats, layered ? A->med.n_layers + 2u : 0u),
                                 stream);
  return e == hipSuccess ? 0 : (int)e;
}

extern "C" int mcfl_reduce_copies(unsigned long long* p, uint64_t n, uint32_t ncopies, hipStream_t stream) {
  if (n == 0 || ncopies <= 1) return 0;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  mcfl_reduce_copies_kernel<<<dim3((unsigned)blocks), dim3(256), 0, stream>>>(p, n, ncopies);
  return (int)hipGetLastError();
}

extern "C" int mcfl_accumulate_copies(unsigned long long* p, uint64_t n, uint32_t ncopies, unsigned long long* acc,
                                      hipStream_t stream) {
  if (n == 0) return 0;
  if (!p || !acc || ncopies == 0) return -1;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  mcfl_accumulate_copies_kernel<<<dim3((unsigned)blocks), dim3(256), 0, stream>>>(p, n, ncopies, acc);
  return (int)hipGetLastError();
}

extern "C" int mcfl_selfcheck(int which, uint32_t p, uint64_t lo, uint64_t hi, uint64_t* mismatches,
                              uint64_t* first) {
  if (which < 0 || which > MCFL_SC_PIXEL || lo > hi || hi > (1ull << 32) || !mismatches || !first) return -1;
  unsigned long long* d = nullptr;
  if (hipMalloc((void**)&d, 16) != hipSuccess) return -1;
  const unsigned long long init[2] = {0ull, ~0ull};
  hipError_t e = hipMemcpy(d, init, 16, hipMemcpyHostToDevice);
  if (e == hipSuccess && hi > lo) {
    mcfl_selfcheck_kernel<<<dim3(4096), dim3(256)>>>(which, p, lo, hi, d, d + 1);
    e = hipGetLastError();
  }
  unsigned long long res[2] = {0ull, 0ull};
  if (e == hipSuccess) e = hipMemcpy(res, d, 16, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return (int)e;
  *mismatches = res[0];
  *first = res[1];
  return 0;
}

extern "C" int mcfl_math_eval(int fn, const uint32_t* in, uint32_t* out, uint64_t n) {
  if (fn < 0 || fn > MCFL_ME_DRAW_OC || !in || !out) return -1;
  if (n == 0) return 0;
  const uint64_t nout = fn == MCFL_ME_SINCOS ? 2 * n : n;
  uint32_t *din = nullptr, *dout = nullptr;
  hipError_t e = hipMalloc((void**)&din, 4 * n);
  if (e == hipSuccess) e = hipMalloc((void**)&dout, 4 * nout);
  if (e == hipSuccess) e = hipMemcpy(din, in, 4 * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    mcfl_math_eval_kernel<<<dim3(4096), dim3(256)>>>(fn, din, dout, n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(out, dout, 4 * nout, hipMemcpyDeviceToHost);
  (void)hipFree(din);
  (void)hipFree(dout);
  return e == hipSuccess ? 0 : (int)e;
}

extern "C" int mcfl_fill_u64(unsigned long long* p, uint64_t n, hipStream_t stream) {
  if (n == 0) return 0;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  mcfl_fill_u64_kernel<<<dim3((unsigned)blocks), dim3(256), 0, stream>>>(p, n);
  return (int)hipGetLastError();
}

/* Device address of this GPU's scratch array, looked up once per device
   (the device must be current) and kept: the array's name is a host-side
   stand-in and is never passed to a kernel or a copy. */
extern "C" int mcfl_scratch_address(int device, void** out, uint64_t* bytes) {
  static std::mutex mu;
  static void* addr[256] = {};
  if (!out || device < 0 || device >= 256) return -1;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return e != hipSuccess ? (int)e : -1;
    addr[device] = p;
  }
  *out = addr[device];
  if (bytes) *bytes = MCFL_NA_BYTES;
  return 0;
}
