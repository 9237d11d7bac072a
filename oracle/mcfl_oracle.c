/*
 * mcfl_oracle.c -- CPU restatement of the reference's method-2 transport.
 * TEST INFRASTRUCTURE ONLY (see mcfl_oracle.h).  Every function cites the
 * reference lines it restates.  Build with -ffp-contract=off (oracle/Makefile).
 */
#include "mcfl_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../cudamcfl_amd/csrc/mcfl_mwc.h"

static uint32_t g_max_steps = ORC_MAX_STEPS_PER_PHOTON;
void orc_set_max_steps(uint32_t n) { g_max_steps = n ? n : ORC_MAX_STEPS_PER_PHOTON; }

#define WEIGHTI 429497u     /* CUDAMCFL.h:33 */
#define CHANCE 0.1f         /* CUDAMCFL.h:34 */
#define C_CMFS 2.99792E-5   /* CUDAMCFL.h:29, cm per fs */
#define PI_F 3.141592654f   /* CUDAMCFL.h:22 */

/* Per-material constants.  mua_mutr is the product the reference forms in
 * CUDAMCFLtransport.cu:427 (mua*mutr, left to right); the g-terms are the
 * pieces of the Henyey-Greenstein inversion at :596-597. */
typedef struct {
  float mutr, mua_mutr, n, g, onemg, onemg2, onepg2, twog, inv2g;
} mat_t;

typedef struct {
  const orc_sim* s;
  mat_t mt[256];
  mat_t inc; /* method 1: the spherical inclusion */
  uint32_t last;
  size_t nvox, npix;
} ctx_t;

static void mat_init(mat_t* m, const orc_bulk* b) {
  float g = b->g;
  float g2 = g * g;
  m->mutr = b->mutr;
  m->mua_mutr = b->mua * b->mutr;
  m->n = b->n;
  m->g = g;
  m->onemg = 1.0f - g;
  m->onemg2 = 1.0f - g2;
  m->onepg2 = 1.0f + g2;
  m->twog = 2.0f * g;
  m->inv2g = (g != 0.0f) ? 1.0f / (2.0f * g) : 0.0f;
}

static void ctx_init(ctx_t* c, const orc_sim* s) {
  memset(c, 0, sizeof *c);
  c->s = s;
  c->last = s->n_bulks + 1u;
  for (uint32_t i = 0; i < s->n_bulks + 2u && i < 256u; ++i) mat_init(&c->mt[i], &s->bulks[i]);
  if (s->layered) mat_init(&c->inc, &s->inc);
  c->nvox = (size_t)s->nx * (size_t)s->ny * (size_t)s->nz;
  c->npix = (size_t)s->det_nx * (size_t)(s->det_ny + 1);
}

/* Voxel index of a point known to be inside the lateral window and above the
 * bottom (CUDAMCFLtransport.cu:323-325).  Components are clamped to the grid:
 * float rounding can give ix == nx for x just below the half width, which the
 * reference would read one row over (or past the end of the array). */
static uint32_t vindex(const orc_sim* s, float x, float y, float z) {
  uint32_t ix = mcfl_f2u_rz((x + s->halfx) * s->gs);
  uint32_t iy = mcfl_f2u_rz((y + s->halfy) * s->gs);
  uint32_t iz = mcfl_f2u_rz(z * s->gs);
  if (ix > (uint32_t)s->nx - 1u) ix = (uint32_t)s->nx - 1u;
  if (iy > (uint32_t)s->ny - 1u) iy = (uint32_t)s->ny - 1u;
  if (iz > (uint32_t)s->nz - 1u) iz = (uint32_t)s->nz - 1u;
  return ix + (uint32_t)s->nx * (iy + (uint32_t)s->ny * iz);
}

static int inside(const orc_sim* s, float x, float y, float z) {
  return mcfl_fabs(x) < s->halfx && mcfl_fabs(y) < s->halfy && z < s->esp;
}

/* ---------------------------------------------------------------- RNG init */

/* CUDAMCFLrng.cu:31-90.  The carry is computed in double exactly as :78. */
int orc_init_rng(uint64_t* x, uint32_t* a, uint32_t n_rng, const char* primes_file, uint64_t xinit) {
  FILE* fp = fopen(primes_file, "r");
  if (!fp) return 1;
  unsigned int begin = 0u, fora, t1, t2;
  if (fscanf(fp, "%u %u %u", &begin, &t1, &t2) != 3) { fclose(fp); return 1; }
  if ((xinit == 0ull) | (((unsigned int)(xinit >> 32)) >= (begin - 1)) | (((unsigned int)xinit) >= 0xfffffffful)) {
    fclose(fp);
    return 1;
  }
  for (uint32_t i = 0; i < n_rng; ++i) {
    if (fscanf(fp, "%u %u %u", &fora, &t1, &t2) != 3) { fclose(fp); return 1; }
    a[i] = fora;
    x[i] = 0;
    while ((x[i] == 0) | (((unsigned int)(x[i] >> 32)) >= (fora - 1)) | (((unsigned int)x[i]) >= 0xfffffffful)) {
      xinit = (xinit & 0xffffffffull) * (begin) + (xinit >> 32);
      x[i] = (unsigned int)floor((((double)((unsigned int)xinit)) / (double)0x100000000) * fora);
      x[i] = x[i] << 32;
      xinit = (xinit & 0xffffffffull) * (begin) + (xinit >> 32);
      x[i] += (unsigned int)xinit;
    }
  }
  fclose(fp);
  return 0;
}

/* ------------------------------------------------------------ LaunchPhoton */

static uint32_t launch_label(const ctx_t* c, float x, float y, float z) {
  const orc_sim* s = c->s;
  /* CUDAMCFLtransport.cu:490-503 / :529-543 (method 2) */
  if (z < s->esp) {
    if (mcfl_fabs(x) < s->halfx && mcfl_fabs(y) < s->halfy) return s->labels[vindex(s, x, y, z)];
    return 1u;
  }
  return c->last;
}

/* CUDAMCFLtransport.cu:460-562.  Own transcendental functions replace
 * __logf/__sincosf/__cosf (mcfl_fmath.h). */
static void orc_launch(const ctx_t* c, int kind, float xi, float yi, float zi, orc_photon* p, uint64_t* x,
                       uint32_t a) {
  const orc_sim* s = c->s;
  if (kind == ORC_SRC_ISOTROPIC) {
    float step = 0.999999f / s->gs; /* :470-472 */
    p->x = xi + step * mcfl_rand_oc(x, a);
    p->y = yi + step * mcfl_rand_oc(x, a);
    p->z = zi + step * mcfl_rand_oc(x, a);
    float costheta = (float)(1.0 - 2.0 * (double)mcfl_rand_oc(x, a));          /* :474 */
    float sintheta = (float)sqrt(1.0 - (double)(costheta * costheta));         /* :475 */
    float u = mcfl_rand_oc(x, a);                                              /* :476 */
    float psi = (float)(2.0 * (double)PI_F * (double)u);
    float sn, cospsi;
    mcfl_sincos_turn(u, &sn, &cospsi); /* :477 cos(psi) */
    float sinpsi = mcfl_sqrt((float)(1.0 - (double)(cospsi * cospsi)));       /* :480-483 */
    if (!(psi < PI_F)) sinpsi = -sinpsi;
    p->dx = sintheta * cospsi;
    p->dy = sintheta * sinpsi;
    p->dz = costheta;
    p->weight = 0xFFFFFFFFu; /* :489 */
  } else {
    float rad, sn, cs;
    float u1 = mcfl_rand_oc(x, a);
    float u2 = mcfl_rand_oc(x, a);
    if (kind == ORC_SRC_PENCIL) {
      rad = 0.03f * mcfl_sqrt(-mcfl_log(u1)); /* :510-512 Gaussian spot */
    } else {
      rad = s->src_r * mcfl_sqrt(u1); /* uniform disk (extension) */
    }
    mcfl_sincos_turn(u2, &sn, &cs); /* :513-517 */
    p->x = xi + rad * cs;
    p->y = yi + rad * sn;
    p->z = zi;
    p->dx = 0.0f;
    p->dy = 0.0f;
    p->dz = 1.0f;
    p->weight = s->start_weight; /* :527 */
  }
  /* method 2: label at the entry point; method 1: the layer holding zi (:547-560) */
  p->bulkpos = s->layered ? s->start_layer : launch_label(c, p->x, p->y, p->z);
  p->tof = 0;
}

/* -------------------------------------------------------------------- Spin */

/* CUDAMCFLtransport.cu:585-625.  __fdividef(.,2g) becomes a product with
 * 1/(2g); the draw order, the g == 0 extra draw, the sin/cos swap of :603 and
 * the two incidence branches follow the reference. */
static void orc_spin(const mat_t* m, orc_photon* p, uint64_t* x, uint32_t a) {
  float xi = mcfl_rand_co(x, a);
  float temp = m->onemg2 / mcfl_fma(m->twog, xi, m->onemg);
  float cost = mcfl_fma(-temp, temp, m->onepg2) * m->inv2g;
  if (m->g == 0.0f) cost = 2.0f * mcfl_rand_co(x, a) - 1.0f;
  float sin2t = mcfl_clamp0(mcfl_fma(-cost, cost, 1.0f)); /* sint^2 (:600) */
  float s_psi, c_psi;
  mcfl_rand_sincos(x, a, &s_psi, &c_psi);
  float cosp = s_psi; /* :603 assigns the sine to cosp */
  float sinp = c_psi;
  float dx = p->dx, dy = p->dy, dz = p->dz;
  float temp2 = mcfl_clamp0(mcfl_fma(-dz, dz, 1.0f)); /* temp^2, temp = sqrt(1 - dz^2) (:605) */
  float ndx, ndy, ndz;
  if (temp2 == 0.0f) { /* :607-612 normal incidence */
    float sint = mcfl_sqrt_rs(sin2t);
    ndx = sint * cosp;
    ndy = sint * sinp;
    ndz = mcfl_copysign(cost, dz * cost);
  } else { /* :613-618, with sint/temp and sint*temp from one reciprocal root of sint^2 temp^2,
              grouped as dx (cost + st cosp dz) - dy st sinp, dy (cost + st cosp dz) + dx st sinp */
    float pq = sin2t * temp2;
    float r = pq > 0.0f ? mcfl_rsqrt(pq) : 0.0f;
    float st = sin2t * r;
    float stq = pq * r;
    float c = cost + (st * cosp) * dz, w = st * sinp;
    ndx = mcfl_fma(dx, c, -(dy * w));
    ndy = mcfl_fma(dy, c, dx * w);
    ndz = mcfl_fma(-stq, cosp, dz * cost);
  }
  /* :621-624 renormalisation with a reciprocal square root (first order near 1) */
  float r = mcfl_renorm(mcfl_fma(ndz, ndz, mcfl_fma(ndy, ndy, ndx * ndx)));
  p->dx = ndx * r;
  p->dy = ndy * r;
  p->dz = ndz * r;
}

/* ----------------------------------------------------------------- Reflect */

/* Unpolarised Fresnel reflectance in the reference's closed form
 * (CUDAMCFLtransport.cu:676-679), for 0 < cos_i < 1 without TIR. */
static float fresnel_r(float n1, float n2, float ci) {
  float ci2 = ci * ci;
  float e = (n1 * n1) / (n2 * n2) * (1.0f - ci2);
  float r = 2.0f * mcfl_sqrt(mcfl_fmax(0.0f, (1.0f - ci2) * (1.0f - e) * e * ci * ci));
  e = e + ci2 * (1.0f - 2.0f * e);
  return e * ((1.0f - e - r) / ((1.0f - e + r) * (e + r)));
}

/* CUDAMCFLtransport.cu:628-700 (bulk method 2). Returns 1 if reflected. */
static int orc_reflect(const ctx_t* c, orc_photon* p, uint32_t newlb, uint64_t* x, uint32_t a) {
  float n1 = c->mt[p->bulkpos].n;
  float n2 = c->mt[newlb].n;
  float ci = mcfl_fabs(p->dz);
  if (n1 == n2) { /* :644-649 */
    p->bulkpos = newlb;
    return 0;
  }
  if (n2 * n2 < n1 * n1 * (1.0f - ci * ci)) { /* :651-655 total internal reflection */
    p->dz = -p->dz;
    return 1;
  }
  if (ci == 1.0f) { /* :657-672 normal incidence */
    float r = (n1 - n2) / (n1 + n2);
    if (mcfl_rand_co(x, a) <= r * r) {
      p->dz = -p->dz;
      return 1;
    }
    p->bulkpos = newlb;
    return 0;
  }
  float r = fresnel_r(n1, n2, ci);
  if (mcfl_rand_co(x, a) <= r) { /* :681-686 */
    p->dz = -p->dz;
    return 1;
  }
  r = n1 / n2; /* :689-697 */
  float e = r * r * (1.0f - ci * ci);
  p->dx *= r;
  p->dy *= r;
  p->dz = mcfl_copysign(mcfl_sqrt(mcfl_fmax(0.0f, 1.0f - e)), p->dz);
  p->bulkpos = newlb;
  return 0;
}

/* ---------------------------------------------------- MoveToFirstBoundary */

/* Roots t of |p - t d - c|^2 = r2 (the back step of :792-812, solved in the
 * compact form b -/+ sqrt(b^2 - |d|^2(|p-c|^2 - r2)) over |d|^2 with
 * b = d.(p-c); algebraically identical to the reference's expanded form).
 * The discriminant is clamped at 0. */
static void backstep(float px, float py, float pz, float dx, float dy, float dz, float cx, float cy, float cz,
                     float r2, float* t1, float* t2) {
  float qx = px - cx, qy = py - cy, qz = pz - cz;
  float dd = dx * dx + dy * dy + dz * dz;
  float b = dx * qx + dy * qy + dz * qz;
  float cc = (qx * qx + qy * qy + qz * qz) - r2;
  float disc = mcfl_fmax(0.0f, b * b - dd * cc);
  float sq = mcfl_sqrt(disc);
  *t1 = (b - sq) / dd;
  *t2 = (b + sq) / dd;
}

static uint32_t orc_move_to_boundary(const ctx_t* c, orc_photon* p, uint32_t old_bulk, float max_s) {
  const orc_sim* s = c->s;
  float search_step = 1.0f / (2.0f * s->gs); /* :710 */
  float total_move = 0.0f;
  int64_t index = 1;
  uint32_t present = old_bulk;
  /* Bounded: a photon inside an exit-labelled region (0 or n_bulks+1)
   * marches at a surface with a step of zero and the reference never leaves
   * the loop; any other march ends within about max_s*2gs passes, so twice
   * that only stops the endless ones. */
  const uint32_t passes = (uint32_t)(max_s * 4.0f * s->gs) + 64u;
  uint32_t pass = 0;
  while (present == old_bulk && total_move < max_s && pass++ < passes) { /* :717-756 */
    p->x = mcfl_fma(p->dx, search_step, p->x);
    p->y = mcfl_fma(p->dy, search_step, p->y);
    p->z = mcfl_fma(p->dz, search_step, p->z);
    p->tof += mcfl_d2u64_rz((double)search_step / C_CMFS);
    if (mcfl_fma(p->dz, search_step, p->z) < 0.0f) {
      present = 0;
      index = -1;
      search_step = -p->z / p->dz;
    } else if (mcfl_fma(p->dz, search_step, p->z) > s->esp) {
      present = c->last;
      index = -1;
      search_step = (s->esp - p->z) / p->dz;
    } else if (mcfl_fabs(mcfl_fma(p->dx, search_step, p->x)) >= s->halfx ||
               mcfl_fabs(mcfl_fma(p->dy, search_step, p->y)) >= s->halfy) {
      present = 1;
      index = -1;
    } else {
      index = vindex(s, p->x, p->y, p->z);
      present = s->labels[index];
    }
    total_move += search_step;
  }
  if (index >= 0 && present != old_bulk) { /* :759-844 */
    uint32_t nxy = (uint32_t)s->nx * (uint32_t)s->ny;
    uint32_t iz = (uint32_t)index / nxy;
    uint32_t rem = (uint32_t)index - iz * nxy;
    uint32_t iy = rem / (uint32_t)s->nx;
    uint32_t ix = rem % (uint32_t)s->nx;
    double half = 1.0 / (2.0 * (double)s->gs);
    float xv = (float)((double)((float)ix / s->gs - s->halfx) + half);
    float yv = (float)((double)((float)iy / s->gs - s->halfy) + half);
    float zv = (float)((double)((float)iz / s->gs) + half);
    double rv = 0.707107 * (double)(1.0f / s->gs);
    float r2 = (float)(rv * rv);
    float t1, t2;
    backstep(p->x, p->y, p->z, p->dx, p->dy, p->dz, xv, yv, zv, r2, &t1, &t2);
    float x1 = p->x - p->dx * t1, y1 = p->y - p->dy * t1, z1 = p->z - p->dz * t1;
    float x2 = p->x - p->dx * t2, y2 = p->y - p->dy * t2, z2 = p->z - p->dz * t2;
    float d1 = (x1 - xv) * (x1 - xv) + (y1 - yv) * (y1 - yv) + (z1 - zv) * (z1 - zv);
    float d2 = (x2 - xv) * (x2 - xv) + (y2 - yv) * (y2 - yv) + (z2 - zv) * (z2 - zv);
    if (d1 < d2) {
      p->x = x1; p->y = y1; p->z = z1;
      p->tof -= mcfl_d2u64_rz((double)t1 / C_CMFS);
    } else {
      p->x = x2; p->y = y2; p->z = z2;
      p->tof -= mcfl_d2u64_rz((double)t2 / C_CMFS);
    }
  }
  return present;
}

/* ---------------------------------------------------------- tallies / step */

typedef struct {
  uint64_t *Rd, *Tt, *fhd, *axyz, *time, *Rd2, *Tt2;
  uint64_t tot[ORC_NTOT];
  int scaled;        /* image increments are floor(w*exit_scale + 0.5) */
  double exit_scale;
} tl_t;

static void exit_tally(const ctx_t* c, const orc_photon* p, int top, tl_t* t) {
  const orc_sim* s = c->s;
  float size_x = (s->det_dx * (float)s->det_nx) * 0.5f; /* :259-260 */
  float size_y = (s->det_dy * (float)s->det_ny) * 0.5f;
  uint64_t* img = top ? t->Rd : t->Tt;
  uint64_t* img2 = top ? t->Rd2 : t->Tt2;
  t->tot[top ? ORC_T_R : ORC_T_T] += p->weight;
  if (mcfl_fabs(p->x - s->det_x0) < size_x && mcfl_fabs(p->y - s->det_y0) < size_y) { /* :371-377 / :398-404 */
    uint32_t iy = mcfl_f2u_rz((p->y - s->det_y0 + size_y) / s->det_dy);
    uint32_t ix = mcfl_f2u_rz((p->x - s->det_x0 + size_x) / s->det_dx);
    uint64_t idx = (uint64_t)iy * (uint64_t)s->det_nx + ix;
    if (idx < c->npix) {
      if (t->scaled)
        img[idx] += (uint64_t)((double)p->weight * t->exit_scale + 0.5);
      else
        img[idx] += p->weight;
      if (img2) img2[idx] += ((uint64_t)p->weight * (uint64_t)p->weight) >> 32;
    }
  }
  /* :379-390 / :406-417, outside the image window test */
  if (s->do_temp && s->t_rort == (top ? 0u : 1u) && p->tof < s->t_max) {
    for (uint32_t xp = 0; xp < s->t_ndx; ++xp)
      for (uint32_t yp = 0; yp < s->t_ndy; ++yp) {
        float ddx = p->x - s->t_posx[xp], ddy = p->y - s->t_posy[yp];
        if (ddx * ddx + ddy * ddy < s->t_r * s->t_r) {
          uint32_t bin = mcfl_f2u_rz(((float)p->tof / (float)s->t_max) * (float)s->t_bins);
          if (bin < s->t_bins) t->time[xp + s->t_ndx * (yp + s->t_ndy * (uint64_t)bin)] += p->weight;
        }
      }
  }
}

/* PhotonSurvive, CUDAMCFLtransport.cu:850-864.  Returns 0 when the photon dies. */
static int survive(orc_photon* p, uint64_t* x, uint32_t a, tl_t* t) {
  if (p->weight > WEIGHTI) return 1;
  if (p->weight == 0u) return 0;
  if (mcfl_rand_co(x, a) < CHANCE) {
    uint32_t nw = mcfl_f2u_rn((float)p->weight / CHANCE);
    t->tot[ORC_T_ROUL_GAIN] += (uint64_t)(nw - p->weight);
    p->weight = nw;
    return 1;
  }
  t->tot[ORC_T_ROUL_LOSS] += p->weight;
  p->weight = 0u;
  return 0;
}

/* Inside the method-1 inclusion (CUDAMCFLtransport.cu:83-86, 201-204). */
static int in_sphere(const orc_sim* s, float x, float y, float z) {
  float dx = x - s->inc_x, dy = y - s->inc_y, dz = z - s->inc_z;
  return (dx * dx + dy * dy) + dz * dz <= s->inc_r * s->inc_r;
}

/* One iteration of the MCd loop body (bulk method 1, CUDAMCFLtransport.cu:79-226):
 * step sampled only inside the inclusion, deterministic mutr outside (:99);
 * layer bounds clip the step (:104-115); the step length is zeroed on any
 * layer change, so a reflected or transmitted photon neither absorbs nor
 * scatters at the interface (:142); absorption rounds to nearest with no kill
 * at zero (:205-223). */
static int orc_step_layered(const ctx_t* c, orc_photon* p, uint64_t* x, uint32_t a, tl_t* t) {
  const orc_sim* s = c->s;
  const uint32_t L = p->bulkpos;
  const mat_t* ml = &c->mt[L];
  float st;
  t->tot[ORC_T_STEPS] += 1;
  if (in_sphere(s, p->x, p->y, p->z))
    st = (c->inc.mutr != FLT_MAX) ? -mcfl_log(mcfl_rand_oc(x, a)) * c->inc.mutr : 2.0f; /* :87-91 */
  else
    st = (ml->mutr != FLT_MAX) ? ml->mutr : 2.0f; /* :94-100 */
  int nl = (int)L;
  if (mcfl_fma(st, p->dz, p->z) < s->z_min[L]) { /* :106-109 */
    nl--;
    st = (s->z_min[L] - p->z) / p->dz;
  }
  if (mcfl_fma(st, p->dz, p->z) >= s->z_max[L]) { /* :112-115 */
    nl++;
    st = (s->z_max[L] - p->z) / p->dz;
  }
  p->x = mcfl_fma(p->dx, st, p->x); /* :118-120 */
  p->y = mcfl_fma(p->dy, st, p->y);
  p->z = mcfl_fma(p->dz, st, p->z);
  p->tof += mcfl_d2u64_rz((double)(st * ml->n) / C_CMFS); /* :123 */
  if (p->z > s->z_max[L]) p->z = s->z_max[L];             /* :125-126 */
  if (p->z < s->z_min[L]) p->z = s->z_min[L];
  int in = inside(s, p->x, p->y, p->z);
  uint32_t vi = in ? vindex(s, p->x, p->y, p->z) : 0u;
  if (s->fhd_on && in) t->fhd[vi] += p->weight; /* :129-138 */
  if (nl != (int)L) {                           /* :141-196 */
    st = 0.0f;
    if (!orc_reflect(c, p, (uint32_t)nl, x, a)) {
      if (nl == 0) { /* :147-170 diffuse reflectance */
        exit_tally(c, p, 1, t);
        p->weight = 0u;
      }
      if (nl > (int)s->n_bulks) { /* :172-194 diffuse transmittance */
        exit_tally(c, p, 0, t);
        p->weight = 0u;
      }
    }
  }
  if (st > 0.0f) { /* :200-224, the material at the end point */
    const mat_t* m = in_sphere(s, p->x, p->y, p->z) ? &c->inc : &c->mt[p->bulkpos];
    uint32_t w = mcfl_f2u_rn(m->mua_mutr * (float)p->weight);
    p->weight -= w; /* the reference's underflow test on unsigned values is vacuous */
    t->tot[ORC_T_A] += w;
    if (s->axyz_on && in) t->axyz[vi] += w;
    orc_spin(m, p, x, a);
  }
  return survive(p, x, a, t);
}

/* One iteration of the MCd3D loop body up to and including the roulette
 * (CUDAMCFLtransport.cu:289-435).  Returns 0 when the photon has died. */
static int orc_step(const ctx_t* c, orc_photon* p, uint64_t* x, uint32_t a, tl_t* t) {
  const orc_sim* s = c->s;
  if (s->layered) return orc_step_layered(c, p, x, a, t);
  const mat_t* m = &c->mt[p->bulkpos];
  float st;
  uint32_t nb;
  t->tot[ORC_T_STEPS] += 1;
  if (m->mutr != FLT_MAX) {
    st = -mcfl_log(mcfl_rand_oc(x, a)) * m->mutr; /* :290 */
    nb = p->bulkpos;
    if (mcfl_fma(st, p->dz, p->z) < 0.0f) { /* :297-301 */
      nb = 0;
      st = -p->z / p->dz;
    }
    if (mcfl_fma(st, p->dz, p->z) > s->esp) { /* :304-308 */
      nb = c->last;
      st = (s->esp - p->z) / p->dz;
    }
    p->x = mcfl_fma(p->dx, st, p->x); /* :311-313 */
    p->y = mcfl_fma(p->dy, st, p->y);
    p->z = mcfl_fma(p->dz, st, p->z);
    p->tof += mcfl_d2u64_rz((double)(st * m->n) / C_CMFS); /* :316 */
    if (nb != 0 && nb != c->last) {                        /* :319-333 */
      if (inside(s, p->x, p->y, p->z))
        nb = s->labels[vindex(s, p->x, p->y, p->z)];
      else
        nb = 1;
    }
  } else {
    st = 100.0f; /* :338-339 */
    nb = orc_move_to_boundary(c, p, p->bulkpos, st);
  }
  if (nb != p->bulkpos) orc_reflect(c, p, nb, x, a); /* :352-355 */

  int in = inside(s, p->x, p->y, p->z);
  uint32_t vi = in ? vindex(s, p->x, p->y, p->z) : 0u;
  if (s->fhd_on && in) t->fhd[vi] += p->weight; /* :358-366 */

  if (p->bulkpos == 0) { /* :369-394 */
    exit_tally(c, p, 1, t);
    p->weight = 0u;
    st = 0.0f;
  }
  if (p->bulkpos == c->last) { /* :396-421 */
    exit_tally(c, p, 0, t);
    p->weight = 0u;
    st = 0.0f;
  }
  if (st > 0.0f) { /* :425-433 */
    const mat_t* mm = &c->mt[p->bulkpos];
    uint32_t w = mcfl_f2u_ru(mm->mua_mutr * (float)p->weight);
    if (w > 0u) {
      p->weight -= w;
      t->tot[ORC_T_A] += w;
      if (s->axyz_on && in) t->axyz[vi] += w;
    } else {
      t->tot[ORC_T_ZERO_KILL] += p->weight;
      p->weight = 0u;
    }
    orc_spin(mm, p, x, a);
  }
  return survive(p, x, a, t);
}

/* Run one photon from its launch to its death. */
static void run_photon(const ctx_t* c, orc_photon* p, uint64_t* x, uint32_t a, tl_t* t) {
  t->tot[ORC_T_PHOTONS] += 1;
  t->tot[ORC_T_LAUNCHED_W] += p->weight;
  uint32_t n = 0;
  while (orc_step(c, p, x, a, t)) {
    if (++n >= g_max_steps) {
      t->tot[ORC_T_ABORTED] += p->weight;
      p->weight = 0u;
      break;
    }
  }
}

/* --------------------------------------------------- thread-private tallies */

static int tl_alloc(tl_t* t, const ctx_t* c, const orc_tallies* out, int scaled, double exit_scale) {
  memset(t, 0, sizeof *t);
  const orc_sim* s = c->s;
  t->scaled = scaled;
  t->exit_scale = exit_scale;
  t->Rd = calloc(c->npix, 8);
  t->Tt = calloc(c->npix, 8);
  if (out->Rd2) t->Rd2 = calloc(c->npix, 8);
  if (out->Tt2) t->Tt2 = calloc(c->npix, 8);
  if (s->fhd_on) t->fhd = calloc(c->nvox, 8);
  if (s->axyz_on) t->axyz = calloc(c->nvox, 8);
  if (s->do_temp) t->time = calloc((size_t)s->t_ndx * s->t_ndy * s->t_bins + 1, 8);
  return t->Rd && t->Tt;
}

static void add_arr(uint64_t* dst, const uint64_t* src, size_t n) {
  if (dst && src)
    for (size_t i = 0; i < n; ++i) dst[i] += src[i];
}

static void tl_merge_free(tl_t* t, const ctx_t* c, orc_tallies* out) {
  const orc_sim* s = c->s;
  add_arr(out->Rd, t->Rd, c->npix);
  add_arr(out->Tt, t->Tt, c->npix);
  add_arr(out->Rd2, t->Rd2, c->npix);
  add_arr(out->Tt2, t->Tt2, c->npix);
  if (s->fhd_on) add_arr(out->fhd, t->fhd, c->nvox);
  if (s->axyz_on) add_arr(out->axyz, t->axyz, c->nvox);
  if (s->do_temp) add_arr(out->time, t->time, (size_t)s->t_ndx * s->t_ndy * s->t_bins);
  for (int i = 0; i < ORC_NTOT; ++i) out->tot[i] += t->tot[i];
  free(t->Rd); free(t->Tt); free(t->Rd2); free(t->Tt2); free(t->fhd); free(t->axyz); free(t->time);
}

static int nthreads_eff(int n) {
#ifdef _OPENMP
  return n > 0 ? n : omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

/* ------------------------------------------------------------ run drivers */

int orc_run_photons(const orc_sim* sim, const uint32_t* a_tab, uint32_t P, uint64_t key, uint64_t k0, uint64_t k1,
                    double exit_scale, orc_tallies* out, int nthreads) {
  if (P == 0 || (P & (P - 1)) != 0) return -1;
  ctx_t* c = malloc(sizeof(ctx_t));
  ctx_init(c, sim);
  int nt = nthreads_eff(nthreads);
  int err = 0;
#pragma omp parallel num_threads(nt)
  {
    tl_t t;
    if (!tl_alloc(&t, c, out, exit_scale > 0.0, exit_scale)) err = 1;
#pragma omp for schedule(dynamic, 64)
    for (int64_t k = (int64_t)k0; k < (int64_t)k1; ++k) {
      uint32_t a = a_tab[(uint64_t)k & (P - 1u)];
      uint64_t x = mcfl_photon_state(key, (uint64_t)k, a);
      orc_photon p;
      orc_launch(c, sim->src_kind, sim->xi, sim->yi, sim->zi, &p, &x, a);
      run_photon(c, &p, &x, a, &t);
    }
#pragma omp critical
    tl_merge_free(&t, c, out);
  }
  free(c);
  return err;
}

/* CUDAMCFLmain.cu:117-165 with the global photon counter replaced by fixed
 * per-lane quotas, which fixes the trajectory set. */
int orc_run_lanes(const orc_sim* sim, uint64_t* xs, const uint32_t* as, uint32_t L, uint64_t N, orc_tallies* out,
                  int nthreads) {
  if (N < L) return -2; /* the reference never terminates (N - NUM_THREADS underflows) */
  ctx_t* c = malloc(sizeof(ctx_t));
  ctx_init(c, sim);
  int nt = nthreads_eff(nthreads);
  int err = 0;
  uint64_t q = N / L, rem = N % L;
#pragma omp parallel num_threads(nt)
  {
    tl_t t;
    if (!tl_alloc(&t, c, out, 0, 0.0)) err = 1;
#pragma omp for schedule(dynamic, 16)
    for (int64_t i = 0; i < (int64_t)L; ++i) {
      uint64_t quota = q + ((uint64_t)i < rem ? 1u : 0u);
      uint64_t x = xs[i];
      uint32_t a = as[i];
      orc_photon p;
      uint64_t xl = x; /* LaunchPhoton_Global works on a copy (:574-581) */
      orc_launch(c, sim->src_kind, sim->xi, sim->yi, sim->zi, &p, &xl, a);
      for (uint64_t j = 0; j < quota; ++j) {
        if (j > 0) orc_launch(c, sim->src_kind, sim->xi, sim->yi, sim->zi, &p, &x, a); /* :439 */
        run_photon(c, &p, &x, a, &t);
      }
      xs[i] = x; /* CUDAMCFLmem.cu:42 copies the advanced states back */
    }
#pragma omp critical
    tl_merge_free(&t, c, out);
  }
  free(c);
  return err;
}

/* Source corner of voxel v (CUDAMCFLmain.cu:481-483). */
static void voxel_corner(const orc_sim* s, uint32_t v, float* xi, float* yi, float* zi) {
  uint32_t nxy = (uint32_t)s->nx * (uint32_t)s->ny;
  uint32_t iz = v / nxy, rem = v - iz * nxy, iy = rem / (uint32_t)s->nx, ix = rem % (uint32_t)s->nx;
  *xi = ((float)ix / s->gs) - s->halfx;
  *yi = ((float)iy / s->gs) - s->halfy;
  *zi = ((float)iz / s->gs);
}

int orc_run_fluor_jobs(const orc_sim* sim, const uint32_t* a_tab, uint32_t P, uint32_t njobs, const uint32_t* vox,
                       const uint64_t* keys, const double* scale, uint64_t n_per_job, orc_tallies* out,
                       int nthreads) {
  if (P == 0 || (P & (P - 1)) != 0) return -1;
  ctx_t* c = malloc(sizeof(ctx_t));
  ctx_init(c, sim);
  int nt = nthreads_eff(nthreads);
  int err = 0;
  int64_t total = (int64_t)njobs * (int64_t)n_per_job;
#pragma omp parallel num_threads(nt)
  {
    tl_t t;
    if (!tl_alloc(&t, c, out, 1, 1.0)) err = 1;
#pragma omp for schedule(dynamic, 64)
    for (int64_t g = 0; g < total; ++g) {
      uint32_t j = (uint32_t)(g / (int64_t)n_per_job);
      uint64_t k = (uint64_t)(g % (int64_t)n_per_job);
      t.exit_scale = scale[j];
      float xi, yi, zi;
      voxel_corner(sim, vox[j], &xi, &yi, &zi);
      uint32_t a = a_tab[k & (P - 1u)];
      uint64_t x = mcfl_photon_state(keys[j], k, a);
      orc_photon p;
      orc_launch(c, ORC_SRC_ISOTROPIC, xi, yi, zi, &p, &x, a);
      run_photon(c, &p, &x, a, &t);
    }
#pragma omp critical
    tl_merge_free(&t, c, out);
  }
  free(c);
  return err;
}

/* The reference's emission loop on its lanes (CUDAMCFLmain.cu:471-594,
 * DoOneSimulationFl :196-288) with the images in fixed point: voxels vox[j]
 * in the reference's loop order, each one simulation of nv photons on the L
 * lane streams that continue from voxel to voxel; in every simulation a
 * lane's first photon replays its launch draws (LaunchPhoton_Global does not
 * store the advanced state, CUDAMCFLtransport.cu:574-581) and the lane runs
 * nv/L photons (the first nv%L lanes one more, as orc_run_lanes).  Exits add
 * floor(w * scale[j] + 0.5).  x is advanced in place. */
int orc_run_fluor_lanes(const orc_sim* sim, uint64_t* xs, const uint32_t* as, uint32_t L, uint32_t nvox,
                        const uint32_t* vox, const double* scale, uint64_t nv, orc_tallies* out, int nthreads) {
  if (nv < L) return -2; /* DoOneSimulationFl fails */
  ctx_t* c = malloc(sizeof(ctx_t));
  ctx_init(c, sim);
  int nt = nthreads_eff(nthreads);
  int err = 0;
  const uint64_t q = nv / L, rem = nv % L;
#pragma omp parallel num_threads(nt)
  {
    tl_t t;
    if (!tl_alloc(&t, c, out, 1, 1.0)) err = 1;
#pragma omp for schedule(dynamic, 4)
    for (int64_t i = 0; i < (int64_t)L; ++i) {
      const uint64_t quota = q + ((uint64_t)i < rem ? 1u : 0u);
      uint64_t x = xs[i];
      const uint32_t a = as[i];
      for (uint32_t j = 0; j < nvox; ++j) {
        float xi, yi, zi;
        voxel_corner(sim, vox[j], &xi, &yi, &zi);
        t.exit_scale = scale[j];
        orc_photon p;
        uint64_t xl = x;
        orc_launch(c, ORC_SRC_ISOTROPIC, xi, yi, zi, &p, &xl, a);
        for (uint64_t k = 0; k < quota; ++k) {
          if (k > 0) orc_launch(c, ORC_SRC_ISOTROPIC, xi, yi, zi, &p, &x, a);
          run_photon(c, &p, &x, a, &t);
        }
      }
      xs[i] = x;
    }
#pragma omp critical
    tl_merge_free(&t, c, out);
  }
  free(c);
  return err;
}

uint64_t orc_fluor_reference(const orc_sim* sim, const double* Fx, uint64_t* xs, const uint32_t* as, uint32_t L,
                             uint64_t nv, double* FlR, double* FlT, int nthreads) {
  ctx_t* c = malloc(sizeof(ctx_t));
  ctx_init(c, sim);
  int nt = nthreads_eff(nthreads);
  const orc_sim* s = sim;
  size_t npix = c->npix;
  size_t xy = (size_t)s->det_nx * (size_t)s->det_ny;
  uint64_t* R = malloc(npix * 8);
  uint64_t* T = malloc(npix * 8);
  double dxdy = (double)s->det_dx * (double)s->det_dy;
  uint64_t simulated = 0;
  for (size_t i = 0; i < xy; ++i) FlR[i] = FlT[i] = 0.0;
  for (int ix = 0; ix < s->nx; ++ix)
    for (int iy = 0; iy < s->ny; ++iy)
      for (int iz = 0; iz < s->nz; ++iz) {
        uint32_t v = (uint32_t)ix + (uint32_t)s->nx * ((uint32_t)iy + (uint32_t)iz * (uint32_t)s->ny);
        const orc_bulk* b = &s->bulks[s->labels[v]];
        if (!(b->albedof < 1.0f)) { /* :551,558-561: voxelw = 0, status 1 */
          simulated += 1;
          continue;
        }
        if (nv < L) continue; /* DoOneSimulationFl fails (status 0), x not copied back */
        orc_sim sv = *s;
        sv.src_kind = ORC_SRC_ISOTROPIC;
        voxel_corner(s, v, &sv.xi, &sv.yi, &sv.zi);
        orc_tallies t;
        memset(&t, 0, sizeof t);
        memset(R, 0, npix * 8);
        memset(T, 0, npix * 8);
        t.Rd = R;
        t.Tt = T;
        sv.fhd_on = 0;
        sv.axyz_on = 0;
        sv.do_temp = 0;
        orc_run_lanes(&sv, xs, as, L, nv, &t, nt);
        simulated += nv;
        double voxelw = ((double)b->eY * (double)(1 - b->albedof) * Fx[v]) / (double)(nv * 0xFFFFFFFFull);
        for (size_t ij = 0; ij < xy; ++ij) { /* :572-577 */
          double tr = voxelw * (double)R[ij];
          double tt = voxelw * (double)T[ij];
          if (FlR[ij] + tr < DBL_MAX) FlR[ij] += tr / dxdy;
          if (FlT[ij] + tt < DBL_MAX) FlT[ij] += tt / dxdy;
        }
      }
  free(R);
  free(T);
  free(c);
  return simulated;
}

/* ---------------------------------------------------------- unit exposure */

int orc_reflect_test(float n1, float n2, float dz_in, float xi, float* dz_out, float* dx_scale) {
  orc_sim s;
  memset(&s, 0, sizeof s);
  orc_bulk b[3];
  memset(b, 0, sizeof b);
  b[0].n = n1;
  b[1].n = n2;
  b[2].n = n2;
  s.n_bulks = 1;
  s.bulks = b;
  ctx_t* c = malloc(sizeof(ctx_t));
  ctx_init(c, &s);
  orc_photon p;
  memset(&p, 0, sizeof p);
  p.dx = 1.0f;
  p.dz = dz_in;
  p.bulkpos = 0;
  /* a state whose next [0,1) draw is xi: x_lo*a + c with a = 1, c = 0 */
  uint32_t u = (uint32_t)((double)xi * 4294967296.0);
  uint64_t x = (uint64_t)u;
  int r = orc_reflect(c, &p, 1, &x, 1u);
  *dz_out = p.dz;
  *dx_scale = p.dx;
  free(c);
  return r;
}

float orc_fresnel_test(float n1, float n2, float ci) { return fresnel_r(n1, n2, ci); }

void orc_spin_test(float g, const float* din, uint64_t* x, uint32_t a, float* dout) {
  orc_bulk b;
  memset(&b, 0, sizeof b);
  b.g = g;
  b.mutr = 1.0f;
  orc_sim s;
  memset(&s, 0, sizeof s);
  s.n_bulks = 0;
  orc_bulk bb[2] = {b, b};
  s.bulks = bb;
  ctx_t* c = malloc(sizeof(ctx_t));
  ctx_init(c, &s);
  orc_photon p;
  memset(&p, 0, sizeof p);
  p.dx = din[0];
  p.dy = din[1];
  p.dz = din[2];
  orc_spin(&c->mt[0], &p, x, a);
  dout[0] = p.dx;
  dout[1] = p.dy;
  dout[2] = p.dz;
  free(c);
}

void orc_backstep_test(const float* p, const float* d, const float* cc, float r2, float* t1, float* t2) {
  backstep(p[0], p[1], p[2], d[0], d[1], d[2], cc[0], cc[1], cc[2], r2, t1, t2);
}

/* n photons launched by orc_launch from the source (kind, s->xi, s->yi,
 * s->zi) on photon streams: photon k uses multiplier a_tab[k & (P-1)] and
 * the state mcfl_photon_state(key, k, a).  pos_dir[6k..6k+5] = x, y, z, dx,
 * dy, dz; weight[k], label[k] the launch weight and entry label. */
void orc_launch_sample(const orc_sim* s, int kind, const uint32_t* a_tab, uint32_t P, uint64_t key, uint64_t n,
                       float* pos_dir, uint32_t* weight, uint32_t* label) {
  ctx_t* c = malloc(sizeof(ctx_t));
  ctx_init(c, s);
  for (uint64_t k = 0; k < n; ++k) {
    uint32_t a = a_tab[k & (P - 1u)];
    uint64_t x = mcfl_photon_state(key, k, a);
    orc_photon p;
    orc_launch(c, kind, s->xi, s->yi, s->zi, &p, &x, a);
    float* o = pos_dir + 6 * k;
    o[0] = p.x; o[1] = p.y; o[2] = p.z; o[3] = p.dx; o[4] = p.dy; o[5] = p.dz;
    weight[k] = p.weight;
    label[k] = p.bulkpos;
  }
  free(c);
}

/* survive() on a photon of weight w with generator state *x, multiplier a
 * (advanced by the draws taken).  Returns 1 if the photon lives; *w_out its
 * weight afterwards, gain/loss the roulette totals it adds. */
int orc_survive_test(uint32_t w, uint64_t* x, uint32_t a, uint32_t* w_out, uint64_t* gain, uint64_t* loss) {
  orc_photon p;
  memset(&p, 0, sizeof p);
  p.weight = w;
  tl_t t;
  memset(&t, 0, sizeof t);
  int r = survive(&p, x, a, &t);
  *w_out = p.weight;
  *gain = t.tot[ORC_T_ROUL_GAIN];
  *loss = t.tot[ORC_T_ROUL_LOSS];
  return r;
}

/* survive() on n photons of weight w, photon k drawing from the stream
 * (key, k) with multiplier a_tab[k & (P-1)]: the number that live and the
 * sum of their weights afterwards. */
void orc_survive_batch(uint32_t w, const uint32_t* a_tab, uint32_t P, uint64_t key, uint64_t n, uint64_t* lived,
                       uint64_t* weight_sum) {
  uint64_t nl = 0, sw = 0;
  for (uint64_t k = 0; k < n; ++k) {
    uint32_t a = a_tab[k & (P - 1u)];
    uint64_t x = mcfl_photon_state(key, k, a);
    uint32_t wo;
    uint64_t g, l;
    if (orc_survive_test(w, &x, a, &wo, &g, &l)) {
      ++nl;
      sw += wo;
    }
  }
  *lived = nl;
  *weight_sum = sw;
}

/* orc_move_to_boundary on a photon at pd[0..2] moving along pd[3..5] inside a
 * region of label old_label.  Returns the label found; out[0..2] the
 * position afterwards, *tof the time of flight it accumulated (from 2^40, so
 * the back step cannot wrap). */
uint32_t orc_march_test(const orc_sim* s, const float* pd, uint32_t old_label, float max_s, float* out,
                        uint64_t* tof) {
  ctx_t* c = malloc(sizeof(ctx_t));
  ctx_init(c, s);
  orc_photon p;
  memset(&p, 0, sizeof p);
  p.x = pd[0]; p.y = pd[1]; p.z = pd[2]; p.dx = pd[3]; p.dy = pd[4]; p.dz = pd[5];
  p.bulkpos = old_label;
  p.tof = 1ull << 40;
  uint32_t lab = orc_move_to_boundary(c, &p, old_label, max_s);
  out[0] = p.x; out[1] = p.y; out[2] = p.z;
  *tof = p.tof - (1ull << 40);
  free(c);
  return lab;
}

/* ----------------------------------------------- math / generator exposure */

void orc_math_eval(int fn, const float* in, float* out, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    float s, c;
    switch (fn) {
      case 0: out[i] = mcfl_log(in[i]); break;
      case 1: mcfl_sincos_turn(in[i], &s, &c); out[i] = s; break;
      case 2: mcfl_sincos_turn(in[i], &s, &c); out[i] = c; break;
      case 3: { uint32_t u; memcpy(&u, &in[i], 4); out[i] = mcfl_u2f_rz(u); } break;
      case 4: out[i] = mcfl_rsqrt(in[i]); break;
      case 5: out[i] = mcfl_sqrt_rs(in[i]); break;
      case 6: out[i] = mcfl_renorm(in[i]); break;
      default: out[i] = 0.0f;
    }
  }
}

void orc_mwc_draws(uint64_t x, uint32_t a, size_t n, uint64_t* states, float* co) {
  for (size_t i = 0; i < n; ++i) {
    co[i] = mcfl_rand_co(&x, a);
    states[i] = x;
  }
}

uint64_t orc_photon_state(uint64_t key, uint64_t k, uint32_t a) { return mcfl_photon_state(key, k, a); }

/* Float bit patterns b in [lo, hi) for which the kernel's time-of-flight
   quotient (mcfl_fmath.h mcfl_tof_div) differs from the IEEE division used
   here (CUDAMCFLtransport.cu:316). */
uint64_t orc_tof_div_mismatches(uint32_t lo, uint32_t hi) {
  uint64_t bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
  for (int64_t b = (int64_t)lo; b < (int64_t)hi; ++b) {
    const float f = mcfl_bitsf((uint32_t)b);
    const double q = (double)f;
    if (!(mcfl_tof_div(q) == q / C_CMFS) || mcfl_tof_inc(f) != mcfl_d2u64_rz(q / C_CMFS)) ++bad;
  }
  return bad;
}
