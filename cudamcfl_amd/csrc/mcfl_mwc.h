/*
 * mcfl_mwc.h -- the multiply-with-carry generator and photon stream seeding,
 * shared by the HIP kernel and host C code.
 *
 * Generator (CUDAMCFLrng.cu:17-29): packed state x = c*2^32 + x_lo,
 *     x <- x_lo * a + c
 * draw in [0,1) = (float, rounded toward zero)(uint32)x * 2^-32 and the (0,1]
 * draw is 1 - that.  For a safe-prime multiplier (p = a*2^32-1 and (p-1)/2
 * prime) one step is multiplication by a modulo p, so the n-th state is
 * a^n * x0 mod p.
 *
 * Photon streams (no counterpart in the reference): photon k of a job with
 * 64-bit key K draws from multiplier a[k mod P] (P a power of two, the
 * per-lane multipliers of the primes table) starting at a state derived from
 * splitmix64(K + (k+1)*golden).  The state is then independent of which lane
 * or which launch runs the photon, so a job's results do not depend on
 * scheduling.  Validity of the derived state follows the reference's seeding
 * rules (CUDAMCFLrng.cu:56-60,72): not (0,0), carry < a-1, low word < 2^32-1;
 * the carry is floor(u * a / 2^32) for a uniform 32-bit u, the integer form of
 * CUDAMCFLrng.cu:78.
 */
#ifndef MCFL_MWC_H
#define MCFL_MWC_H

#include "mcfl_fmath.h"

MCFL_HD uint32_t mcfl_mwc_next(uint64_t* x, uint32_t a) {
  *x = (uint64_t)(uint32_t)(*x) * (uint64_t)a + (*x >> 32);
  return (uint32_t)(*x);
}

/* [0,1) */
MCFL_HD float mcfl_rand_co(uint64_t* x, uint32_t a) { return mcfl_u2f_rz(mcfl_mwc_next(x, a)) * 0x1p-32f; }
/* (0,1]: 1 - that, one FMA (the product with 2^-32 is exact, so this is the
   same rounding as 1.0f - mcfl_rand_co) */
MCFL_HD float mcfl_rand_oc(uint64_t* x, uint32_t a) {
  return mcfl_fma(mcfl_u2f_rz(mcfl_mwc_next(x, a)), -0x1p-32f, 1.0f);
}

/* sin and cos of 2 pi times a [0,1) draw: the quarter-turn argument
   4 * draw = u * 2^-30 straight from the converted integer (exact) */
MCFL_HD void mcfl_rand_sincos(uint64_t* x, uint32_t a, float* sn, float* cs) {
  mcfl_sincos_quarter(mcfl_u2f_rz(mcfl_mwc_next(x, a)) * 0x1p-30f, sn, cs);
}

MCFL_HD uint64_t mcfl_mix64(uint64_t z) {
  z ^= z >> 30;
  z *= 0xBF58476D1CE4E5B9ull;
  z ^= z >> 27;
  z *= 0x94D049BB133111EBull;
  z ^= z >> 31;
  return z;
}

#define MCFL_GOLDEN64 0x9E3779B97F4A7C15ull

/* Key of stream number `stream` under user seed `seed` (job keys, voxel keys). */
MCFL_HD uint64_t mcfl_key(uint64_t seed, uint64_t stream) {
  return mcfl_mix64(mcfl_mix64(seed ^ 0xD1B54A32D192ED03ull) + (stream + 1ull) * MCFL_GOLDEN64);
}

/* A state x is usable with multiplier a when it is not 0, its carry is below
   a-1 and its low word below 2^32-1 (CUDAMCFLrng.cu:56-60,72). */
MCFL_HD int mcfl_mwc_state_ok(uint64_t x, uint32_t a) {
  return x != 0ull && (uint32_t)(x >> 32) < a - 1u && (uint32_t)x != 0xFFFFFFFFu;
}

/* Valid MWC state for photon k of the job with key `key`, multiplier a. */
MCFL_HD uint64_t mcfl_photon_state(uint64_t key, uint64_t k, uint32_t a) {
  uint64_t h = key + (k + 1ull) * MCFL_GOLDEN64;
  for (;;) {
    uint64_t z = mcfl_mix64(h);
    uint32_t lo = (uint32_t)z;
    uint32_t c = (uint32_t)(((uint64_t)(uint32_t)(z >> 32) * (uint64_t)a) >> 32);
    const uint64_t x = ((uint64_t)c << 32) | lo;
    if (mcfl_mwc_state_ok(x, a)) return x;
    h += MCFL_GOLDEN64;
  }
}

#endif /* MCFL_MWC_H */
