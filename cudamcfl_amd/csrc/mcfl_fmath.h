/*
 * mcfl_fmath.h -- deterministic FP32 arithmetic shared by the HIP transport
 * kernel and by host C code.
 *
 * The reference evaluates its transport with CUDA fast-math intrinsics
 * (__logf, __sincosf, __fdividef; CUDAMCFLtransport.cu:290,596-603 and
 * Compile750:1 "--use_fast_math").  Those are not reproducible off NVIDIA
 * hardware, so this build defines its own transcendental functions from
 * published methods, built only from IEEE-754 operations that every
 * conforming implementation rounds identically (+, -, *, fused multiply-add,
 * correctly rounded division and square root, exact integer bit operations).
 * Compiled with -ffp-contract=off on both sides, a kernel and a host program
 * that use this header execute bit-identical FP32 arithmetic, so integer
 * tallies of a fixed photon set agree exactly between CPU and GPU.
 *
 *  - logarithm: exponent/mantissa split with the mantissa folded into
 *    (sqrt(1/2), sqrt(2)] by one integer subtraction of the bits of
 *    sqrt(1/2), log1p(f) = f + f^2 R(f) with R a degree-8
 *    polynomial fitted by least squares on Chebyshev nodes (near-minimax,
 *    Remez-style approximation), Cody-Waite two-constant ln2 split.
 *    Max relative error about 2 ulp on (0, 1].
 *  - reciprocal square root: integer-shift seed with Lomont's constant
 *    0x5F375A86 (seed error up to 3.4%) and three Newton-Raphson steps
 *    written with FMA; about 1 ulp.  Square roots on the per-step path are
 *    x * rsqrt(x); renormalisation factors of near-unit vectors use the
 *    first-order expansion of 1/sqrt around 1.
 *  - sine/cosine of 2*pi*t ("turns"): exact quadrant reduction 4t - rint(4t)
 *    (no large-argument reduction is needed because the argument is a
 *    fraction of a turn), then odd/even degree-9/8 polynomials in the
 *    reduced variable fitted the same way.  Max error about 1 ulp.
 *
 * This header is valid C99 and HIP C++.
 */
#ifndef MCFL_FMATH_H
#define MCFL_FMATH_H

#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define MCFL_HD static __host__ __device__ __forceinline__
#else
#include <string.h>
#define MCFL_HD static inline
#endif

MCFL_HD uint32_t mcfl_fbits(float f) {
#if defined(__HIPCC__) || defined(__HIP__)
  return __builtin_bit_cast(uint32_t, f);
#else
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
#endif
}

MCFL_HD float mcfl_bitsf(uint32_t u) {
#if defined(__HIPCC__) || defined(__HIP__)
  return __builtin_bit_cast(float, u);
#else
  float f;
  memcpy(&f, &u, 4);
  return f;
#endif
}

MCFL_HD float mcfl_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
MCFL_HD double mcfl_fmad(double a, double b, double c) { return __builtin_fma(a, b, c); }
MCFL_HD float mcfl_sqrt(float x) { return __builtin_sqrtf(x); }
MCFL_HD float mcfl_rint(float x) { return __builtin_rintf(x); }
MCFL_HD float mcfl_ceil(float x) { return __builtin_ceilf(x); }
MCFL_HD float mcfl_fabs(float x) { return __builtin_fabsf(x); }
MCFL_HD float mcfl_copysign(float m, float s) { return __builtin_copysignf(m, s); }
MCFL_HD float mcfl_fmax(float a, float b) { return a > b ? a : b; }
/* max(x, 0) as one maximum instruction (IEEE maxNum: a NaN x gives 0).  For
   x that is never -0, where maxNum may return either zero: the call sites
   take x = 1 - c*c from one FMA, whose exact zero rounds to +0. */
MCFL_HD float mcfl_clamp0(float x) { return __builtin_fmaxf(x, 0.0f); }

/* Natural logarithm for positive normal floats.  u = 2^e m with m in
   [0x3F3504F4, 0x3FB504F3] (as bits, about [sqrt(1/2), sqrt(2)]): the bits of
   u less those of the lowest m keep m's mantissa field in their low 23 bits
   and e in the bits above, an arithmetic shift's worth of exponent.  (The
   same e and m as splitting u at 1 and halving a mantissa above
   0x3FB504F3; both forms agree for every normal positive float.) */
MCFL_HD float mcfl_log(float u) {
  const uint32_t t = mcfl_fbits(u) - 0x3F3504F4u;
  const int e = (int32_t)t >> 23;
  const float f = mcfl_bitsf((t & 0x007FFFFFu) + 0x3F3504F4u) - 1.0f; /* exact (Sterbenz) */
  float r = -0.07473210990428925f;
  r = mcfl_fma(r, f, 0.128195121884346f);
  r = mcfl_fma(r, f, -0.1326199471950531f);
  r = mcfl_fma(r, f, 0.14196686446666718f);
  r = mcfl_fma(r, f, -0.16608595848083496f);
  r = mcfl_fma(r, f, 0.200011745095253f);
  r = mcfl_fma(r, f, -0.2500157654285431f);
  r = mcfl_fma(r, f, 0.3333333730697632f);
  r = mcfl_fma(r, f, -0.49999988079071045f);
  float y = mcfl_fma(f * f, r, f);
  float ef = (float)e;
  /* ln2 = 0.693145751953125 (16 significant bits) + 1.4286068203094172e-06 */
  return mcfl_fma(ef, 0.693145751953125f, mcfl_fma(ef, 1.4286068203094172e-06f, y));
}

/* 1/sqrt(x) for positive normal x. */
MCFL_HD float mcfl_rsqrt(float x) {
  float y = mcfl_bitsf(0x5F375A86u - (mcfl_fbits(x) >> 1));
  const float h = 0.5f * x;
  float e = mcfl_fma(-h * y, y, 0.5f); /* y <- y (1.5 - x y^2 / 2) */
  y = mcfl_fma(y, e, y);
  e = mcfl_fma(-h * y, y, 0.5f);
  y = mcfl_fma(y, e, y);
  e = mcfl_fma(-h * y, y, 0.5f);
  return mcfl_fma(y, e, y);
}

/* sqrt(x) for x = 0 or a positive normal x, as x * rsqrt(x): within 2 ulp. */
MCFL_HD float mcfl_sqrt_rs(float x) { return x > 0.0f ? x * mcfl_rsqrt(x) : 0.0f; }

/* 1/sqrt(x) as a renormalisation factor: for |x - 1| < 2^-12 the first-order
   expansion 1.5 - x/2 (one FMA) is off by 3/8 (x-1)^2 < 2^-25, below half an
   ulp of the result; farther from 1 the full reciprocal root. */
MCFL_HD float mcfl_renorm(float x) {
  return mcfl_fabs(x - 1.0f) < 0x1p-12f ? mcfl_fma(x, -0.5f, 1.5f) : mcfl_rsqrt(x);
}

/* sin((pi/2) y) and cos((pi/2) y) for y = 4t in [0, 4]. */
MCFL_HD void mcfl_sincos_quarter(float y, float* sn, float* cs) {
  float j = mcfl_rint(y);
  float r = y - j; /* exact, |r| <= 1/2; angle = (pi/2)(j + r) */
  int q = (int)j & 3;
  float r2 = r * r;
  float ps = mcfl_fma(0.0001582550466991961f, r2, -0.0046812924556434155f);
  ps = mcfl_fma(ps, r2, 0.07969258725643158f);
  ps = mcfl_fma(ps, r2, -0.6459640860557556f);
  ps = mcfl_fma(ps, r2, 1.5707963705062866f);
  ps = ps * r;
  float pc = mcfl_fma(0.0009037023992277682f, r2, -0.02086012065410614f);
  pc = mcfl_fma(pc, r2, 0.25366920232772827f);
  pc = mcfl_fma(pc, r2, -1.2337005138397217f);
  pc = mcfl_fma(pc, r2, 1.0f);
  const float s = (q & 1) ? pc : ps;
  const float c = (q & 1) ? ps : pc;
  /* the quadrant's signs as sign-bit flips: sin negative in quadrants 2, 3;
     cos in 1, 2 */
  *sn = mcfl_bitsf(mcfl_fbits(s) ^ ((uint32_t)(q & 2) << 30));
  *cs = mcfl_bitsf(mcfl_fbits(c) ^ ((uint32_t)((q + 1) & 2) << 30));
}

/* sin(2*pi*t) and cos(2*pi*t). */
MCFL_HD void mcfl_sincos_turn(float t, float* sn, float* cs) { mcfl_sincos_quarter(4.0f * t, sn, cs); }

/* ---- conversions with the saturating semantics of the GPU instructions ---- */

/* uint32 -> float, rounding toward zero.  On the host: keep the 24 leading
   bits, then the conversion is exact (0xFFFFFF00 >> clz(u) has ones from the
   leading bit down 24 places, bits above the leading one are zero in u
   anyway; u | 1 keeps clz defined at u = 0).  On the GPU: the conversion
   instruction with the FP32 rounding field of the MODE register (bits 1:0)
   set to toward-zero around it, one vector instruction and two scalar ones
   instead of five vector ones.  The three are one asm statement: the
   compiler does not order the conversion after a MODE write made with the
   builtin, and has moved it out of such a window.  Both forms agree for all
   2^32 inputs (checked on the GPU). */
MCFL_HD float mcfl_u2f_rz(uint32_t u) {
#if defined(__HIP_DEVICE_COMPILE__)
  float f;
  asm volatile(
      "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 0, 2), 3\n\t"
      "v_cvt_f32_u32 %0, %1\n\t"
      "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 0, 2), 0"
      : "=v"(f)
      : "v"(u));
  return f;
#else
  return (float)(u & (0xFFFFFF00u >> __builtin_clz(u | 1u)));
#endif
}

/* float -> uint32 toward zero; NaN and negatives give 0, >= 2^32 saturates.
   On the GPU that is what the conversion instruction does by itself (equal
   to the host form for all 2^32 inputs, checked on the GPU); written as asm
   because the C cast of an out-of-range value is undefined. */
MCFL_HD uint32_t mcfl_f2u_rz(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t r;
  asm("v_cvt_u32_f32 %0, %1" : "=v"(r) : "v"(f));
  return r;
#else
  return (f >= 4294967296.0f) ? 0xFFFFFFFFu : ((f > 0.0f) ? (uint32_t)f : 0u);
#endif
}
MCFL_HD uint32_t mcfl_f2u_ru(float f) { return mcfl_f2u_rz(mcfl_ceil(f)); }
MCFL_HD uint32_t mcfl_f2u_rn(float f) { return mcfl_f2u_rz(mcfl_rint(f)); }

/* double -> uint64 toward zero with the same saturation rules. */
MCFL_HD uint64_t mcfl_d2u64_rz(double d) {
  return (d >= 18446744073709551616.0) ? 0xFFFFFFFFFFFFFFFFull : ((d > 0.0) ? (uint64_t)d : 0ull);
}

/* Light speed in tissue units, cm per fs (CUDAMCFL.h:29). */
#define MCFL_CCMFS 2.99792E-5

/* q / MCFL_CCMFS for a float q (the time-of-flight increments, CUDAMCFLtransport.cu:316):
   the product with the rounded reciprocal and one FMA correction of its residual
   (Markstein's correction), three double operations instead of a division.  It
   equals the correctly rounded quotient for every float q; the tests check all
   2^31 non-negative finite floats (the formula is odd in q). */
MCFL_HD double mcfl_tof_div(double q) {
  const double r = 1.0 / MCFL_CCMFS;
  const double y = q * r;
  return mcfl_fmad(mcfl_fmad(-y, MCFL_CCMFS, q), r, y);
}

/* The time-of-flight increment of a float q = step * n: the quotient q / c
   truncated to uint64 with mcfl_d2u64_rz's rules.  Below 2^32 fs (steps
   shorter than 1.29e5 cm) that is one double -> uint32 conversion; NaN and
   larger values take the general conversion. */
MCFL_HD uint64_t mcfl_tof_inc(float q) {
  const double t = mcfl_tof_div((double)q);
  if (t < 4294967296.0) return t > 0.0 ? (uint64_t)(uint32_t)t : 0ull;
  return mcfl_d2u64_rz(t);
}

#endif /* MCFL_FMATH_H */
