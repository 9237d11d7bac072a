/*
 * gen_safeprimes.c -- regenerates safeprimes_base32.txt, the table of
 * multiply-with-carry (MWC) multipliers read by the RNG initialiser.
 *
 * The reference reads this file (CUDAMCFLrng.cu:44-52,69) but does not ship
 * it.  Its content is therefore defined here by a rule:
 *
 *   a qualifies  <=>  p = a*2^32 - 1 is prime  and  (p-1)/2 = a*2^31 - 1 is prime
 *
 * i.e. p is a safe prime.  For such a the MWC generator
 *     s' = (s mod 2^32) * a + floor(s / 2^32)
 * is multiplication by a modulo p on the packed state s = c*2^32 + x, and
 * because 2^32 is a quadratic residue mod p its period is (p-1)/2.
 *
 * Lines are written in DESCENDING order of a, starting at the largest
 * qualifying a < 2^32.  Each line is "a hi32(p) lo32(p)", three unsigned
 * 32-bit decimals: the reader uses only the first column and skips the other
 * two (CUDAMCFLrng.cu:52,69 read "%u %u %u").  The first line seeds the
 * per-lane initialisation generator, the next lines are per-lane multipliers.
 *
 * Primality: deterministic Miller-Rabin for 64-bit integers with the first
 * twelve prime bases (correct for all n < 3.3e24), after trial division.
 *
 * Build and run:  gcc -O2 -fopenmp gen_safeprimes.c -o gen && ./gen 32769 > safeprimes_base32.txt
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

typedef unsigned __int128 u128;

static uint64_t mulmod(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)((u128)a * b % m); }

static uint64_t powmod(uint64_t b, uint64_t e, uint64_t m) {
  uint64_t r = 1 % m;
  b %= m;
  while (e) {
    if (e & 1) r = mulmod(r, b, m);
    b = mulmod(b, b, m);
    e >>= 1;
  }
  return r;
}

static int is_prime_u64(uint64_t n) {
  static const uint32_t small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89, 97};
  if (n < 2) return 0;
  for (unsigned i = 0; i < sizeof small / sizeof small[0]; ++i) {
    if (n == small[i]) return 1;
    if (n % small[i] == 0) return 0;
  }
  uint64_t d = n - 1;
  int s = 0;
  while ((d & 1) == 0) { d >>= 1; ++s; }
  static const uint64_t bases[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
  for (unsigned i = 0; i < sizeof bases / sizeof bases[0]; ++i) {
    uint64_t x = powmod(bases[i], d, n);
    if (x == 1 || x == n - 1) continue;
    int composite = 1;
    for (int r = 1; r < s; ++r) {
      x = mulmod(x, x, n);
      if (x == n - 1) { composite = 0; break; }
    }
    if (composite) return 0;
  }
  return 1;
}

int qualifies(uint32_t a) {
  uint64_t p = ((uint64_t)a << 32) - 1u;
  uint64_t q = ((uint64_t)a << 31) - 1u;
  return is_prime_u64(q) && is_prime_u64(p);
}

int main(int argc, char** argv) {
  long want = argc > 1 ? atol(argv[1]) : 32769;
  const long block = 1 << 16;
  uint32_t top = 0xFFFFFFFFu;
  long have = 0;
  unsigned char* hit = malloc(block);
  while (have < want) {
#pragma omp parallel for schedule(dynamic, 256)
    for (long i = 0; i < block; ++i) hit[i] = qualifies(top - (uint32_t)i);
    for (long i = 0; i < block && have < want; ++i) {
      if (!hit[i]) continue;
      uint32_t a = top - (uint32_t)i;
      uint64_t p = ((uint64_t)a << 32) - 1u;
      printf("%u %u %u\n", a, (uint32_t)(p >> 32), (uint32_t)p);
      ++have;
    }
    top -= (uint32_t)block;
  }
  free(hit);
  return 0;
}
