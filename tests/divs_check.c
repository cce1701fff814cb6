/* The shared-reciprocal division of the kernels (Math<false>::divs in hush_amd/csrc/hush_device.h): with
   y = RN(1/b), q0 = RN(a y) and one FMA correction q = RN(q0 + RN(a - b q0) y) (the residual a - b q0 is
   exact in an FMA), the result must be the correctly rounded a / b (Markstein's one-step correction).  The
   kernels end with v_div_fixup_f32, which for finite non-zero b and finite a returns the correction with the
   sign of a b; here that is copysign(q, q0), q0 carrying the sign of a b (a zero quotient's sign can be lost
   by the correction's addition of two zeros).  The fixup's special cases (zero, infinite or NaN operands) are
   checked on the GPU.  Checks, against IEEE division (bit for bit, so -0 != +0):
     * random operands over wide exponent ranges (argv[1] pairs);
     * every quotient of small integers, and +-0 over divisors of both signs;
     * every numerator significand (2^23) at three exponents for divisors with the hard significands (all
       ones, all ones minus 1 and 2, one ulp above 1, exactly 1, 1.5, 1.5 minus one ulp) at three exponents,
       and for argv[2] random divisors (one numerator exponent each);
     * numerators of 2^-102 .. 2^-80 over divisors of every normal exponent (argv[1] pairs), normal quotients.
   Prints the mismatch count. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float divs(float a, float b, float y) {
  float q0 = a * y;
  return copysignf(fmaf(fmaf(-b, q0, a), y, q0), q0);
}
static uint64_t st = 88172645463325252ull;
static uint64_t rnd(void) { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return st; }
static float fb(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static float rf(int emin, int emax) {
  uint32_t u = ((uint32_t)(emin + (int)(rnd() % (uint64_t)(emax - emin + 1)) + 127) << 23) | (uint32_t)(rnd() & 0x7fffff);
  float f = fb(u);
  return (rnd() & 1) ? -f : f;
}
static long sweep(float b, int aexp) {
  long bad = 0;
  const float y = 1.0f / b;
  for (uint32_t m = 0; m < (1u << 23); m++) {
    const float a = fb(((uint32_t)(aexp + 127) << 23) | m), q = divs(a, b, y), r = a / b;
    bad += memcmp(&q, &r, 4) != 0;
  }
  return bad;
}
int main(int argc, char **argv) {
  long n = argc > 1 ? atol(argv[1]) : 10000000, nb = argc > 2 ? atol(argv[2]) : 16, bad = 0;
  for (long i = 0; i < n; i++) {
    float a = rf(-60, 60), b = rf(-40, 40), q = divs(a, b, 1.0f / b), r = a / b;
    bad += memcmp(&q, &r, 4) != 0;
  }
  for (int i = 0; i < 4096; i++)
    for (int j = 1; j < 64; j++) {
      float a = (float)i, b = (float)j, q = divs(a, b, 1.0f / b), r = a / b;
      bad += memcmp(&q, &r, 4) != 0;
    }
  for (int i = 0; i < 100000; i++) {
    const float z = (i & 1) ? -0.0f : 0.0f, b = rf(-40, 40), q = divs(z, b, 1.0f / b), r = z / b;
    bad += memcmp(&q, &r, 4) != 0;
  }
  const uint32_t hard[] = {0x7fffff, 0x7ffffe, 0x7ffffd, 0x000001, 0x000000, 0x400000, 0x3fffff};
  for (int s = 0; s < 7; s++)
    for (int be = -3; be <= 3; be += 3)
      for (int ae = -1; ae <= 1; ae++) bad += sweep(fb(((uint32_t)(be + 127) << 23) | hard[s]), ae);
  for (long k = 0; k < nb; k++)
    bad += sweep(fb(((uint32_t)((int)(rnd() % 40) - 20 + 127) << 23) | (uint32_t)(rnd() & 0x7fffff)), (int)(rnd() % 5) - 2);
  /* the low end of the domain: numerators down to 2^-102 (below it the residual a - b q0 can fall under
     the subnormal grid and the correction can miss by an ulp) over divisors of any normal exponent, every
     normal quotient counted */
  for (long i = 0; i < n; i++) {
    const float a = rf(-102, -80), b = rf(-126, 125), r = a / b;
    if (!isnormal(r)) continue;
    const float q = divs(a, b, 1.0f / b);
    bad += memcmp(&q, &r, 4) != 0;
  }
  printf("%ld\n", bad);
  return 0;
}
