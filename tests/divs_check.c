/* The shared-reciprocal division of the kernels (Math<false>::divs in hush_amd/csrc/hush_device.h): with
   y = RN(1/b), q0 = RN(a y) and two FMA corrections q' = RN(q + RN(a - b q) y), the result must be the
   correctly rounded a / b.  Checks random operands over wide exponent ranges and every small-integer
   quotient against IEEE division; prints the mismatch count. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float divs(float a, float b, float y) {
  float q0 = a * y;
  float q1 = fmaf(fmaf(-b, q0, a), y, q0);
  return fmaf(fmaf(-b, q1, a), y, q1);
}
static uint64_t st = 88172645463325252ull;
static uint64_t rnd(void) { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return st; }
static float rf(int emin, int emax) {
  uint32_t u = ((uint32_t)(emin + (int)(rnd() % (uint64_t)(emax - emin + 1)) + 127) << 23) | (uint32_t)(rnd() & 0x7fffff);
  float f;
  memcpy(&f, &u, 4);
  return (rnd() & 1) ? -f : f;
}
int main(int argc, char **argv) {
  long n = argc > 1 ? atol(argv[1]) : 10000000, bad = 0;
  for (long i = 0; i < n; i++) {
    float a = rf(-60, 60), b = rf(-40, 40), q = divs(a, b, 1.0f / b), r = a / b;
    bad += memcmp(&q, &r, 4) != 0;
  }
  for (int i = 0; i < 4096; i++)
    for (int j = 1; j < 64; j++) {
      float a = (float)i, b = (float)j, q = divs(a, b, 1.0f / b), r = a / b;
      bad += memcmp(&q, &r, 4) != 0;
    }
  printf("%ld\n", bad);
  return 0;
}
