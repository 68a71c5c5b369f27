/* Exhaustive check of the kernels' voxel-coordinate division (div_rn in the HIP source):
 * q = RN(a y), r = fma(-q, b, a), q' = fma(r, y, q) with y = RN(1 / b), against IEEE single
 * division a / b, for every float a in [amin, amax) and the given divisors b.  Prints the
 * number of mismatches per divisor and the first mismatching a.  Negative a follow by
 * symmetry of round-to-nearest (every step changes sign with a); a = 0 gives +0.
 * Build: gcc -O2 -ffp-contract=off -fopenmp div_rn_check.c -lm
 * Run:   ./a.out amin amax b [b ...] */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float from_bits(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}
static uint32_t to_bits(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}

int main(int argc, char** argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s amin amax b [b ...]\n", argv[0]);
    return 2;
  }
  const uint32_t begin = to_bits(strtof(argv[1], NULL));
  const uint32_t end = to_bits(strtof(argv[2], NULL));
  int bad_total = 0;
  for (int k = 3; k < argc; ++k) {
    const float b = strtof(argv[k], NULL);
    const volatile float one = 1.0f;
    const float y = one / b;
    long long bad = 0;
    uint32_t first = 0xffffffffu;
#pragma omp parallel for reduction(+ : bad) schedule(static, 1 << 20)
    for (long long i = begin; i < (long long)end; ++i) {
      const float a = from_bits((uint32_t)i);
      const float q = a * y;
      const float r = fmaf(-q, b, a);
      const float qr = fmaf(r, y, q);
      const volatile float bv = b;
      const float ref = a / bv;
      if (to_bits(qr) != to_bits(ref)) {
        ++bad;
#pragma omp critical
        if ((uint32_t)i < first) first = (uint32_t)i;
      }
    }
    printf("%.9g %lld %u\n", (double)b, bad, first);
    bad_total += bad != 0;
  }
  return bad_total ? 1 : 0;
}
