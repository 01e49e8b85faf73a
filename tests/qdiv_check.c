/* Check, against IEEE binary32 division, the division sequence of the FP32
 * fitting kernel (bm_qdiv in brian_amd/csrc/fit.hip): y = 1/d rounded once,
 * q = a*y, then twice q += (a - q*d)*y with fused multiply-adds.  Every
 * significand of a is tried at the exponents given on the command line,
 * with both signs, for each divisor; the program prints the number of
 * quotients whose bits differ and exits non-zero if any does.
 *   qdiv_check <divisor bits, hex, comma-separated> <exponents, comma-separated> */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float from_bits(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t to_bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

static float qdiv(float a, float d, float y) {
  float q = a * y;
  float r = fmaf(-q, d, a);
  q = fmaf(r, y, q);
  r = fmaf(-q, d, a);
  return fmaf(r, y, q);
}

int main(int argc, char **argv) {
  if (argc != 3) return 2;
  uint32_t ds[64];
  int exps[64], nd = 0, ne = 0;
  for (char *t = strtok(argv[1], ","); t && nd < 64; t = strtok(NULL, ",")) ds[nd++] = (uint32_t)strtoul(t, NULL, 16);
  for (char *t = strtok(argv[2], ","); t && ne < 64; t = strtok(NULL, ",")) exps[ne++] = atoi(t);
  long long bad = 0, n = 0;
  for (int k = 0; k < nd; ++k) {
    const float d = from_bits(ds[k]);
    const float y = 1.0f / d;
    for (int e = 0; e < ne; ++e)
      for (uint32_t m = 0; m < (1u << 23); ++m)
        for (int sg = 0; sg < 2; ++sg) {
          const float a0 = ldexpf(from_bits(0x3f800000u | m), exps[e]);
          const float a = sg ? -a0 : a0;
          ++n;
          if (to_bits(a / d) != to_bits(qdiv(a, d, y))) {
            if (bad < 5) printf("d=%a a=%a: %a != %a\n", d, a, a / d, qdiv(a, d, y));
            ++bad;
          }
        }
  }
  printf("%lld of %lld quotients differ\n", bad, n);
  return bad != 0;
}
