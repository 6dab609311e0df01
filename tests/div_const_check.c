/* Check of the constant-divisor division used by the HIP steppers (div_c in
 * meep-torch_amd/csrc/common.h): for y = RN(1/b), q0 = RN(a*y),
 * r = fma(-q0, b, a), q1 = fma(r, y, q0) must equal the IEEE quotient a / b
 * bit for bit.  Every float significand (2^23) is tried for a spread of
 * exponents across [2^-100, 2^100], both signs, and each divisor given on
 * the command line; prints the mismatch count and exits non-zero on any. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int main(int argc, char** argv) {
  unsigned long long bad = 0, tried = 0;
  for (int bi = 1; bi < argc; ++bi) {
    volatile float b = strtof(argv[bi], NULL);
    float y = 1.0f / b;
    for (int e = -100; e <= 99; e += 9) {
      for (uint32_t m = 0; m < (1u << 23); ++m) {
        for (int sg = 0; sg < 2; ++sg) {
          uint32_t w = ((uint32_t)sg << 31) | ((uint32_t)(e + 127) << 23) | m;
          float a;
          memcpy(&a, &w, 4);
          volatile float va = a;
          float q = va / b;
          float q0 = a * y;
          float r = fmaf(-q0, b, a);
          float q1 = fmaf(r, y, q0);
          uint32_t x1, x2;
          memcpy(&x1, &q, 4);
          memcpy(&x2, &q1, 4);
          ++tried;
          if (x1 != x2) {
            if (bad < 5) printf("mismatch b=%a a=%a q=%a q1=%a\n", (double)b, (double)a, (double)q, (double)q1);
            ++bad;
          }
        }
      }
    }
  }
  /* below 2^-100 the sequence is not exact for every dividend (the kernels
     send those to true division): b = 0.05 fails at a = 0x1.999a78p-128 */
  {
    volatile float b = 0.05f, a = 0x1.999a78p-128f;
    float y = 1.0f / b, q = a / b, q0 = a * y, r = fmaf(-q0, b, a), q1 = fmaf(r, y, q0);
    if (memcmp(&q, &q1, 4) == 0) {
      printf("expected a mismatch below 2^-100\n");
      return 2;
    }
  }
  printf("tried %llu, mismatches %llu\n", tried, bad);
  return bad != 0;
}
