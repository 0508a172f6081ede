/* Compares the splat kernel's quotient r2 / radius2 formed as q0 = r2 y, q = q0 + (r2 - radius2 q0) y
 * (fused, y = the correctly rounded 1 / radius2) with the IEEE quotient, over the dividends in
 * [2^-100, radius2) taken with a stride, for the divisors given on the command line as hex bits.
 * Prints the number of differences. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float fb(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t bf(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

int main(int argc, char** argv) {
  long bad = 0, total = 0;
  const uint32_t stride = (uint32_t)strtoul(argv[1], NULL, 0);
  for (int k = 2; k < argc; k++) {
    const float b = fb((uint32_t)strtoul(argv[k], NULL, 16)), y = 1.0f / b;
    for (uint32_t u = bf(0x1p-100f); u < bf(b); u += stride) {
      const float a = fb(u), q0 = a * y;
      const float q = fmaf(fmaf(-q0, b, a), y, q0);
      bad += bf(q) != bf(a / b);
      total++;
    }
  }
  printf("%ld %ld\n", total, bad);
  return 0;
}
