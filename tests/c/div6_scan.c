/* Exhaustive check of the float division by 6 used by the device QUICK scheme: a reciprocal
 * multiply with one FMA correction (Markstein), kept when the quotient is a normal number and
 * replaced by the IEEE division otherwise.  Prints the number of the 2^32 bit patterns whose
 * result differs in any bit from x/6 (NaN results compare equal whatever their payload).
 * Built with -mfma -ffp-contract=off so that fmaf is the single-rounding hardware FMA. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static float div6_device_form(float x) {
  const float r = 0.16666667f;
  const float q0 = x * r;
  const float q = fmaf(fmaf(-q0, 6.0f, x), r, q0);
  return fpclassify(q) == FP_NORMAL ? q : x / 6.0f;
}

int main(void) {
  unsigned long long bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
  for (long long b = 0; b < (1ll << 32); ++b) {
    const uint32_t u = (uint32_t)b;
    float x;
    memcpy(&x, &u, 4);
    const float want = x / 6.0f, got = div6_device_form(x);
    if (isnan(want) && isnan(got)) continue;
    uint32_t uw, ug;
    memcpy(&uw, &want, 4);
    memcpy(&ug, &got, 4);
    if (uw != ug) ++bad;
  }
  printf("%llu\n", bad);
  return 0;
}
