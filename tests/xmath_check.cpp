// Compares cress::xexpf / cress::xlogf (cress_amd/csrc/xmath.h, host
// build) with glibc's expf/logf and with the binary64 functions rounded to
// binary32, over every stride-th binary32 argument of a range.
//   xmath_check exp|log stride
// prints: count, mismatches vs glibc, mismatches vs binary64-rounded, max ulp
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "xmath.h"

static float bits(unsigned u) { float f; memcpy(&f, &u, 4); return f; }
static long ulpdiff(float a, float b) {
  int ia, ib;
  memcpy(&ia, &a, 4);
  memcpy(&ib, &b, 4);
  if (ia < 0) ia = (int)0x80000000 - ia;
  if (ib < 0) ib = (int)0x80000000 - ib;
  return labs((long)ia - (long)ib);
}

int main(int argc, char **argv) {
  if (argc < 3) return 2;
  const int isexp = strcmp(argv[1], "exp") == 0;
  const unsigned stride = (unsigned)atoi(argv[2]);
  long n = 0, mg = 0, md = 0, mx = 0;
  // exp: x in [-104, 88.8] (both signs); log: every positive finite float
  for (unsigned long u = 0; u < 0x100000000ul; u += stride) {
    const float x = bits((unsigned)u);
    if (isnan(x) || isinf(x)) continue;
    float got, g, d;
    if (isexp) {
      if (!(x >= -104.0f && x <= 88.8f)) continue;
      got = cress::xexpf(x);
      g = expf(x);
      d = (float)exp((double)x);
    } else {
      if (!(x > 0.0f)) continue;
      got = cress::xlogf(x);
      g = logf(x);
      d = (float)log((double)x);
    }
    n++;
    if (memcmp(&got, &g, 4) != 0) mg++;
    if (memcmp(&got, &d, 4) != 0) md++;
    const long ud = ulpdiff(got, g);
    if (ud > mx) mx = ud;
  }
  printf("%ld %ld %ld %ld\n", n, mg, md, mx);
  return 0;
}
