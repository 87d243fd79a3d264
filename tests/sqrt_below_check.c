/* Checks the shear law's Coulomb test as the step kernels decide it without a
 * square root where they can: RN(sqrt(x)) < m from x against RN(m*m)*(1 -/+ 2^-20)
 * for m in [2^-63, 2^63], the correctly rounded root otherwise.  Compared with
 * sqrtf(x) < m on `count` random m (log-uniform over 2^-70..2^70, plus zero,
 * subnormal, FLT_MAX, infinity and NaN) and x drawn around m*m: every float
 * within 64 ulps of m*m, of the square of the midpoint below m and of the two
 * thresholds, and random x.  Prints the number of mismatches.
 * Build: cc -O2 -ffp-contract=off -lm. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t s = 0xD1B54A32D192ED03ull;
static uint64_t next(void) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static int fast(float x, float m) {
  if (m >= 0x1p-63f && m <= 0x1p63f) {
    const float t = m * m;
    if (x < t * (1.0f - 0x1p-20f)) return 1;
    if (x > t * (1.0f + 0x1p-20f)) return 0;
  }
  return sqrtf(x) < m;
}

static long bad = 0, checked = 0;
static void check(float x, float m) {
  if (!(x >= 0.0f) && !isnan(x)) return; /* the kernels pass sums of squares */
  ++checked;
  if (fast(x, m) != (sqrtf(x) < m)) {
    if (bad < 5) printf("mismatch x=%a m=%a\n", x, m);
    ++bad;
  }
}

static void around(float c, float m) {
  float lo = c, hi = c;
  check(c, m);
  for (int k = 0; k < 64; ++k) {
    lo = nextafterf(lo, 0.0f);
    hi = nextafterf(hi, INFINITY);
    check(lo, m);
    check(hi, m);
  }
}

int main(int argc, char** argv) {
  const long count = argc > 1 ? atol(argv[1]) : 100000;
  const float special[] = {0.0f, 0x1p-149f, 0x1p-126f, 0x1p-63f, 0x1p63f, 3.40282347e38f, INFINITY, NAN};
  for (long k = 0; k < count; ++k) {
    float m;
    if (k < (long)(sizeof special / sizeof special[0])) m = special[k];
    else m = (float)ldexp(1.0 + (double)(next() >> 11) * 0x1p-53, (int)(next() % 141) - 70);
    const float t = m * m;
    const float below = (float)(((double)m + (double)nextafterf(m, 0.0f)) / 2);
    around(t, m);
    around((float)((double)below * (double)below), m);
    around(t * (1.0f - 0x1p-20f), m);
    around(t * (1.0f + 0x1p-20f), m);
    for (int j = 0; j < 16; ++j) {
      uint32_t b = (uint32_t)next() & 0x7FFFFFFFu;
      float x;
      memcpy(&x, &b, 4);
      check(x, m);
    }
    check(NAN, m);
    check(INFINITY, m);
  }
  printf("%ld of %ld\n%ld\n", bad, checked, bad);
  return 0;
}
