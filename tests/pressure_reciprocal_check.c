/* Checks the pressure quotient of the collide kernel's contact-list mode:
 * len / den, den = 4*pi*r*r in double, computed as q = RN(len*y) with
 * y = RN(1/den), then twice q <- RN(q + RN(len - q*den)*y) by fused
 * multiply-adds,
 * against the correctly rounded division, on `count` random (len, r) pairs:
 * len any finite float (zero, subnormal, normal up to FLT_MAX), r a float
 * from 2^-40 to 2^40 (log-uniform), plus the extreme radii.
 * With an argument "faithful" it instead counts the inputs whose first
 * quotient RN(len*y) is not faithful (lies beyond the two floats that
 * bracket len/den), the case that the second correction exists for.
 * Prints the number of mismatches.  Build: cc -O2 -ffp-contract=off -lm. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t s = 0x9E3779B97F4A7C15ull;
static uint64_t next(void) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static float any_finite_float(void) {
  for (;;) {
    uint32_t b = (uint32_t)next() & 0x7FFFFFFFu;
    float f;
    memcpy(&f, &b, 4);
    if (isfinite(f)) return f;
  }
}

static double quotient(float len, double den, double y) {
  const double a = (double)len;
  double q = a * y;
  const double r = fma(-q, den, a);
  if (!isfinite(r)) return a / den;
  q = fma(r, y, q);
  return fma(fma(-q, den, a), y, q);
}

/* Whether q is one of the two floats bracketing len/den, given the correctly
 * rounded quotient `want` (whose remainder len - want*den is exact). */
static int faithful(double q, float len, double den, double want) {
  const double rem = fma(-want, den, (double)len);
  if (q == want) return 1;
  if (rem > 0) return q == nextafter(want, INFINITY);
  if (rem < 0) return q == nextafter(want, -INFINITY);
  return 0;
}

int main(int argc, char** argv) {
  const long count = argc > 1 ? atol(argv[1]) : 1000000;
  const int count_unfaithful = argc > 2 && strcmp(argv[2], "faithful") == 0;
  long bad = 0;
  for (long k = 0; k < count; ++k) {
    float r;
    if (k % 1000 == 0) r = (k / 1000) % 2 ? 1.17549435e-38f : 3.40282347e38f;
    else r = (float)ldexp(1.0 + (double)(next() >> 11) * 0x1p-53, (int)(next() % 81) - 40);
    const float len = (k % 7 == 0) ? 0.0f : any_finite_float();
    const double den = 4 * M_PI * (double)r * (double)r;
    const double y = 1.0 / den;
    const double want = (double)len / den;
    if (count_unfaithful) {
      bad += len != 0.0f && !faithful((double)len * y, len, den, want);
      continue;
    }
    const double got = quotient(len, den, y);
    if (memcmp(&want, &got, 8) != 0) {
      if (bad < 5) printf("mismatch len=%a r=%a: %a vs %a\n", len, r, got, want);
      ++bad;
    }
  }
  printf("%ld\n", bad);
  return 0;
}
