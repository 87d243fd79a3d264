// scene_rng.h -- the random draws of the scene generators, shared by the
// gfx950 kernels (scene.hip) and the host driver (driver/prepare.cpp).
//
// The draws use only correctly rounded IEEE double operations (+ - * /,
// sqrt, frexp/ldexp scalings), so the device, the host and the numpy
// restatement in demso_amd/scene.py (jitter_normal) give the same bits:
//   uniform01(k, s, i) = top 53 bits of SplitMix64 (Steele, Lea & Flood
//                        2014) at counter k*G + s*H + (i+1)*G, times 2^-53
//   jitter_normal      = Box-Muller sqrt(-2 ln(1-u1)) cos(2 pi u2), u1 from
//                        stream s, u2 from stream s + 1000003, with ln from
//                        the atanh series and cos from its Taylor series
//                        after an exact reduction to [0, 1/8] turns.
// Compile without FMA contraction (-ffp-contract=off) on both sides.
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define DEM_RNG_FN __host__ __device__ inline
#else
#define DEM_RNG_FN inline
#endif

namespace dem_rng {

constexpr uint64_t kG = 0x9E3779B97F4A7C15ull;
constexpr uint64_t kH = 0xD1B54A32D192ED03ull;
constexpr uint64_t kNormalPair = 1000003ull;  // stream offset of Box-Muller's second uniform

DEM_RNG_FN uint64_t mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

DEM_RNG_FN double uniform01(uint64_t seed, uint64_t stream, uint64_t i) {
  const uint64_t z = seed * kG + stream * kH + (i + 1) * kG;
  return (double)(mix(z) >> 11) * 0x1.0p-53;
}

// ln x for x in [2^-53, 1]: x = f 2^e with f in [sqrt(1/2), sqrt(2)), then
// ln f = 2 atanh(t), t = (f-1)/(f+1), |t| <= 0.1716, series to t^21, and
// e ln 2 in two parts (fdlibm's ln2_hi / ln2_lo split).
DEM_RNG_FN double series_log(double x) {
  int e = 0;
  double f = std::frexp(x, &e);  // f in [0.5, 1)
  if (f < 0.70710678118654752440) {
    f = f * 2.0;
    e = e - 1;
  }
  const double t = (f - 1.0) / (f + 1.0);
  const double t2 = t * t;
  double p = 1.0 / 21.0;
  p = p * t2 + 1.0 / 19.0;
  p = p * t2 + 1.0 / 17.0;
  p = p * t2 + 1.0 / 15.0;
  p = p * t2 + 1.0 / 13.0;
  p = p * t2 + 1.0 / 11.0;
  p = p * t2 + 1.0 / 9.0;
  p = p * t2 + 1.0 / 7.0;
  p = p * t2 + 1.0 / 5.0;
  p = p * t2 + 1.0 / 3.0;
  p = p * t2 + 1.0;
  const double lnf = 2.0 * t * p;
  const double de = (double)e;
  return de * 6.93147180369123816490e-01 + (de * 1.90821492927058770002e-10 + lnf);
}

// cos(2 pi u) for u in [0, 1).  The reductions u -> 1-u, 1/2-u, 1/4-u are
// exact (Sterbenz), leaving v in [0, 1/8] and x = 2 pi v <= pi/4.
DEM_RNG_FN double series_cos2pi(double u) {
  if (u >= 0.5) u = 1.0 - u;
  double sign = 1.0;
  if (u > 0.25) {
    u = 0.5 - u;
    sign = -1.0;
  }
  const bool use_sin = u > 0.125;
  const double v = use_sin ? 0.25 - u : u;
  const double x = 6.28318530717958647692 * v;
  const double x2 = x * x;
  double r;
  if (use_sin) {
    // x - x^3/3! + ... - x^19/19!
    double p = -1.0 / 121645100408832000.0;
    p = p * x2 + 1.0 / 355687428096000.0;
    p = p * x2 - 1.0 / 1307674368000.0;
    p = p * x2 + 1.0 / 6227020800.0;
    p = p * x2 - 1.0 / 39916800.0;
    p = p * x2 + 1.0 / 362880.0;
    p = p * x2 - 1.0 / 5040.0;
    p = p * x2 + 1.0 / 120.0;
    p = p * x2 - 1.0 / 6.0;
    p = p * x2 + 1.0;
    r = x * p;
  } else {
    // 1 - x^2/2! + ... + x^18/18!
    double p = 1.0 / 6402373705728000.0;
    p = p * x2 - 1.0 / 20922789888000.0;
    p = p * x2 + 1.0 / 87178291200.0;
    p = p * x2 - 1.0 / 479001600.0;
    p = p * x2 + 1.0 / 3628800.0;
    p = p * x2 - 1.0 / 40320.0;
    p = p * x2 + 1.0 / 720.0;
    p = p * x2 - 1.0 / 24.0;
    p = p * x2 + 1.0 / 2.0;
    r = 1.0 - x2 * p;
  }
  return sign * r;
}

// Standard normal draw i of stream `stream` under `seed`.
DEM_RNG_FN double jitter_normal(uint64_t seed, uint64_t stream, uint64_t i) {
  const double u1 = uniform01(seed, stream, i);
  const double u2 = uniform01(seed, stream + kNormalPair, i);
  return std::sqrt(-2.0 * series_log(1.0 - u1)) * series_cos2pi(u2);
}

}  // namespace dem_rng

#undef DEM_RNG_FN
