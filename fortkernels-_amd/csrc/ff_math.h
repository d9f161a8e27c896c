// ff_math.h — natural logarithm for the FF frequency-grid kernel.
//
// The grid terms of the FF Coulomb-hole sums take log|ratio| with the ratio
// clamped to [1e-4, 1e4].  The logarithm of the device math library spends
// about 100 FP64 instructions on double-double corrections; this one uses
// about 25, at most 1.3 ulp from the exact value (0.5 ulp near 1), checked
// against long double logl over the clamped range.
//
// Method (the classical reduction, as in fdlibm's e_log.c): x = 2^k (1+f)
// with sqrt(2)/2 <= 1+f < sqrt(2), s = f/(2+f), z = s^2 <= (3 - 2 sqrt 2)^2,
//   log(1+f) = 2 atanh(s) = f - f^2/2 + s (f^2/2 + R),
//   R = z P(z),  P(z) = sum_{j>=0} 2 z^j / (2j+3),
// log x = k ln2_hi + (log(1+f) + k ln2_lo), ln2_hi having 32 significant
// bits so that k ln2_hi is exact.  P is replaced by the degree-6 polynomial
// Q obtained by Chebyshev economization of the series on [0, 0.0295]
// (computed in exact rational arithmetic, then rounded to double):
// |P - Q| <= 3.2e-16 there, below 5e-18 of the logarithm.
//
// Compiled for the device (hipcc) and for the host (g++, to check it).
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define FF_HD __host__ __device__
#else
#define FF_HD
#endif

namespace gpp {
namespace ff {

// 1/d for d in [1.5, 2.5]: the hardware reciprocal refined by one Newton
// step on the device, the IEEE quotient on the host.
FF_HD inline double log_recip(double d) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double r = __builtin_amdgcn_rcp(d);
  return fma(fma(-d, r, 1.0), r, r);
#else
  return 1.0 / d;
#endif
}

// log(x) for positive normal x; NaN in gives NaN out.
FF_HD inline double log_pos(double x) {
  constexpr double kLn2Hi = 0x1.62e42fee00000p-1;   // ln 2, 32 bits
  constexpr double kLn2Lo = 0x1.a39ef35793c76p-33;  // ln 2 - kLn2Hi
  constexpr double kSqrtHalf = 0.70710678118654752440;
  int k;
  double m = frexp(x, &k);  // x = m 2^k, m in [0.5, 1)
  if (m < kSqrtHalf) {
    m += m;
    --k;
  }
  const double f = m - 1.0;  // exact
  const double s = f * log_recip(2.0 + f);
  const double z = s * s;
  double R = 0x1.2b6ada4e6cac0p-3;
  R = fma(R, z, 0x1.39fcd91739d16p-3);
  R = fma(R, z, 0x1.7462c0aff40b1p-3);
  R = fma(R, z, 0x1.c71c62ba3869bp-3);
  R = fma(R, z, 0x1.2492492e143e6p-2);
  R = fma(R, z, 0x1.999999999519ap-2);
  R = fma(R, z, 0x1.5555555555558p-1);
  R *= z;
  const double hfsq = 0.5 * f * f;
  const double r = f - (hfsq - s * (hfsq + R));
  const double dk = (double)k;
  return fma(dk, kLn2Hi, fma(dk, kLn2Lo, r));
}

}  // namespace ff
}  // namespace gpp
