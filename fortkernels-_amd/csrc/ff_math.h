// ff_math.h — natural logarithm for the FF frequency-grid kernel.
//
// The interval terms of ffKernel.f90 need log|(wx - right)/(wx - left)| for
// consecutive grid points, clamped to [log 1e-4, log 1e4]; the kernel forms
// it as the difference of log|wx - zb[f]| at adjacent points (one logarithm
// per point, no division) and clamps the difference.  The logarithm of the
// device math library spends about 100 FP64 instructions on double-double
// corrections; log_abs below takes 17 and one table read (from LDS on the
// device), with an absolute error below 1.5 ulp of max(|log x|, ln 2),
// which is what a difference of two logarithms needs.
//
// Method (table-driven, Tang, ACM TOMS 16(4), 1990): |x| = m 2^k with m in
// [0.5, 1) (frexp, exact for subnormals too); the top 8 fraction bits of m
// pick {invc, logc} (ff_log_table.h), r = m invc - 1 (one fma, |r| < 2^-9),
//   log|x| = (k ln2_hi + logc) + (k ln2_lo + r + r^2 Q(r)),
// ln2_hi having 32 significant bits so that k ln2_hi is exact, Q the cubic
// of ff_log_table.h.
//
// Compiled for the device (hipcc) and for the host (g++, to check it).
#pragma once

#include <cmath>
#include <cstdint>

#include "ff_log_table.h"

#if defined(__HIPCC__)
#define FF_HD __host__ __device__
#else
#define FF_HD
#endif


namespace gpp {
namespace ff {

// log|x| for x != 0: +inf at +-inf, NaN for NaN, a finite value of no
// meaning at 0.  `tab` is kLogTable (or a copy of it).
FF_HD inline double log_abs_nonzero(double x, const double *tab) {
  constexpr double kLn2Hi = 0x1.62e42fee00000p-1;   // ln 2, 32 bits
  constexpr double kLn2Lo = 0x1.a39ef35793c76p-33;  // ln 2 - kLn2Hi
  int k;
  // +-inf: m = inf gives r = inf and a result of +inf; NaN stays NaN
  const double m = frexp(fabs(x), &k);  // |x| = m 2^k, m in [0.5, 1)
  uint64_t bits;
  __builtin_memcpy(&bits, &m, sizeof bits);
  // byte offset of entry i = the top kLogTableBits fraction bits of m
  const unsigned off = (unsigned)(bits >> (52 - kLogTableBits - 4)) &
                       (((1u << kLogTableBits) - 1) << 4);
  const double *e = (const double *)((const char *)tab + off);
  const double invc = e[0], logc = e[1];
  const double r = fma(m, invc, -1.0);
  double q = fma(kLogQ3, r, kLogQ2);
  q = fma(q, r, kLogQ1);
  q = fma(q, r, kLogQ0);
  const double y = fma(q, r * r, r);  // log1p(r)
  const double dk = (double)k;
  return fma(dk, kLn2Hi, logc) + fma(dk, kLn2Lo, y);
}

// log|x| for every x: -inf at 0, +inf at +-inf, NaN for NaN.
FF_HD inline double log_abs(double x, const double *tab) {
  const double l = log_abs_nonzero(x, tab);
  return x == 0.0 ? -INFINITY : l;
}

// log of a ratio clamped to [1e-4, 1e4] (ffKernel.f90:393-396, 399-402),
// taken as the difference t = log|a| - log|b| clamped to [log(1e-4),
// log(1e4)] (the double-rounded logarithms of the bounds as the reference
// computes them); a NaN difference (0/0, inf/inf) passes through as the
// reference's NaN ratio does.
constexpr double kLogRatioLo = -0x1.26bb1bbb55515p+3;  // log(1e-4)
constexpr double kLogRatioHi = 0x1.26bb1bbb55516p+3;   // log(1e4)
FF_HD inline double clamp_log_ratio(double t) {
  if (t < kLogRatioLo) t = kLogRatioLo;
  if (t > kLogRatioHi) t = kLogRatioHi;
  return t;
}

// The interval logarithms of consecutive grid points, log|a_f / a_{f-1}|
// clamped as above, a_f = wx - zb[f] (or wx + zb[f]): one logarithm per
// point, carried to the next interval.  Most differences lie inside the
// bounds and neither end is zero, and for them the step is one compare
// (two with the zero test); the rest take the exact form: log 0 = -inf
// (so x/0 and 0/x clamp to the bounds and 0/0 is NaN), the clamp, NaN
// passed through.
struct IntervalLog {
  double l;  // log|a_{f-1}| (log_abs_nonzero)
#if defined(__HIP_DEVICE_COMPILE__)
  // The zero tests are kept as wave lane masks (scalar registers): one
  // compare per point and no per-lane mask bookkeeping.
  uint64_t zm;  // lanes with a_{f-1} == 0
  __device__ void start(double a, const double *tab) {
    l = log_abs_nonzero(a, tab);
    zm = __builtin_amdgcn_ballot_w64(a == 0.0);
  }
  __device__ double step(double a, const double *tab) {
    const double lr = log_abs_nonzero(a, tab);
    double t = lr - l;
    const uint64_t zr = __builtin_amdgcn_ballot_w64(a == 0.0);
    // -kLogRatioLo < kLogRatioHi: |t| <= -kLogRatioLo is inside both bounds
    const uint64_t ex =
        __builtin_amdgcn_ballot_w64(!(fabs(t) <= -kLogRatioLo)) | zr | zm;
    if (ex != 0) {
      const uint64_t bit = 1ull << __lane_id();
      const double te = clamp_log_ratio((zr & bit ? -INFINITY : lr) -
                                        (zm & bit ? -INFINITY : l));
      t = ex & bit ? te : t;
    }
    l = lr;
    zm = zr;
    return t;
  }
#else
  bool z;  // a_{f-1} == 0
  void start(double a, const double *tab) {
    l = log_abs_nonzero(a, tab);
    z = a == 0.0;
  }
  double step(double a, const double *tab) {
    const double lr = log_abs_nonzero(a, tab);
    const bool zr = a == 0.0;
    double t = lr - l;
    if (!(fabs(t) <= -kLogRatioLo) || zr || z)
      t = clamp_log_ratio((zr ? -INFINITY : lr) - (z ? -INFINITY : l));
    l = lr;
    z = zr;
    return t;
  }
#endif
};

}  // namespace ff
}  // namespace gpp
