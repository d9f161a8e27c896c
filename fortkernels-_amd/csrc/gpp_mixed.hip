// gpp_mixed.hip — the unmasked (gppker semantics) mixed-precision
// non-occupied kernel of one GPP sigma evaluation; the kernel and its block
// are in gpp_device_common.h (the bgw-masked variants and the batched launch
// are instantiated by gpp_device.hip).  It is a source of its own so that it
// can be built with the compiler's ILP-first scheduler (the Makefile's
// MIXED_FLAGS): measured on the MI355X at 512/2/32768/20 (band-resolved
// inputs), that schedule runs it in 22.7 ms against 23.6 ms with the
// default one; it slowed the FP64 kernel (32.3-32.5 against 31.8-32.0 ms)
// and the masked mixed one (36.2 against 34.7 ms), which gpp_device.hip
// builds with the default scheduler.
#include "gpp_device_common.h"

namespace gpp {
namespace {

template <int K>
hipError_t launch_k(const DevProblem &p, int n1b, int n1e, bool kahan,
                    int waves, double thresh, double *partials, double *n1log,
                    unsigned long long *fb, int glo, int gcnt, int grid,
                    hipStream_t s) {
  const dim3 blk(kWave * waves);
  const bool lg = n1log != nullptr;
  const float th = (float)thresh;
#define GPP_MIX(L, KA)                                                      \
  unocc_mixed_kernel<K, false, L, KA><<<grid, blk, 0, s>>>(p, n1b, n1e, glo, \
                                                           gcnt, partials,  \
                                                           n1log, th, fb)
  if (kahan) {
    if (lg) GPP_MIX(true, true); else GPP_MIX(false, true);
  } else {
    if (lg) GPP_MIX(true, false); else GPP_MIX(false, false);
  }
#undef GPP_MIX
  return hipGetLastError();
}

}  // namespace

hipError_t launch_unocc_mixed(const DevProblem &p, int n1_begin, int n1_end,
                              int k_ig, bool kahan, int waves,
                              double res_thresh, double *partials,
                              double *n1log,
                              unsigned long long *fallback_count, int grp_lo,
                              int grp_cnt, int grid, hipStream_t s) {
  switch (k_ig) {
#define GPP_CASE(KK)                                                        \
  case KK:                                                                  \
    return launch_k<KK>(p, n1_begin, n1_end, kahan, waves, res_thresh,      \
                        partials, n1log, fallback_count, grp_lo, grp_cnt,   \
                        grid, s);
    GPP_CASE(4)
    GPP_CASE(8)
    GPP_CASE(16)
    GPP_CASE(32)
#undef GPP_CASE
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace gpp
