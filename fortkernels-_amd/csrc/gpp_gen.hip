// gpp_gen.hip — synthetic inputs generated directly in device memory, for
// shapes whose arrays are too large to stage from the host every run (the
// 2048/16/65536/20 instance holds 11 GB).  Same counter-based draws as the
// host generator (gpp_inputs.cpp): the uniform-derived arrays (I_eps,
// wtilde, the near-resonance placement) are bit-identical to the host's;
// the normal variates (aqsm, aqsn) use the device's log/sin/cos and agree
// with the host's to a few ulp.  The small tables (vcoul, index maps,
// frequencies, sorted band energies) are computed on the host and copied,
// so they are identical.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "gpp_guard.h"
#include "../../include/gpp_sigma.h"
#include "gpp_rng.h"

namespace {

using namespace gpp::rng;

constexpr int kThreads = 256;

// The host generator evaluates a + b*u as a rounded multiply then a rounded
// add; so must the device (no fused multiply-add) to give the same bits.
#pragma clang fp contract(off)

unsigned grid_for(size_t n) {
  size_t b = (n + kThreads - 1) / kThreads;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (unsigned)b;
}

__global__ void fill_const_kernel(double2 *a, size_t n, double2 v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    a[i] = v;
}

__global__ void normal_kernel(uint64_t seed, uint64_t stream, size_t n,
                              double scale, double2 *out) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const double u1 = u01_open0(draw(seed, stream, 2 * i));
    const double u2 = u01(draw(seed, stream, 2 * i + 1));
    const double r = sqrt(-2.0 * log(u1));
    const double t = 6.283185307179586 * u2;
    out[i] = make_double2(r * cos(t) * scale, r * sin(t) * scale);
  }
}

__global__ void ieps_wtilde_kernel(uint64_t seed, size_t n, double2 *ieps,
                                   double2 *wt) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    ieps[i] = make_double2(u01(draw(seed, kIeps, 2 * i)),
                           -0.1 + 0.2 * u01(draw(seed, kIeps, 2 * i + 1)));
    wt[i] = make_double2(0.5 + 2.5 * u01(draw(seed, kWtilde, 2 * i)),
                         -0.05 + 0.1 * u01(draw(seed, kWtilde, 2 * i + 1)));
  }
}

__global__ void resonance_kernel(uint64_t seed, size_t n, int NB,
                                 const double *__restrict__ wx,
                                 double2 *wt) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    if (u01(draw(seed, kResPick, i)) >= 0.05) continue;
    const uint64_t x = draw(seed, kResWhere, 2 * i);
    const int n1 = (int)((x >> 11) % (uint64_t)NB);
    const int iw = (int)((x >> 3) % 3);
    const double off = -1e-4 + 2e-4 * u01(draw(seed, kResWhere, 2 * i + 1));
    wt[i].x = wx[n1 * 3 + iw] + off;
  }
}

}  // namespace

extern "C" int gpp_fill_inputs_device(
    int config, unsigned long long seed, int number_bands, int ncouls,
    int ngpown, int nodes_per_group, gpp_c128 *aqsmtemp, gpp_c128 *aqsntemp,
    gpp_c128 *I_eps_array, gpp_c128 *wtilde_array, double *vcoul,
    int *inv_igp_index, int *indinv, double *wx_array, void *stream) try {
  const int NB = number_bands, NC = ncouls, NG = ngpown;
  const bool vlin = config & GPP_FILL_VCOUL_LINEAR;
  config &= ~GPP_FILL_VCOUL_LINEAR;
  if (config < 1 || config > 3 || NB < 0 || NC <= 0 || NG < 0)
    return GPP_EINVAL;
  if (!aqsmtemp || !aqsntemp || !I_eps_array || !wtilde_array || !vcoul ||
      !inv_igp_index || !indinv || !wx_array)
    return GPP_EINVAL;
  if (config != 1 && (nodes_per_group <= 0 ||
                      (long long)(NG > 0 ? NG - 1 : 0) * nodes_per_group >= NC))
    return GPP_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  const size_t nbnc = (size_t)NB * NC, ngnc = (size_t)NG * NC;
  double2 *aqsm = (double2 *)aqsmtemp, *aqsn = (double2 *)aqsntemp;
  double2 *ieps = (double2 *)I_eps_array, *wt = (double2 *)wtilde_array;

  // Small tables on the host, exactly as gpp_fill_inputs computes them.
  std::vector<double> hv(NC), hwx(config == 3 ? (size_t)NB * 3 : 3);
  std::vector<int> hinv(NG > 0 ? NG : 1), hind(NC);
  for (int i = 0; i < NC; ++i) hind[i] = i;
  if (config == 1) {
    for (int i = 0; i < NC; ++i) hv[i] = 1.0;
    for (int k = 0; k < NG; ++k)
      hinv[k] = (int)((long long)(k + 1) * NC / NG);
    for (int iw = 0; iw < 3; ++iw) hwx[iw] = 10.0 - 6.0 + (iw - 1);
  } else {
    for (int ig = 0; ig < NC; ++ig)
      hv[ig] = 1.0 / (0.01 + std::pow((double)ig / NC, 2.0 / 3.0));
    for (int i = 0; i < NG; ++i) hinv[i] = i * nodes_per_group;
    if (config == 2) {
      for (int iw = 0; iw < 3; ++iw) hwx[iw] = 4.0 + (iw - 1);
    } else {
      std::vector<double> en(NB);
      for (int n = 0; n < NB; ++n)
        en[n] = -20.0 + 40.0 * u01(draw(seed, kEnergy, n));
      std::sort(en.begin(), en.end());
      for (int n = 0; n < NB; ++n)
        for (int iw = 0; iw < 3; ++iw) {
          double w = 4.0 - en[n] + (iw - 1);
          if (std::fabs(w) < 1e-6) w = 1e-6;
          hwx[(size_t)n * 3 + iw] = w;
        }
    }
  }
  if (vlin)
    for (int i = 0; i < NC; ++i) hv[i] = (double)i / NC;
#define GEN_OR_RET(expr)                                   \
  do {                                                     \
    if ((expr) != hipSuccess) return GPP_EDEVICE;          \
  } while (0)
  GEN_OR_RET(hipMemcpyAsync(vcoul, hv.data(), hv.size() * 8,
                            hipMemcpyHostToDevice, s));
  GEN_OR_RET(hipMemcpyAsync(indinv, hind.data(), hind.size() * 4,
                            hipMemcpyHostToDevice, s));
  if (NG > 0)
    GEN_OR_RET(hipMemcpyAsync(inv_igp_index, hinv.data(), (size_t)NG * 4,
                              hipMemcpyHostToDevice, s));
  GEN_OR_RET(hipMemcpyAsync(wx_array, hwx.data(), hwx.size() * 8,
                            hipMemcpyHostToDevice, s));

  if (config == 1) {
    const double2 c = make_double2(0.5, 0.5);
    if (nbnc) {
      fill_const_kernel<<<grid_for(nbnc), kThreads, 0, s>>>(aqsm, nbnc, c);
      fill_const_kernel<<<grid_for(nbnc), kThreads, 0, s>>>(aqsn, nbnc, c);
    }
    if (ngnc) {
      fill_const_kernel<<<grid_for(ngnc), kThreads, 0, s>>>(ieps, ngnc, c);
      fill_const_kernel<<<grid_for(ngnc), kThreads, 0, s>>>(wt, ngnc, c);
    }
  } else {
    const double sc = 1.0 / std::sqrt((double)NC);
    if (nbnc) {
      normal_kernel<<<grid_for(nbnc), kThreads, 0, s>>>(seed, kAqsm, nbnc, sc,
                                                        aqsm);
      normal_kernel<<<grid_for(nbnc), kThreads, 0, s>>>(seed, kAqsn, nbnc, sc,
                                                        aqsn);
    }
    if (ngnc)
      ieps_wtilde_kernel<<<grid_for(ngnc), kThreads, 0, s>>>(seed, ngnc, ieps,
                                                             wt);
    if (config == 3 && NB > 0 && ngnc)
      resonance_kernel<<<grid_for(ngnc), kThreads, 0, s>>>(seed, ngnc, NB,
                                                           wx_array, wt);
  }
  GEN_OR_RET(hipGetLastError());
  GEN_OR_RET(hipStreamSynchronize(s));
#undef GEN_OR_RET
  return GPP_OK;
} catch (...) {
  return gpp::exception_code(nullptr);
}
