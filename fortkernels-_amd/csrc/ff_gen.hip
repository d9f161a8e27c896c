// ff_gen.hip — synthetic inputs of the full-frequency (FF) kernels, in host
// memory (ff_fill_inputs) or generated in device memory
// (ff_fill_inputs_device: the I_eps arrays of the mini-app's example runs
// hold tens of GB).  Every large-array element is a pure function of
// (seed, stream, index) built from uniform draws with one rounding per
// operation, so host and device produce the same bits; the small tables
// (vcoul, index maps, ekq, grid, pref, dFreqBrd, schDt_matrix) are computed
// on the host and copied.
//
// config 1: the mini-app's fills (ffKernel.cpp:385-432, ffKernel.f90:87-188):
//   aqsm = aqsn = 0.5+0.5i, I_epsR = 0.5+0.5i, I_epsA = 0.5-0.5i, vcoul = 1,
//   identity index maps, ekq[i] = -10+i, dFreqGrid[k] = 2k, pref from the
//   grid, schDt_matrix = 0, dFreqBrd = 0.5+0.1i (C++) or 0.1i (F90).
// config 2: aqsm, aqsn components (2u-1)/sqrt(NC); I_eps[f][G'][G] =
//   (a + b t_f) + i (c + d t_f), t_f = f/(nFreq-1), with per-(G',G) uniform
//   a in [0.5,1), b in [-0.5,0.5), c in [-0.1,0.1), d in [-0.05,0.05);
//   vcoul[G] = 1/(0.01 + (G/NC)^(2/3)); inv_igp_index[i] = i*(NC/NG);
//   ekq sorted in [-10,10); dFreqGrid increments in [1.5,2.5);
//   dFreqBrd = (0.5u, 0.05+0.1u'); schDt_matrix components 2u-1.
// Scalars (both): freqevalmin 0, freqevalstep 0.5, pref_zb 0.5/3.14, occ 1.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "gpp_guard.h"
#include "../../include/ff_sigma.h"
#include "ff_internal.h"
#include "gpp_rng.h"

// One rounding per operation on host and device alike.
#pragma clang fp contract(off)

namespace {

using namespace gpp::rng;

struct EpsCoef {
  double a, b, c, d;
};

GPP_HD inline EpsCoef eps_coef(uint64_t seed, uint64_t stream, size_t e) {
  EpsCoef k;
  k.a = 0.5 + 0.5 * u01(draw(seed, stream, 4 * e));
  k.b = u01(draw(seed, stream, 4 * e + 1)) - 0.5;
  k.c = 0.2 * u01(draw(seed, stream, 4 * e + 2)) - 0.1;
  k.d = 0.1 * u01(draw(seed, stream, 4 * e + 3)) - 0.05;
  return k;
}

GPP_HD inline double sym(uint64_t seed, uint64_t stream, size_t i) {
  return 2.0 * u01(draw(seed, stream, i)) - 1.0;
}

__global__ void fill_eps_kernel(int config, uint64_t seed, int nFreq,
                                size_t per_slice, double tstep,
                                double2 *__restrict__ epsR,
                                double2 *__restrict__ epsA) {
  const size_t n = per_slice * (size_t)nFreq;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    if (config == 1) {
      epsR[i] = make_double2(0.5, 0.5);
      epsA[i] = make_double2(0.5, -0.5);
      continue;
    }
    const size_t e = i % per_slice;
    const double t = (double)(i / per_slice) * tstep;
    const EpsCoef r = eps_coef(seed, kFFEpsR, e), a = eps_coef(seed, kFFEpsA, e);
    epsR[i] = make_double2(r.a + r.b * t, r.c + r.d * t);
    epsA[i] = make_double2(a.a + a.b * t, a.c + a.d * t);
  }
}

__global__ void fill_aqs_kernel(int config, uint64_t seed, size_t n,
                                double scale, double2 *__restrict__ aqsm,
                                double2 *__restrict__ aqsn) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    if (config == 1) {
      aqsm[i] = aqsn[i] = make_double2(0.5, 0.5);
      continue;
    }
    aqsm[i] = make_double2(sym(seed, kFFAqsm, 2 * i) * scale,
                           sym(seed, kFFAqsm, 2 * i + 1) * scale);
    aqsn[i] = make_double2(sym(seed, kFFAqsn, 2 * i) * scale,
                           sym(seed, kFFAqsn, 2 * i + 1) * scale);
  }
}

unsigned grid_for(size_t n) {
  size_t b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  return (unsigned)(b < 1 ? 1 : b);
}

template <class F>
void parallel_for(size_t n, F f) {
  unsigned nt = std::thread::hardware_concurrency();
  if (nt == 0) nt = 1;
  if (nt > 16) nt = 16;
  if (n < (size_t)1 << 16) nt = 1;
  if (nt == 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> th;
  const size_t chunk = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; ++t) {
    const size_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    th.emplace_back([=] { f(b, e); });
  }
  for (auto &x : th) x.join();
}

bool bad_shape(int config, int semantics, int NB, int NC, int NG, int nF) {
  return (config != 1 && config != 2) ||
         (semantics != FF_SEM_CPP && semantics != FF_SEM_F90) || NB < 0 ||
         NC <= 0 || NG < 0 || NG > NC || nF < 2;
}

// The small tables, host memory.
void small_tables(int config, int semantics, uint64_t seed, int NB, int NC,
                  int NG, int nF, double *vcoul, int *inv, int *indinv,
                  double *ekq, double *grid, double *pref, gpp_c128 *brd,
                  gpp_c128 *schdt, double *scal) {
  const int stride = NG > 0 ? NC / NG : 1;
  for (int i = 0; i < NC; ++i) {
    indinv[i] = i;
    vcoul[i] = config == 1
                   ? 1.0
                   : 1.0 / (0.01 + std::pow((double)i / NC, 2.0 / 3.0));
  }
  for (int i = 0; i < NG; ++i) inv[i] = config == 1 ? i : i * stride;
  for (int i = 0; i < NB; ++i)
    ekq[i] = config == 1 ? -10.0 + i : -10.0 + 20.0 * u01(draw(seed, kFFEkq, i));
  if (config == 2) std::sort(ekq, ekq + NB);
  grid[0] = 0.0;
  for (int k = 1; k < nF; ++k)
    grid[k] = config == 1 ? 2.0 * k
                          : grid[k - 1] + 1.5 + u01(draw(seed, kFFGrid, k));
  for (int k = 0; k < nF; ++k)
    pref[k] = k < nF - 1 ? (grid[k + 1] - grid[k]) / 3.14 : pref[k - 1];
  pref[0] *= 0.5;
  pref[nF - 1] *= 0.5;
  for (int k = 0; k < nF; ++k) {
    if (config == 1) {
      brd[k].re = semantics == FF_SEM_CPP ? 0.5 : 0.0;
      brd[k].im = 0.1;
    } else {
      brd[k].re = 0.5 * u01(draw(seed, kFFBrd, 2 * (uint64_t)k));
      brd[k].im = 0.05 + 0.1 * u01(draw(seed, kFFBrd, 2 * (uint64_t)k + 1));
    }
  }
  for (size_t i = 0; i < (size_t)NB * nF; ++i) {
    schdt[i].re = config == 1 ? 0.0 : sym(seed, kFFSchDt, 2 * i);
    schdt[i].im = config == 1 ? 0.0 : sym(seed, kFFSchDt, 2 * i + 1);
  }
  scal[0] = 0.0;
  scal[1] = 0.5;
  scal[2] = 0.5 / 3.14;
  scal[3] = 1.0;
}

}  // namespace

namespace gpp {
namespace ff {

hipError_t launch_fill_eps(int config, uint64_t seed, int nFreq, int NG,
                           int NC, double2 *epsR, double2 *epsA,
                           hipStream_t s) {
  const size_t per = (size_t)NG * NC;
  if (per == 0 || nFreq <= 0) return hipSuccess;
  const double tstep = nFreq > 1 ? 1.0 / (nFreq - 1) : 0.0;
  fill_eps_kernel<<<grid_for(per * nFreq), 256, 0, s>>>(config, seed, nFreq,
                                                        per, tstep, epsR, epsA);
  return hipGetLastError();
}

hipError_t launch_fill_aqs(int config, uint64_t seed, int NB, int NC,
                           double2 *aqsm, double2 *aqsn, hipStream_t s) {
  const size_t n = (size_t)NB * NC;
  if (n == 0) return hipSuccess;
  fill_aqs_kernel<<<grid_for(n), 256, 0, s>>>(config, seed, n,
                                              1.0 / std::sqrt((double)NC),
                                              aqsm, aqsn);
  return hipGetLastError();
}

}  // namespace ff
}  // namespace gpp

extern "C" {

int ff_fill_inputs(int config, int semantics, unsigned long long seed,
                   int NB, int NC, int NG, int nF, gpp_c128 *aqsm,
                   gpp_c128 *aqsn, gpp_c128 *epsR, gpp_c128 *epsA,
                   double *vcoul, int *inv, int *indinv, double *ekq,
                   double *grid, double *pref, gpp_c128 *brd,
                   gpp_c128 *schdt, double *scal) try {
  if (bad_shape(config, semantics, NB, NC, NG, nF)) return GPP_EINVAL;
  if (!aqsm || !aqsn || !epsR || !epsA || !vcoul || !inv || !indinv ||
      !ekq || !grid || !pref || !brd || !schdt || !scal)
    return GPP_EINVAL;
  small_tables(config, semantics, seed, NB, NC, NG, nF, vcoul, inv, indinv,
               ekq, grid, pref, brd, schdt, scal);
  const double scale = 1.0 / std::sqrt((double)NC);
  parallel_for((size_t)NB * NC, [&](size_t b, size_t e) {
    for (size_t i = b; i < e; ++i) {
      if (config == 1) {
        aqsm[i] = aqsn[i] = gpp_c128{0.5, 0.5};
        continue;
      }
      aqsm[i] = gpp_c128{sym(seed, kFFAqsm, 2 * i) * scale,
                         sym(seed, kFFAqsm, 2 * i + 1) * scale};
      aqsn[i] = gpp_c128{sym(seed, kFFAqsn, 2 * i) * scale,
                         sym(seed, kFFAqsn, 2 * i + 1) * scale};
    }
  });
  const size_t per = (size_t)NG * NC;
  const double tstep = 1.0 / (nF - 1);
  parallel_for(per * nF, [&](size_t b, size_t e) {
    for (size_t i = b; i < e; ++i) {
      if (config == 1) {
        epsR[i] = gpp_c128{0.5, 0.5};
        epsA[i] = gpp_c128{0.5, -0.5};
        continue;
      }
      const size_t el = i % per;
      const double t = (double)(i / per) * tstep;
      const EpsCoef r = eps_coef(seed, kFFEpsR, el), a = eps_coef(seed, kFFEpsA, el);
      epsR[i] = gpp_c128{r.a + r.b * t, r.c + r.d * t};
      epsA[i] = gpp_c128{a.a + a.b * t, a.c + a.d * t};
    }
  });
  return GPP_OK;
} catch (...) {
  return gpp::exception_code(nullptr);
}

int ff_fill_inputs_device(int config, int semantics, unsigned long long seed,
                          int NB, int NC, int NG, int nF, gpp_c128 *aqsm,
                          gpp_c128 *aqsn, gpp_c128 *epsR, gpp_c128 *epsA,
                          double *vcoul, int *inv, int *indinv, double *ekq,
                          double *grid, double *pref, gpp_c128 *brd,
                          gpp_c128 *schdt, double *scal, void *stream) try {
  if (bad_shape(config, semantics, NB, NC, NG, nF)) return GPP_EINVAL;
  if (!aqsm || !aqsn || !epsR || !epsA || !vcoul || !inv || !indinv ||
      !ekq || !grid || !pref || !brd || !schdt || !scal)
    return GPP_EINVAL;
  std::vector<double> hv(NC), hek(NB), hg(nF), hp(nF);
  std::vector<int> hinv(NG), hind(NC);
  std::vector<gpp_c128> hb(nF), hs((size_t)NB * nF);
  small_tables(config, semantics, seed, NB, NC, NG, nF, hv.data(), hinv.data(),
               hind.data(), hek.data(), hg.data(), hp.data(), hb.data(),
               hs.data(), scal);
  hipStream_t s = (hipStream_t)stream;
  const hipMemcpyKind k = hipMemcpyHostToDevice;
  if (hipMemcpyAsync(vcoul, hv.data(), hv.size() * 8, k, s) != hipSuccess ||
      hipMemcpyAsync(inv, hinv.data(), hinv.size() * 4, k, s) != hipSuccess ||
      hipMemcpyAsync(indinv, hind.data(), hind.size() * 4, k, s) != hipSuccess ||
      hipMemcpyAsync(ekq, hek.data(), hek.size() * 8, k, s) != hipSuccess ||
      hipMemcpyAsync(grid, hg.data(), hg.size() * 8, k, s) != hipSuccess ||
      hipMemcpyAsync(pref, hp.data(), hp.size() * 8, k, s) != hipSuccess ||
      hipMemcpyAsync(brd, hb.data(), hb.size() * 16, k, s) != hipSuccess ||
      hipMemcpyAsync(schdt, hs.data(), hs.size() * 16, k, s) != hipSuccess)
    return GPP_EDEVICE;
  if (gpp::ff::launch_fill_aqs(config, seed, NB, NC, (double2 *)aqsm,
                               (double2 *)aqsn, s) != hipSuccess ||
      gpp::ff::launch_fill_eps(config, seed, nF, NG, NC, (double2 *)epsR,
                               (double2 *)epsA, s) != hipSuccess)
    return GPP_EDEVICE;
  // The host tables above are freed on return: finish their copies first.
  return hipStreamSynchronize(s) == hipSuccess ? GPP_OK : GPP_EDEVICE;
} catch (...) {
  return gpp::exception_code(nullptr);
}

}  // extern "C"
