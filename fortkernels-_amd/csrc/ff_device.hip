// ff_device.hip — gfx950 kernels of the full-frequency (FF) self-energy
// terms (FFKer_cori).
//
// Every I_eps term of the mini-app (ssxDittt_kernel ffKernel.cpp:80,
// achsDtemp_Kernel :107, schDttt_corKernel1/2 :250,279 and their F90
// counterparts) is a sum
//   sum_{G',G} aqsn[n1][G] conj(aqsm[n1][igp(G')]) vcoul[igp] X[G'][G]
// with X a linear interpolation between two frequency slices of I_epsR,
// I_epsA or their difference.  By linearity each is a combination of the
// per-slice sums
//   T(n1, slice) = sum_{G'} cG[n1][G'] sum_G aqsn[n1][G] slice[G'][G],
//   cG[n1][G'] = vcoul[igp] conj(aqsm[n1][igp]),
// so the device computes T once for every (band, slice) the evaluation
// frequencies touch and the host forms the interpolations.  A slice is
// NG*NC complex numbers read once per pass for up to 8 bands: the pass is
// HBM-bound (4 FMA per band per 16 bytes).
//
// The frequency-grid terms (achDtemp_Kernel :155, ffKernel.f90:335-384)
// touch no large array: one thread per (n1, iw) runs the grid.
#include "ff_internal.h"

namespace gpp {
namespace ff {
namespace {

constexpr int kWave = 64;

template <typename T>
using GlobalPtr = __attribute__((address_space(1))) const T *;

__device__ __forceinline__ double2 ldg(GlobalPtr<double2> q) {
  const GlobalPtr<double> d = (GlobalPtr<double>)q;
  return make_double2(d[0], d[1]);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
  return v;
}

// 1/x by v_rcp_f64 plus one Newton step for positive normal x; IEEE
// division otherwise, so zeros and non-finite values come out as the
// reference's division gives them.
__device__ __forceinline__ double recip(double x) {
  if (__builtin_expect(!__builtin_amdgcn_class(x, 0x100), 0)) return 1.0 / x;
  const double q = __builtin_amdgcn_rcp(x);
  return fma(q, fma(-x, q, 1.0), q);
}

// a / b formed as Complex.h:208 does: a*conj(b) / |b|^2.
__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
  const double r = recip(b.x * b.x + b.y * b.y);
  return make_double2((a.x * b.x + a.y * b.y) * r, (a.y * b.x - a.x * b.y) * r);
}

__global__ void cg_kernel(const double2 *__restrict__ aqsm,
                          const double *__restrict__ vcoul,
                          const int *__restrict__ inv,
                          const int *__restrict__ indinv, int NB, int NC,
                          int NG, double2 *__restrict__ cG) {
  const size_t n = (size_t)NB * NG;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int b = (int)(i / NG), g = (int)(i % NG);
    const int indigp = inv[g];
    int igp = (indigp >= 0 && indigp < NC) ? indinv[indigp] : -1;
    if (igp >= NC) igp = -1;
    if (igp < 0) {
      cG[i] = make_double2(0.0, 0.0);
    } else {
      const double2 a = aqsm[(size_t)b * NC + igp];
      const double v = vcoul[igp];
      cG[i] = make_double2(v * a.x, -v * a.y);
    }
  }
}

// Block: job blockIdx.y, row group g and G chunk c from blockIdx.x.  Each
// thread owns one G and runs the group's G' rows:
//   W[b] += cG[band_b][G'] * slice[G'][G]     (cG wave-uniform: scalar loads)
// then forms aqsn[band_b][G] * W[b], summed over the block into one record.
__global__ __launch_bounds__(kSliceThreads) void slice_kernel(
    const SliceJob *__restrict__ jobs, const double2 *__restrict__ cG,
    const double2 *__restrict__ aqsn, int NC, int NG, int ngroups,
    int nchunk, double *__restrict__ rec) {
  __shared__ double red[kSliceThreads / kWave][kRecSlots];
  const SliceJob &job = jobs[blockIdx.y];
  const int nb = job.nb;
  const GlobalPtr<double2> sl = (GlobalPtr<double2>)job.slice;
  const int c = blockIdx.x % nchunk, grp = blockIdx.x / nchunk;
  const int rows = (NG + ngroups - 1) / ngroups;
  const int r0 = grp * rows, r1 = min(NG, r0 + rows);
  const int ig = c * kSliceThreads + threadIdx.x;
  const bool live = ig < NC;
  const int igc = live ? ig : NC - 1;
  int band[kMaxJobBands];
#pragma unroll
  for (int b = 0; b < kMaxJobBands; ++b) band[b] = job.band[b];
  double wr[kMaxJobBands], wi[kMaxJobBands];
#pragma unroll
  for (int b = 0; b < kMaxJobBands; ++b) wr[b] = wi[b] = 0.0;
#pragma unroll 2
  for (int row = r0; row < r1; ++row) {
    const double2 y = ldg(sl + (size_t)row * NC + igc);
#pragma unroll
    for (int b = 0; b < kMaxJobBands; ++b) {
      if (b < nb) {
        const double2 cg = cG[(size_t)band[b] * NG + row];
        wr[b] = fma(cg.x, y.x, fma(-cg.y, y.y, wr[b]));
        wi[b] = fma(cg.x, y.y, fma(cg.y, y.x, wi[b]));
      }
    }
  }
  double v[kRecSlots];
#pragma unroll
  for (int b = 0; b < kMaxJobBands; ++b) {
    double pr = 0.0, pi = 0.0;
    if (b < nb && live) {
      const double2 a = aqsn[(size_t)band[b] * NC + ig];
      pr = a.x * wr[b] - a.y * wi[b];
      pi = a.x * wi[b] + a.y * wr[b];
    }
    v[2 * b] = pr;
    v[2 * b + 1] = pi;
  }
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < kRecSlots; ++k) {
    const double t = wave_sum(v[k]);
    if (lane == 0) red[wave][k] = t;
  }
  __syncthreads();
  if (threadIdx.x < kRecSlots) {
    double s = 0.0;
    for (int w = 0; w < kSliceThreads / kWave; ++w) s += red[w][threadIdx.x];
    rec[((size_t)blockIdx.y * ngroups * nchunk + blockIdx.x) * kRecSlots +
        threadIdx.x] = s;
  }
}

constexpr int kRedThreads = 256;

// One block per job: fixed-order (strided Neumaier, then a tree) sum of its
// records, slot by slot.
__global__ __launch_bounds__(kRedThreads) void slice_reduce_kernel(
    const double *__restrict__ rec, int per_job, double *__restrict__ T) {
  __shared__ double buf[kRedThreads];
  const double *r = rec + (size_t)blockIdx.x * per_job * kRecSlots;
  for (int s = 0; s < kRecSlots; ++s) {
    double acc = 0.0, comp = 0.0;
    for (int i = threadIdx.x; i < per_job; i += kRedThreads) {
      const double x = r[(size_t)i * kRecSlots + s];
      const double t = acc + x;
      comp += (fabs(acc) >= fabs(x)) ? ((acc - t) + x) : ((x - t) + acc);
      acc = t;
    }
    buf[threadIdx.x] = acc + comp;
    __syncthreads();
    for (int h = kRedThreads / 2; h > 0; h >>= 1) {
      if (threadIdx.x < h) buf[threadIdx.x] += buf[threadIdx.x + h];
      __syncthreads();
    }
    if (threadIdx.x == 0) T[(size_t)blockIdx.x * kRecSlots + s] = buf[0];
    __syncthreads();
  }
}

// |x| clamped to [1e-4, 1e4] as ffKernel.f90:369-371 (NaN passes through).
__device__ __forceinline__ double clamp_ratio(double x) {
  if (x < 1e-4) x = 1e-4;
  if (x > 1e4) x = 1e4;
  return x;
}

// One thread per (n1 = blockIdx.y, iw).  SEM 0 (ffKernel.cpp:200-206):
//   ach -= (i pref[f] schDt[n1][f]) / (wx - coh[f]),  every f,
//   wx = freqevalmin - ekq[n1] + (iw-1)*freqevalstep.
// SEM 1 (ffKernel.f90:335-384), wx = freqevalmin - ekq[n1] + iw*freqevalstep:
//   f != nFreq-1: ach  -= i pref schDt / (wx - coh); corb -= ... / (wx - cor)
//   f != 0:       sch2 -= i pref_zb avg * (-log|ratio|) ; cor -= i pref_zb lin3
// The grid values are wave-uniform (scalar loads).
template <int SEM>
__global__ __launch_bounds__(256) void freq_kernel(FreqArgs a,
                                                   double *__restrict__ part) {
  const int n1 = blockIdx.y;
  const int iw = blockIdx.x * blockDim.x + threadIdx.x;
  if (iw >= a.NE) return;
  const bool occ = n1 < a.NV;
  const double ek = a.ekq[n1];
  const double wx = SEM == 0 ? a.fmin - ek + (iw - 1) * a.fstep
                             : a.fmin - ek + iw * a.fstep;
  const double2 *sd = a.schdt + (size_t)n1 * a.nFreq;
  double achr = 0.0, achi = 0.0, cbr = 0.0, cbi = 0.0;
  double s2r = 0.0, s2i = 0.0, cr = 0.0, ci = 0.0;
  // occupied lin3 argument: -freqevalmin - ekq + (iw-1)*freqevalstep (f90:375)
  const double wocc = -a.fmin - ek + iw * a.fstep;
  double2 sd_prev = make_double2(0.0, 0.0);
  for (int f = 0; f < a.nFreq; ++f) {
    const double zb = a.grid[f];
    const double2 br = a.brd[f];
    const double2 sdt = sd[f];
    const double2 num = make_double2(-a.pref[f] * sdt.y, a.pref[f] * sdt.x);
    if (SEM == 0 || f != a.nFreq - 1) {
      const double2 t1 = cdiv(num, make_double2(wx - (zb - br.x), br.y));
      achr -= t1.x;
      achi -= t1.y;
      if (SEM == 1) {
        const double corr = (occ ? -zb : zb) - br.x;
        const double2 t2 = cdiv(num, make_double2(wx - corr, br.y));
        cbr -= t2.x;
        cbi -= t2.y;
      }
    }
    if (SEM == 1 && f != 0) {
      const double right = zb, left = a.grid[f - 1];
      const double2 avg =
          make_double2(0.5 * (sdt.x + sd_prev.x), 0.5 * (sdt.y + sd_prev.y));
      const double2 lin = make_double2(sdt.x - sd_prev.x, sdt.y - sd_prev.y);
      const double rl = 1.0 / (right - left);
      const double2 lin2 = make_double2(lin.x * rl, lin.y * rl);
      double it = -log(clamp_ratio(fabs((wx - right) / (wx - left))));
      // sch2 -= i pref_zb * avg * it
      s2r -= -a.pref_zb * avg.y * it;
      s2i -= a.pref_zb * avg.x * it;
      double ar;
      if (occ) {
        it = log(clamp_ratio(fabs((wx + right) / (wx + left))));
        ar = wocc - left;
      } else {
        ar = wx - left;
      }
      const double l3r = (sd_prev.x + lin2.x * ar) * it + lin.x;
      const double l3i = (sd_prev.y + lin2.y * ar) * it + lin.y;
      // cor -= i pref_zb * lin3
      cr -= -a.pref_zb * l3i;
      ci -= a.pref_zb * l3r;
    }
    sd_prev = sdt;
  }
  double *o = part + ((size_t)n1 * a.NE + iw) * kFreqSlots;
  o[0] = achr;
  o[1] = achi;
  o[2] = cbr;
  o[3] = cbi;
  o[4] = s2r;
  o[5] = s2i;
  o[6] = cr;
  o[7] = ci;
}

__global__ void freq_reduce_kernel(const double *__restrict__ part, int NB,
                                   int NE, double *__restrict__ out) {
  const int iw = blockIdx.x * blockDim.x + threadIdx.x;
  if (iw >= NE) return;
  double s[kFreqSlots] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int n1 = 0; n1 < NB; ++n1)
#pragma unroll
    for (int k = 0; k < kFreqSlots; ++k)
      s[k] += part[((size_t)n1 * NE + iw) * kFreqSlots + k];
#pragma unroll
  for (int k = 0; k < kFreqSlots; ++k) out[(size_t)iw * kFreqSlots + k] = s[k];
}

}  // namespace

int slice_chunks(int NC) { return (NC + kSliceThreads - 1) / kSliceThreads; }

hipError_t launch_cg(const double2 *aqsm, const double *vcoul,
                     const int *inv_igp_index, const int *indinv, int NB,
                     int NC, int NG, double2 *cG, hipStream_t s) {
  const size_t n = (size_t)NB * NG;
  if (n == 0) return hipSuccess;
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  cg_kernel<<<(unsigned)blocks, 256, 0, s>>>(aqsm, vcoul, inv_igp_index,
                                             indinv, NB, NC, NG, cG);
  return hipGetLastError();
}

hipError_t launch_slices(const SliceJob *jobs, int njobs, const double2 *cG,
                         const double2 *aqsn, int NC, int NG, int ngroups,
                         double *rec, hipStream_t s) {
  if (njobs <= 0 || NG <= 0) return hipSuccess;
  if (njobs > 65535 || ngroups < 1) return hipErrorInvalidValue;
  const int nchunk = slice_chunks(NC);
  const long long bx = (long long)nchunk * ngroups;
  if (bx > 0x7fffffffLL) return hipErrorInvalidValue;
  slice_kernel<<<dim3((unsigned)bx, (unsigned)njobs), kSliceThreads, 0, s>>>(
      jobs, cG, aqsn, NC, NG, ngroups, nchunk, rec);
  return hipGetLastError();
}

hipError_t launch_slice_reduce(const double *rec, int njobs, int per_job,
                               double *T, hipStream_t s) {
  if (njobs <= 0) return hipSuccess;
  slice_reduce_kernel<<<njobs, kRedThreads, 0, s>>>(rec, per_job, T);
  return hipGetLastError();
}

hipError_t launch_freq(const FreqArgs &a, int sem, double *part,
                       hipStream_t s) {
  if (a.NB <= 0 || a.NE <= 0) return hipSuccess;
  if (a.NB > 65535) return hipErrorInvalidValue;
  const dim3 grid((a.NE + 255) / 256, a.NB);
  if (sem == 0)
    freq_kernel<0><<<grid, 256, 0, s>>>(a, part);
  else
    freq_kernel<1><<<grid, 256, 0, s>>>(a, part);
  return hipGetLastError();
}

hipError_t launch_freq_reduce(const double *part, int NB, int NE, double *out,
                              hipStream_t s) {
  if (NE <= 0) return hipSuccess;
  freq_reduce_kernel<<<(NE + 255) / 256, 256, 0, s>>>(part, NB, NE, out);
  return hipGetLastError();
}

}  // namespace ff
}  // namespace gpp
