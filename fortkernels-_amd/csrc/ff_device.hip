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
// The frequency-grid terms (achDtemp_Kernel :155, ffKernel.f90:348-411)
// touch no large array: one thread per (n1, iw) runs the grid.
#include <algorithm>

#include "ff_internal.h"
#include "ff_math.h"

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

// The same with b.y * b.y given (by2, rounded as above).
__device__ __forceinline__ double2 cdiv_y2(double2 a, double2 b, double by2) {
  const double r = recip(fma(b.x, b.x, by2));
  return make_double2((a.x * b.x + a.y * b.y) * r, (a.y * b.x - a.x * b.y) * r);
}

// 1/d1 and 1/d2 from one reciprocal of d1*d2 (v_rcp_f64 + one Newton step);
// each through recip() when the product is not a positive normal number.
__device__ __forceinline__ void recip2(double d1, double d2, double &r1,
                                       double &r2) {
  const double p = d1 * d2;
  if (__builtin_expect(!__builtin_amdgcn_class(p, 0x100), 0)) {
    r1 = recip(d1);
    r2 = recip(d2);
    return;
  }
  double q = __builtin_amdgcn_rcp(p);
  q = fma(q, fma(-p, q, 1.0), q);
  r1 = q * d2;
  r2 = q * d1;
}

// Coefficient table of a run of passes: for pass j (jobs[j]), row G' and
// slot b < 8,
//   cGj[(j*NG + G')*8 + b] = vcoul[igp] conj(aqsm[band_b][igp])
// with igp = indinv[inv_igp_index[G']], and 0 for skipped rows (igp outside
// [0, NC)) and for slots b >= nb.  One pass's row is 128 contiguous bytes:
// two scalar loads in the slice kernel.
__global__ void job_cg_kernel(const SliceJob *__restrict__ jobs, int njobs,
                              const double2 *__restrict__ aqsm,
                              const double *__restrict__ vcoul,
                              const int *__restrict__ inv,
                              const int *__restrict__ indinv, int NC, int NG,
                              double2 *__restrict__ cGj) {
  const size_t n = (size_t)njobs * NG * kMaxJobBands;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int b = (int)(i % kMaxJobBands);
    const size_t jr = i / kMaxJobBands;
    const int g = (int)(jr % NG), j = (int)(jr / NG);
    double2 v = make_double2(0.0, 0.0);
    if (b < jobs[j].nb) {
      const int indigp = inv[g];
      int igp = (indigp >= 0 && indigp < NC) ? indinv[indigp] : -1;
      if (igp >= NC) igp = -1;
      if (igp >= 0) {
        const double2 a = aqsm[(size_t)jobs[j].band[b] * NC + igp];
        const double vc = vcoul[igp];
        v = make_double2(vc * a.x, -vc * a.y);
      }
    }
    cGj[i] = v;
  }
}

// Block: pass blockIdx.y, row group g and G chunk c from blockIdx.x.  Each
// thread owns one G and runs the group's G' rows:
//   W[b] += cGj[pass][G'][b] * slice[G'][G]   (wave-uniform: scalar loads)
// for all 8 slots (the unused ones have zero coefficients), then forms
// aqsn[band_b][G] * W[b], summed over the block into one record.
__global__ __launch_bounds__(kSliceThreads) void slice_kernel(
    const SliceJob *__restrict__ jobs, const double2 *__restrict__ cGj,
    const double2 *__restrict__ aqsn, int NC, int NG, int ngroups,
    int nchunk, double *__restrict__ rec) {
  __shared__ double red[kSliceThreads / kWave][kRecSlots];
  const SliceJob &job = jobs[blockIdx.y];
  const int nb = job.nb;
  const GlobalPtr<double2> sl = (GlobalPtr<double2>)job.slice;
  const double2 *cg = cGj + (size_t)blockIdx.y * NG * kMaxJobBands;
  const int c = blockIdx.x % nchunk, grp = blockIdx.x / nchunk;
  const int rows = (NG + ngroups - 1) / ngroups;
  const int r0 = grp * rows, r1 = min(NG, r0 + rows);
  // kSliceG G per thread, kSliceThreads apart (coalesced loads)
  int ig[kSliceG], igc[kSliceG];
  bool live[kSliceG];
#pragma unroll
  for (int u = 0; u < kSliceG; ++u) {
    ig[u] = c * kSliceThreads * kSliceG + u * kSliceThreads + threadIdx.x;
    live[u] = ig[u] < NC;
    igc[u] = live[u] ? ig[u] : NC - 1;
  }
  double wr[kSliceG][kMaxJobBands], wi[kSliceG][kMaxJobBands];
#pragma unroll
  for (int u = 0; u < kSliceG; ++u)
#pragma unroll
    for (int b = 0; b < kMaxJobBands; ++b) wr[u][b] = wi[u][b] = 0.0;
#pragma unroll 1
  for (int row = r0; row < r1; ++row) {
    double2 y[kSliceG];
#pragma unroll
    for (int u = 0; u < kSliceG; ++u) y[u] = ldg(sl + (size_t)row * NC + igc[u]);
    const double2 *cr = cg + (size_t)row * kMaxJobBands;
#pragma unroll
    for (int b = 0; b < kMaxJobBands; ++b) {
      const double2 k = cr[b];
#pragma unroll
      for (int u = 0; u < kSliceG; ++u) {
        wr[u][b] = fma(k.x, y[u].x, fma(-k.y, y[u].y, wr[u][b]));
        wi[u][b] = fma(k.x, y[u].y, fma(k.y, y[u].x, wi[u][b]));
      }
    }
  }
  double v[kRecSlots];
#pragma unroll
  for (int b = 0; b < kMaxJobBands; ++b) {
    double pr = 0.0, pi = 0.0;
#pragma unroll
    for (int u = 0; u < kSliceG; ++u)
      if (b < nb && live[u]) {
        const double2 a = aqsn[(size_t)job.band[b] * NC + ig[u]];
        pr += a.x * wr[u][b] - a.y * wi[u][b];
        pi += a.x * wi[u][b] + a.y * wr[u][b];
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
constexpr int kRedJobsPerBlock = kRedThreads / kWave;
static_assert(kRecSlots == 16 && kFreqSlots == 8,
              "the reduction shuffles assume 16 record and 8 grid slots");

// One wave per job: lane = (sub-stream lane>>4, slot lane&15).  Each lane
// sums its sub-stream of the job's records (records sub, sub+4, ...) with
// Neumaier compensation; the four sub-streams are then added in a fixed
// order (xor 16, then xor 32).
__global__ __launch_bounds__(kRedThreads) void slice_reduce_kernel(
    const double *__restrict__ rec, int njobs, int per_job,
    double *__restrict__ T) {
  const int job = blockIdx.x * kRedJobsPerBlock + threadIdx.x / kWave;
  if (job >= njobs) return;
  const int lane = threadIdx.x & (kWave - 1);
  const int slot = lane & (kRecSlots - 1), sub = lane / kRecSlots;
  const double *r = rec + (size_t)job * per_job * kRecSlots;
  double acc = 0.0, comp = 0.0;
  for (int i = sub; i < per_job; i += kWave / kRecSlots) {
    const double x = r[(size_t)i * kRecSlots + slot];
    const double t = acc + x;
    comp += (fabs(acc) >= fabs(x)) ? ((acc - t) + x) : ((x - t) + acc);
    acc = t;
  }
  double v = acc + comp;
  v += __shfl_xor(v, 16, kWave);
  v += __shfl_xor(v, 32, kWave);
  if (lane < kRecSlots) T[(size_t)job * kRecSlots + slot] = v;
}

// The per-(n1, f) values of the grid terms that do not depend on iw, formed
// once (one thread each) with the operations of the reference's loop body
// (ffKernel.f90:363-382, ffKernel.cpp:200-206); row layout:
//   0,1  i pref[f] schDt[n1][f]        2  zb - Re brd        3  cor shift
//   4    Im brd      5  right = zb     6  left = zb[f-1]     7  (Im brd)^2
//   8,9  i pref_zb avg (sch2 factor)   10,11 schDt[f-1]
//   12,13 lin/(right-left)             14,15 lin = schDt[f] - schDt[f-1]
// (ffKernel.cpp semantics: 3 = r1 r4 and 5 = r0 r4 instead.)
__global__ void freq_prep_kernel(FreqArgs a, int sem,
                                 double *__restrict__ rows) {
  const size_t n = (size_t)a.NB * a.nFreq;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int n1 = (int)(i / a.nFreq), f = (int)(i % a.nFreq);
    const bool occ = n1 < a.NV;
    const double zb = a.grid[f];
    const double2 br = a.brd[f];
    const double2 sdt = a.schdt[i];
    double r[kRowDoubles];
    r[0] = -a.pref[f] * sdt.y;
    r[1] = a.pref[f] * sdt.x;
    r[2] = zb - br.x;
    r[3] = (occ ? -zb : zb) - br.x;
    r[4] = br.y;
    r[5] = zb;
    r[7] = br.y * br.y;
    if (f > 0) {
      const double2 sp = a.schdt[i - 1];
      const double left = a.grid[f - 1];
      const double2 avg =
          make_double2(0.5 * (sdt.x + sp.x), 0.5 * (sdt.y + sp.y));
      const double2 lin = make_double2(sdt.x - sp.x, sdt.y - sp.y);
      const double rl = 1.0 / (zb - left);
      r[6] = left;
      r[8] = -a.pref_zb * avg.y;
      r[9] = a.pref_zb * avg.x;
      r[10] = sp.x;
      r[11] = sp.y;
      r[12] = lin.x * rl;
      r[13] = lin.y * rl;
      r[14] = lin.x;
      r[15] = lin.y;
    } else {
      r[6] = 0.0;
#pragma unroll
      for (int k = 8; k < kRowDoubles; ++k) r[k] = 0.0;
    }
    if (sem == 0) {
      // ffKernel.cpp reads neither the cor shift nor the interval ends; its
      // pole term uses the frequency-independent products i pref schDt *
      // Im brd instead (rows 3 and 5 then hold r1 r4 and r0 r4).
      r[3] = r[1] * r[4];
      r[5] = r[0] * r[4];
    }
    double2 *o = (double2 *)(rows + i * kRowDoubles);
#pragma unroll
    for (int k = 0; k < kRowDoubles / 2; ++k)
      o[k] = make_double2(r[2 * k], r[2 * k + 1]);
  }
}

// One thread per (n1 = blockIdx.y, iw).  SEM 0 (ffKernel.cpp:200-206):
//   ach -= (i pref[f] schDt[n1][f]) / (wx - coh[f]),  every f,
//   wx = freqevalmin - ekq[n1] + (iw-1)*freqevalstep.
// SEM 1 (ffKernel.f90:361-409), wx = freqevalmin - ekq[n1] + iw*freqevalstep:
//   f != nFreq-1: ach  -= i pref schDt / (wx - coh); corb -= ... / (wx - cor)
//   f != 0:       sch2 -= i pref_zb avg * (-log|ratio|) ; cor -= i pref_zb lin3
// A band's rows are wave-uniform (scalar loads).  blockIdx.z picks a
// contiguous range of f (the grid is split so that small NE*NB still fills
// the chip); part[((z*NB + n1)*NE + iw)*8].
// OCC: the launch covers occupied bands only (n1 = blockIdx.y) or
// unoccupied ones only (n1 = NV + blockIdx.y), so the band kind is a
// compile-time constant and the unoccupied loop carries one pole division
// and one logarithm per grid point, not two.
template <int SEM, bool OCC>
__global__ __launch_bounds__(256) void freq_kernel(FreqArgs a, int fchunk,
                                                   const double *__restrict__ rows,
                                                   double *__restrict__ part) {
  const int n1 = OCC ? (int)blockIdx.y : a.NV + (int)blockIdx.y;
  const int iw = blockIdx.x * blockDim.x + threadIdx.x;
  // the logarithm's table (4 KiB), read per lane from LDS
  __shared__ double logtab[SEM == 1 ? 2 << kLogTableBits : 1];
  if constexpr (SEM == 1) {
    for (int i = threadIdx.x; i < (2 << kLogTableBits); i += blockDim.x)
      logtab[i] = kLogTable[i];
    __syncthreads();
  }
  if (iw >= a.NE) return;
  const int f0 = blockIdx.z * fchunk;
  const int f1 = min(a.nFreq, f0 + fchunk);
  constexpr bool occ = OCC;
  const double ek = a.ekq[n1];
  const double wx = SEM == 0 ? a.fmin - ek + (iw - 1) * a.fstep
                             : a.fmin - ek + iw * a.fstep;
  const double *rb = rows + (size_t)n1 * a.nFreq * kRowDoubles;
  double achr = 0.0, achi = 0.0, cbr = 0.0, cbi = 0.0;
  double s2r = 0.0, s2i = 0.0, cr = 0.0, ci = 0.0;
  // occupied lin3 argument: -freqevalmin - ekq + (iw-1)*freqevalstep (f90:403)
  const double wocc = -a.fmin - ek + iw * a.fstep;
  int f = f0;
  if constexpr (SEM == 0) {
    // ffKernel.cpp: pole terms only; two grid points share a reciprocal
    for (; f + 1 < f1; f += 2) {
      const double *r = rb + (size_t)f * kRowDoubles;
      const double *rn = r + kRowDoubles;
      const double xa = wx - r[2], xb = wx - rn[2];
      double qa, qb;
      recip2(fma(xa, xa, r[7]), fma(xb, xb, rn[7]), qa, qb);
      achr -= fma(r[0], xa, r[3]) * qa;
      achi -= fma(r[1], xa, -r[5]) * qa;
      achr -= fma(rn[0], xb, rn[3]) * qb;
      achi -= fma(rn[1], xb, -rn[5]) * qb;
    }
  }
  if constexpr (SEM == 0) {
    for (; f < f1; ++f) {
      const double *r = rb + (size_t)f * kRowDoubles;
      const double2 t1 = cdiv_y2(make_double2(r[0], r[1]),
                                 make_double2(wx - r[2], r[4]), r[7]);
      achr -= t1.x;
      achi -= t1.y;
    }
  } else {
    // Pole terms, f != nFreq-1 (f90:386-387); two grid points share a
    // reciprocal.  Unoccupied bands: the cor pole is the coh pole (row 3 ==
    // row 2), so corb equals ach and is copied after the loops.
    const int fp1 = min(f1, a.nFreq - 1);
    for (; f + 1 < fp1; f += 2) {
      const double *r = rb + (size_t)f * kRowDoubles;
      const double *rn = r + kRowDoubles;
      const double xa = wx - r[2], xb = wx - rn[2];
      double qa, qb;
      recip2(fma(xa, xa, r[7]), fma(xb, xb, rn[7]), qa, qb);
      // -(i pref schDt) / (x + i brd.im) = -num conj(x + i brd.im) q
      achr -= fma(r[0], xa, r[1] * r[4]) * qa;
      achi -= fma(r[1], xa, -r[0] * r[4]) * qa;
      achr -= fma(rn[0], xb, rn[1] * rn[4]) * qb;
      achi -= fma(rn[1], xb, -rn[0] * rn[4]) * qb;
      if (OCC) {
        const double ya = wx - r[3], yb = wx - rn[3];
        recip2(fma(ya, ya, r[7]), fma(yb, yb, rn[7]), qa, qb);
        cbr -= fma(r[0], ya, r[1] * r[4]) * qa;
        cbi -= fma(r[1], ya, -r[0] * r[4]) * qa;
        cbr -= fma(rn[0], yb, rn[1] * rn[4]) * qb;
        cbi -= fma(rn[1], yb, -rn[0] * rn[4]) * qb;
      }
    }
    for (; f < fp1; ++f) {
      const double *r = rb + (size_t)f * kRowDoubles;
      const double2 num = make_double2(r[0], r[1]);
      const double2 t1 = cdiv_y2(num, make_double2(wx - r[2], r[4]), r[7]);
      achr -= t1.x;
      achi -= t1.y;
      if (OCC) {
        const double2 t2 = cdiv_y2(num, make_double2(wx - r[3], r[4]), r[7]);
        cbr -= t2.x;
        cbi -= t2.y;
      }
    }
    // Interval terms, f != 0 (f90:391-409): log|(wx - right)/(wx - left)|
    // as the difference of log|wx - zb| at the interval's ends, carried from
    // one grid point to the next (right = zb[f] = row 5, left = zb[f-1]).
    const int fi0 = max(f0, 1);
    if (fi0 < f1) {
      const double zl = rb[(size_t)(fi0 - 1) * kRowDoubles + 5];
      double lw = log_abs(wx - zl, logtab);              // log|wx - left|
      double lo = OCC ? log_abs(wx + zl, logtab) : 0.0;  // log|wx + left|
      for (f = fi0; f < f1; ++f) {
        const double *r = rb + (size_t)f * kRowDoubles;
        const double right = r[5], left = r[6];
        const double lr = log_abs(wx - right, logtab);
        double it = -clamp_log_ratio(lr - lw);
        lw = lr;
        // sch2 -= i pref_zb * avg * it
        s2r -= r[8] * it;
        s2i -= r[9] * it;
        double ar;
        if (OCC) {
          const double lro = log_abs(wx + right, logtab);
          it = clamp_log_ratio(lro - lo);
          lo = lro;
          ar = wocc - left;
        } else {
          ar = wx - left;
        }
        const double l3r = (r[10] + r[12] * ar) * it + r[14];
        const double l3i = (r[11] + r[13] * ar) * it + r[15];
        // cor -= i pref_zb * lin3
        cr -= -a.pref_zb * l3i;
        ci -= a.pref_zb * l3r;
      }
    }
  }
  if (SEM == 1 && !OCC) {
    cbr = achr;
    cbi = achi;
  }
  double *o =
      part + (((size_t)blockIdx.z * a.NB + n1) * a.NE + iw) * kFreqSlots;
  o[0] = achr;
  o[1] = achi;
  o[2] = cbr;
  o[3] = cbi;
  o[4] = s2r;
  o[5] = s2i;
  o[6] = cr;
  o[7] = ci;
}

// One wave per iw: lane = (row stream lane>>3, slot lane&7); each lane sums
// rows sub, sub+8, ... of part in order, then the eight streams are added
// in a fixed order (xor 8, 16, 32).
__global__ __launch_bounds__(256) void freq_reduce_kernel(
    const double *__restrict__ part, int nrow, int NE, double *__restrict__ out) {
  const int iw = blockIdx.x * (256 / kWave) + threadIdx.x / kWave;
  if (iw >= NE) return;
  const int lane = threadIdx.x & (kWave - 1);
  const int slot = lane & (kFreqSlots - 1), sub = lane / kFreqSlots;
  double v = 0.0;
  for (int r = sub; r < nrow; r += kWave / kFreqSlots)
    v += part[((size_t)r * NE + iw) * kFreqSlots + slot];
  v += __shfl_xor(v, 8, kWave);
  v += __shfl_xor(v, 16, kWave);
  v += __shfl_xor(v, 32, kWave);
  if (lane < kFreqSlots) out[(size_t)iw * kFreqSlots + slot] = v;
}

}  // namespace

int slice_chunks(int NC) {
  return (NC + kSliceThreads * kSliceG - 1) / (kSliceThreads * kSliceG);
}

hipError_t launch_cg(const SliceJob *jobs, int njobs, const double2 *aqsm,
                     const double *vcoul, const int *inv_igp_index,
                     const int *indinv, int NC, int NG, double2 *cGj,
                     hipStream_t s) {
  const size_t n = (size_t)njobs * NG * kMaxJobBands;
  if (n == 0) return hipSuccess;
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  job_cg_kernel<<<(unsigned)blocks, 256, 0, s>>>(
      jobs, njobs, aqsm, vcoul, inv_igp_index, indinv, NC, NG, cGj);
  return hipGetLastError();
}

hipError_t launch_slices(const SliceJob *jobs, int njobs, const double2 *cGj,
                         const double2 *aqsn, int NC, int NG, int ngroups,
                         double *rec, hipStream_t s) {
  if (njobs <= 0 || NG <= 0) return hipSuccess;
  if (njobs > 65535 || ngroups < 1) return hipErrorInvalidValue;
  const int nchunk = slice_chunks(NC);
  const long long bx = (long long)nchunk * ngroups;
  if (bx > 0x7fffffffLL) return hipErrorInvalidValue;
  slice_kernel<<<dim3((unsigned)bx, (unsigned)njobs), kSliceThreads, 0, s>>>(
      jobs, cGj, aqsn, NC, NG, ngroups, nchunk, rec);
  return hipGetLastError();
}

hipError_t launch_slice_reduce(const double *rec, int njobs, int per_job,
                               double *T, hipStream_t s) {
  if (njobs <= 0) return hipSuccess;
  slice_reduce_kernel<<<(njobs + kRedJobsPerBlock - 1) / kRedJobsPerBlock,
                        kRedThreads, 0, s>>>(rec, njobs, per_job, T);
  return hipGetLastError();
}

int freq_splits(int NB, int NE, int nFreq) {
  // Every (split, n1, iw) thread runs the same number of grid points, so the
  // busiest SIMD sets the time: aim for about 64 waves per SIMD (256 CUs x 4
  // SIMDs), which keeps the imbalance near 1/64; at least 16 grid points per
  // split, at most 64 splits and 256 MiB of partial sums.
  const long long threads = (long long)NB * (((long long)NE + 255) / 256 * 256);
  if (threads <= 0 || nFreq <= 0) return 1;
  long long sp = (256LL * 4 * 64 * 64) / threads;
  sp = std::min<long long>(sp, std::min(64, nFreq / 16));
  const long long part_bytes = (long long)NB * NE * kFreqSlots * 8;
  if (part_bytes > 0) sp = std::min<long long>(sp, (256LL << 20) / part_bytes);
  return (int)std::max<long long>(1, sp);
}

hipError_t launch_freq(const FreqArgs &a, int sem, int splits, double *rows,
                       double *part, hipStream_t s) {
  if (a.NB <= 0 || a.NE <= 0 || a.nFreq <= 0) return hipSuccess;
  if (a.NB > 65535 || splits < 1 || splits > 65535) return hipErrorInvalidValue;
  const size_t n = (size_t)a.NB * a.nFreq;
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  freq_prep_kernel<<<(unsigned)blocks, 256, 0, s>>>(a, sem, rows);
  const int fchunk = (a.nFreq + splits - 1) / splits;
  const int nocc = std::min(a.NV, a.NB), nun = a.NB - nocc;
  const unsigned gx = (unsigned)((a.NE + 255) / 256);
  if (nocc > 0) {
    const dim3 grid(gx, nocc, splits);
    if (sem == 0)
      freq_kernel<0, true><<<grid, 256, 0, s>>>(a, fchunk, rows, part);
    else
      freq_kernel<1, true><<<grid, 256, 0, s>>>(a, fchunk, rows, part);
  }
  if (nun > 0) {
    const dim3 grid(gx, nun, splits);
    if (sem == 0)
      freq_kernel<0, false><<<grid, 256, 0, s>>>(a, fchunk, rows, part);
    else
      freq_kernel<1, false><<<grid, 256, 0, s>>>(a, fchunk, rows, part);
  }
  return hipGetLastError();
}

hipError_t launch_freq_reduce(const double *part, int nrow, int NE,
                              double *out, hipStream_t s) {
  if (NE <= 0) return hipSuccess;
  freq_reduce_kernel<<<(NE + 3) / 4, 256, 0, s>>>(part, nrow, NE, out);
  return hipGetLastError();
}

}  // namespace ff
}  // namespace gpp
