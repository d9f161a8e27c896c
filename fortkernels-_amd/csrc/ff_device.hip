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
// (ffKernel.f90:363-382, ffKernel.cpp:200-206), in two row arrays of
// kRowHalf doubles per (n1, f) so that each grid loop reads one 64-byte row
// per point with one scalar load.  Pole rows (both semantics):
//   0,1  num = i pref[f] schDt[n1][f]     2  coh shift zb - Re brd
//   3    (Im brd)^2       4  num.im Im brd    5  num.re Im brd
//   6    cor shift (ffKernel.f90: -zb - Re brd for occupied bands, else 2)
//   7    Im brd
// (the unoccupied and ffKernel.cpp loops read 0-5 only).
// Interval rows (ffKernel.f90 semantics; f = 0 holds only 0):
//   0    right = zb[f]    1,2  i pref_zb avg (sch2 factor)
//   3,4  pref_zb schDt[f-1]      5,6  pref_zb lin / (right - left)
__global__ void freq_prep_kernel(FreqArgs a, int sem,
                                 double *__restrict__ prow,
                                 double *__restrict__ irow) {
  const size_t n = (size_t)a.NB * a.nFreq;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int n1 = (int)(i / a.nFreq), f = (int)(i % a.nFreq);
    const bool occ = n1 < a.NV;
    const double zb = a.grid[f];
    const double2 br = a.brd[f];
    const double2 sdt = a.schdt[i];
    double p[kRowHalf];
    p[0] = -a.pref[f] * sdt.y;
    p[1] = a.pref[f] * sdt.x;
    p[2] = zb - br.x;
    p[3] = br.y * br.y;
    p[4] = p[1] * br.y;
    p[5] = p[0] * br.y;
    p[6] = (sem == 1 && occ) ? -zb - br.x : p[2];
    p[7] = br.y;
    double2 *o = (double2 *)(prow + i * kRowHalf);
#pragma unroll
    for (int k = 0; k < kRowHalf / 2; ++k)
      o[k] = make_double2(p[2 * k], p[2 * k + 1]);
    if (sem == 0) continue;
    double q[kRowHalf];
#pragma unroll
    for (int k = 0; k < kRowHalf; ++k) q[k] = 0.0;
    q[0] = zb;
    if (f > 0) {
      const double2 sp = a.schdt[i - 1];
      const double left = a.grid[f - 1];
      const double2 avg =
          make_double2(0.5 * (sdt.x + sp.x), 0.5 * (sdt.y + sp.y));
      const double2 lin = make_double2(sdt.x - sp.x, sdt.y - sp.y);
      const double rl = 1.0 / (zb - left);
      q[1] = -a.pref_zb * avg.y;
      q[2] = a.pref_zb * avg.x;
      q[3] = a.pref_zb * sp.x;
      q[4] = a.pref_zb * sp.y;
      q[5] = a.pref_zb * (lin.x * rl);
      q[6] = a.pref_zb * (lin.y * rl);
    }
    o = (double2 *)(irow + i * kRowHalf);
#pragma unroll
    for (int k = 0; k < kRowHalf / 2; ++k)
      o[k] = make_double2(q[2 * k], q[2 * k + 1]);
  }
}

template <typename T>
using ConstPtr = __attribute__((address_space(4))) const T *;

// One 64-byte row (kRowHalf doubles) of a grid-term row array.
struct Row {
  double v[kRowHalf];
};

// The row at p, read whole at the point of the call.  p goes through an
// empty asm so that the loop optimizer keeps the pointer at the row's
// start (the loads then take non-negative immediate offsets, with no
// address arithmetic per load); the constant address space (the rows are
// written by an earlier launch and only read here) keeps them scalar.
__device__ __forceinline__ Row load_row(const double *p) {
  uint64_t a = (uint64_t)p;
  asm volatile("" : "+s"(a));
  const ConstPtr<double> q = (ConstPtr<double>)a;
  Row r;
#pragma unroll
  for (int k = 0; k < kRowHalf; ++k) r.v[k] = q[k];
  return r;
}

// -num / (x + i Im brd) = -num (x - i Im brd) q, q = 1/(x^2 + (Im brd)^2),
// with the pole row's products num Im brd: re -= num.re (x q) + (num.im
// Im brd) q, im -= num.im (x q) - (num.re Im brd) q.
__device__ __forceinline__ void pole_term(const Row &r, double x, double q,
                                          double &re, double &im) {
  const double u = x * q;
  re = fma(-r.v[0], u, re);
  re = fma(-r.v[4], q, re);
  im = fma(-r.v[1], u, im);
  im = fma(r.v[5], q, im);
}

// The pair loops form -(x^2 + (Im brd)^2) = fma(x, -x, -(Im brd)^2): with
// the negations the compiler keeps the three-operand fma, whose addend may
// be the row's scalar register, instead of copying the scalar into the
// destination of a two-operand fmac.  The reciprocals of the negated sums
// are the negated reciprocals, and pole_term_neg takes them.
__device__ __forceinline__ double neg_den(double x, double b2) {
  return fma(x, -x, -b2);
}

// pole_term with nq = -q.
__device__ __forceinline__ void pole_term_neg(const Row &r, double x,
                                              double nq, double &re,
                                              double &im) {
  const double nu = x * nq;
  re = fma(r.v[0], nu, re);
  re = fma(r.v[4], nq, re);
  im = fma(r.v[1], nu, im);
  im = fma(-r.v[5], nq, im);
}

// 1/d1 and 1/d2 from one reciprocal of d1*d2 (v_rcp_f64 + one Newton
// step); the test for a product that is not a positive normal number is
// taken once per wave (a uniform branch), and the lanes that fail it redo
// their pair with recip().
__device__ __forceinline__ void recip2_wave(double d1, double d2, double &r1,
                                            double &r2) {
  const double p = d1 * d2;
  double q = __builtin_amdgcn_rcp(p);
  q = fma(q, fma(-p, q, 1.0), q);
  r1 = q * d2;
  r2 = q * d1;
  const bool odd = !__builtin_amdgcn_class(p, 0x100);
  if (__builtin_expect(__builtin_amdgcn_ballot_w64(odd) != 0, 0)) {
    if (odd) {
      r1 = recip(d1);
      r2 = recip(d2);
    }
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
// OCC: the body for occupied bands (n1 < NV) or unoccupied ones, a
// compile-time constant, so that the unoccupied loop carries one pole
// division and one logarithm per grid point, not two.
template <int SEM, bool OCC>
__device__ __forceinline__ void freq_body(const FreqArgs &a, int fchunk,
                                          const double *__restrict__ prow,
                                          const double *__restrict__ irow,
                                          double *__restrict__ part,
                                          const double *logtab) {
  const int n1 = (int)blockIdx.y;
  const int iw = blockIdx.x * blockDim.x + threadIdx.x;
  if constexpr (SEM == 1) {
    // whole waves past NE leave; lanes past NE in a live wave run along (the
    // wave sums below need every lane) and store nothing
    if (iw - (int)(threadIdx.x & (kWave - 1)) >= a.NE) return;
  } else {
    if (iw >= a.NE) return;
  }
  const int f0 = blockIdx.z * fchunk;
  const int f1 = min(a.nFreq, f0 + fchunk);
  const double ek = a.ekq[n1];
  const double wx = SEM == 0 ? a.fmin - ek + (iw - 1) * a.fstep
                             : a.fmin - ek + iw * a.fstep;
  const double *pb = prow + (size_t)n1 * a.nFreq * kRowHalf;
  double achr = 0.0, achi = 0.0, cbr = 0.0, cbi = 0.0;
  double s2r = 0.0, s2i = 0.0, cr = 0.0, ci = 0.0;
  int f = f0;
  if constexpr (SEM == 0) {
    // ffKernel.cpp: pole terms only; two grid points share a reciprocal
    const double *p = pb + (size_t)f * kRowHalf;
    for (; f + 1 < f1; f += 2, p += 2 * kRowHalf) {
      const Row r = load_row(p), rn = load_row(p + kRowHalf);
      const double xa = wx - r.v[2], xb = wx - rn.v[2];
      double qa, qb;
      recip2_wave(neg_den(xa, r.v[3]), neg_den(xb, rn.v[3]), qa, qb);
      pole_term_neg(r, xa, qa, achr, achi);
      pole_term_neg(rn, xb, qb, achr, achi);
    }
    for (; f < f1; ++f) {
      const Row r = load_row(pb + (size_t)f * kRowHalf);
      const double x = wx - r.v[2];
      pole_term(r, x, recip(fma(x, x, r.v[3])), achr, achi);
    }
  } else {
    // Pole terms, f != nFreq-1 (f90:386-387); two grid points share a
    // reciprocal.  Unoccupied bands: the cor pole is the coh pole (row 6 ==
    // row 2), so corb equals ach and is copied after the loops.
    const int fp1 = min(f1, a.nFreq - 1);
    const double *p = pb + (size_t)f * kRowHalf;
    for (; f + 1 < fp1; f += 2, p += 2 * kRowHalf) {
      const Row r = load_row(p), rn = load_row(p + kRowHalf);
      const double xa = wx - r.v[2], xb = wx - rn.v[2];
      double qa, qb;
      recip2_wave(neg_den(xa, r.v[3]), neg_den(xb, rn.v[3]), qa, qb);
      pole_term_neg(r, xa, qa, achr, achi);
      pole_term_neg(rn, xb, qb, achr, achi);
      if (OCC) {
        const double ya = wx - r.v[6], yb = wx - rn.v[6];
        recip2_wave(neg_den(ya, r.v[3]), neg_den(yb, rn.v[3]), qa, qb);
        pole_term_neg(r, ya, qa, cbr, cbi);
        pole_term_neg(rn, yb, qb, cbr, cbi);
      }
    }
    for (; f < fp1; ++f) {
      const Row r = load_row(pb + (size_t)f * kRowHalf);
      const double x = wx - r.v[2];
      pole_term(r, x, recip(fma(x, x, r.v[3])), achr, achi);
      if (OCC) {
        const double y = wx - r.v[6];
        pole_term(r, y, recip(fma(y, y, r.v[3])), cbr, cbi);
      }
    }
    // Interval terms, f != 0 (f90:391-409): log|(wx - right)/(wx - left)|
    // as the difference of log|wx - zb| at the interval's ends, carried from
    // one grid point to the next (right = zb[f], left = zb[f-1]);
    //   sch2 -= i pref_zb avg it
    //   cor  -= i pref_zb ((schDt[f-1] + lin2 ar) it + lin),
    // ar = wx - left (unoccupied: the previous point's wx - right) or
    // -freqevalmin - ekq + (iw-1)*freqevalstep - left (occupied, f90:403),
    // the lin part summed once below (it does not depend on iw), the rest
    // with the row's pref_zb-scaled schDt[f-1] and lin2.
    const int fi0 = max(f0, 1);
    if (fi0 < f1) {
      const double *ib = irow + (size_t)n1 * a.nFreq * kRowHalf;
      const double wocc = -a.fmin - ek + iw * a.fstep;
      const double zl = ib[(size_t)(fi0 - 1) * kRowHalf];
      IntervalLog lw, lo;  // log|wx - zb|, log|wx + zb| (occupied)
      double ar = OCC ? wocc - zl : wx - zl;
      lw.start(wx - zl, logtab);
      if (OCC) lo.start(wx + zl, logtab);
      const double *p = ib + (size_t)fi0 * kRowHalf;
      for (f = fi0; f < f1; ++f, p += kRowHalf) {
        const Row rw = load_row(p);
        const double *r = rw.v;
        const double right = r[0];
        const double xr = wx - right;
        double it = -lw.step(xr, logtab);
        s2r = fma(-r[1], it, s2r);
        s2i = fma(-r[2], it, s2i);
        double anext;
        if (OCC) {
          it = lo.step(wx + right, logtab);
          anext = wocc - right;
        } else {
          anext = xr;
        }
        const double u = ar * it;
        cr = fma(r[4], it, cr);
        cr = fma(r[6], u, cr);
        ci = fma(-r[3], it, ci);
        ci = fma(-r[5], u, ci);
        ar = anext;
      }
      // sum of lin = schDt[f] - schDt[f-1] over the block's intervals, the
      // same for every iw: lanes take every 64th interval, then a wave sum
      const double2 *sd = a.schdt + (size_t)n1 * a.nFreq;
      double lre = 0.0, lim = 0.0;
      for (int g = fi0 + (int)(threadIdx.x & (kWave - 1)); g < f1; g += kWave) {
        const double2 sr = sd[g], sl = sd[g - 1];
        lre += sr.x - sl.x;
        lim += sr.y - sl.y;
      }
      lre = wave_sum(lre);
      lim = wave_sum(lim);
      cr = fma(a.pref_zb, lim, cr);
      ci = fma(-a.pref_zb, lre, ci);
    }
  }
  if (SEM == 1 && !OCC) {
    cbr = achr;
    cbi = achi;
  }
  if (iw >= a.NE) return;
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

// One launch over all bands (n1 = blockIdx.y): the occupied bands' blocks
// take the OCC body, the rest the unoccupied one (a uniform branch per
// block), so the two kinds share the chip instead of running one launch
// after the other.  ffKernel.cpp semantics has one body for both.
template <int SEM>
__global__ __launch_bounds__(256) void freq_kernel(FreqArgs a, int fchunk,
                                                   const double *__restrict__ prow,
                                                   const double *__restrict__ irow,
                                                   double *__restrict__ part) {
  // the logarithm's table (4 KiB), read per lane from LDS
  __shared__ double logtab[SEM == 1 ? 2 << kLogTableBits : 1];
  if constexpr (SEM == 1) {
    for (int i = threadIdx.x; i < (2 << kLogTableBits); i += blockDim.x)
      logtab[i] = kLogTable[i];
    __syncthreads();
    if ((int)blockIdx.y < a.NV)
      freq_body<1, true>(a, fchunk, prow, irow, part, logtab);
    else
      freq_body<1, false>(a, fchunk, prow, irow, part, logtab);
  } else {
    freq_body<0, false>(a, fchunk, prow, irow, part, logtab);
  }
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
  double *prow = rows, *irow = rows + n * kRowHalf;
  freq_prep_kernel<<<(unsigned)blocks, 256, 0, s>>>(a, sem, prow, irow);
  const int fchunk = (a.nFreq + splits - 1) / splits;
  const dim3 grid((unsigned)((a.NE + 255) / 256), a.NB, splits);
  if (sem == 0)
    freq_kernel<0><<<grid, 256, 0, s>>>(a, fchunk, prow, irow, part);
  else
    freq_kernel<1><<<grid, 256, 0, s>>>(a, fchunk, prow, irow, part);
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
