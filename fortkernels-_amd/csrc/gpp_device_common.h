// gpp_device_common.h — device helpers shared by the kernel sources
// (gpp_device.hip, gpp_mixed.hip): loads, reciprocals, wave and block
// reductions, the XCD-aware tile order, and the mixed-precision
// non-occupied block (also run by the batched launch).  Everything is in
// an anonymous namespace: each source gets its own copy.
#pragma once

#include "gpp_internal.h"

namespace gpp {
namespace {

constexpr int kWave = 64;

__device__ __forceinline__ double rcp_f64(double x) {
  return __builtin_amdgcn_rcp(x);
}
__device__ __forceinline__ float rcp_f32(float x) {
  return __builtin_amdgcn_rcpf(x);
}
// v_cmp_class_f64 with only the "positive normal" bit set.
__device__ __forceinline__ bool is_pos_normal(double x) {
  return __builtin_amdgcn_class(x, 0x100);
}
// Loads of complex elements through generic pointers (kernel-argument
// problems) or global-address-space pointers (problems read from memory by
// the batched launch, where that type lets uniform loads become scalar).
template <typename T>
using GlobalPtr = __attribute__((address_space(1))) const T *;
__device__ __forceinline__ double2 ld(const double2 *q) { return *q; }
__device__ __forceinline__ float2 ld(const float2 *q) { return *q; }
__device__ __forceinline__ double2 ld(GlobalPtr<double2> q) {
  const GlobalPtr<double> d = (GlobalPtr<double>)q;
  return make_double2(d[0], d[1]);
}
__device__ __forceinline__ float2 ld(GlobalPtr<float2> q) {
  const GlobalPtr<float> f = (GlobalPtr<float>)q;
  return make_float2(f[0], f[1]);
}

// 1/x: v_rcp_f64 plus one Newton step (about 1e-15 relative) for positive
// normal x, IEEE division otherwise (zero, denormal, inf, NaN), so special
// values come out as the reference's division gives them.
__device__ __forceinline__ double recip(double x) {
  if (__builtin_expect(!is_pos_normal(x), 0)) return 1.0 / x;
  const double q = rcp_f64(x);
  return fma(q, fma(-x, q, 1.0), q);
}

// r[i] = 1/d[i] for three values with one v_rcp_f64 (batch inversion);
// IEEE divisions if the product is not a positive normal number.
__device__ __forceinline__ void recip3(double d0, double d1, double d2,
                                       double &r0, double &r1, double &r2) {
  const double p01 = d0 * d1;
  const double pp = p01 * d2;
  double q = rcp_f64(pp);
  q = fma(q, fma(-pp, q, 1.0), q);  // 1 / (d0 d1 d2)
  const double t = q * d2;          // 1 / (d0 d1)
  r2 = q * p01;
  r0 = t * d1;
  r1 = t * d0;
  if (__builtin_expect(!is_pos_normal(pp), 0)) {
    r0 = 1.0 / d0;
    r1 = 1.0 / d1;
    r2 = 1.0 / d2;
  }
}

__device__ __forceinline__ int uniform(int x) {
  return __builtin_amdgcn_readfirstlane(x);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
  return v;
}

// Sum `n` doubles per lane over the block and write them to rec[0..n).
// Fixed order: lanes by xor-butterfly, then waves 0,1,2,... in sequence.
template <int N>
__device__ __forceinline__ void block_store(double (&v)[N], double *rec) {
  __shared__ double red[16][N];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] = wave_sum(v[i]);
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < N; ++i) red[wave][i] = v[i];
  }
  __syncthreads();
  if (threadIdx.x < N) {
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += red[w][threadIdx.x];
    rec[threadIdx.x] = s;
  }
}

// Per-band log (acht_n1_loc): sum one complex value per lane over the block
// (lanes by butterfly, waves in order) and store it at dst[2*idx].  Called
// by every thread of the block; two barriers.
__device__ __forceinline__ void n1log_store(double re, double im, double *dst,
                                            size_t idx) {
  __shared__ double lg[16][2];
  re = wave_sum(re);
  im = wave_sum(im);
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
  if (lane == 0) {
    lg[wave][0] = re;
    lg[wave][1] = im;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sr = 0.0, si = 0.0;
    for (int w = 0; w < nw; ++w) {
      sr += lg[w][0];
      si += lg[w][1];
    }
    dst[2 * idx] = sr;
    dst[2 * idx + 1] = si;
  }
  __syncthreads();
}

// Per-band log of a band loop: the sum of one complex value per lane over
// the block for every band, stored at dst[2 * (base + j)], with one barrier
// per kBandLogChunk bands instead of two per band (n1log_store).  Each wave
// puts its sum for band j into an LDS slot; at the end of every chunk of
// kBandLogChunk bands (and after the last band) the block waits once and
// threads 0 .. chunk-1 add one band's slots over the waves, in wave order,
// and store it.  The slots are double-buffered by chunk: a chunk's slots
// are written again two chunks later, after the barrier that ends the
// chunk in between, by which time every thread has stored its band.
// A wave's sum: G = 1, a butterfly over the lanes per band (as
// n1log_store); G > 1, the values of G bands are staged in LDS and summed
// together, kWave / G lanes per band each adding G entries and then a
// butterfly over those lanes: about 10 instructions per band for G = 4
// instead of about 50, but G x 1 KB of LDS per wave.  The lanes of a wave
// execute their LDS accesses in order, so no fence is needed between the
// staging writes and the reads.
constexpr int kBandLogChunk = 64;
template <int G>
struct BandLog {
  static_assert(kBandLogChunk % G == 0 && kWave % G == 0, "group size");
  double (*slot)[kBandLogChunk][4][2];  // [2][band in chunk][wave][re, im]
  double2 (*stage)[kWave];              // G > 1: this wave's [G][kWave]
  double *dst;
  size_t base;
  int nbands;

  // Called by every thread of the block for band j = 0 .. nbands-1 in order.
  __device__ __forceinline__ void put(int j, double re, double im) const {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int c = j / kBandLogChunk, i = j % kBandLogChunk;
    if constexpr (G == 1) {
      re = wave_sum(re);
      im = wave_sum(im);
      if (lane == 0) {
        slot[c & 1][i][wave][0] = re;
        slot[c & 1][i][wave][1] = im;
      }
    } else {
      const int g = j % G;
      stage[g][lane] = make_double2(re, im);
      if (g == G - 1 || j == nbands - 1) {  // wave-uniform
        constexpr int L = kWave / G;  // lanes per band
        const int band = lane / L, part = lane % L;
        double sr = 0.0, si = 0.0;
        if (band <= g) {
#pragma unroll
          for (int q = 0; q < G; ++q) {
            const double2 v = stage[band][part * G + q];
            sr += v.x;
            si += v.y;
          }
        }
#pragma unroll
        for (int m = L / 2; m >= 1; m >>= 1) {
          sr += __shfl_xor(sr, m, kWave);
          si += __shfl_xor(si, m, kWave);
        }
        if (part == 0 && band <= g) {
          slot[c & 1][i - g + band][wave][0] = sr;
          slot[c & 1][i - g + band][wave][1] = si;
        }
      }
    }
    if (i == kBandLogChunk - 1 || j == nbands - 1) {  // block-uniform
      __syncthreads();
      if ((int)threadIdx.x <= i) {
        const int nw = blockDim.x / kWave;
        double sr = 0.0, si = 0.0;
        for (int w = 0; w < nw; ++w) {
          sr += slot[c & 1][threadIdx.x][w][0];
          si += slot[c & 1][threadIdx.x][w][1];
        }
        const size_t k = base + (size_t)c * kBandLogChunk + threadIdx.x;
        dst[2 * k] = sr;
        dst[2 * k + 1] = si;
      }
    }
  }
};

// Order in which a launch over grp_cnt G' groups by T = gridDim.x /
// grp_cnt G-vector tiles runs its tiles.  Blocks are handed to the XCDs
// round-robin (block x runs on XCD x % 8) and each XCD has its own L2, so
// the blocks are renumbered for each XCD to run a contiguous range of the
// launch, and that range is walked in super-tiles of kSuperGrp groups by
// kSuperTile tiles: the blocks resident on one XCD at a time then share a
// few aqsmG column slabs and aqsn row ranges instead of spanning every
// group.  At 512/2/32768/20 this cuts the fetches of a per-term launch from
// 12.8 to 7.5 GB (FP64) and of the band-product launch from 8.7 to 5.0 GB,
// at the same launch times (the kernels are bound by arithmetic).  The
// order does not change any result: each tile writes its own record.
struct TileIdx {
  int tile, grp;  // G-vector tile, G' group within the launch
};
__device__ __forceinline__ TileIdx tile_order(int grp_cnt) {
  constexpr int kXcd = 8, kSuperGrp = 8, kSuperTile = 16;
  const int n = gridDim.x, q = n / kXcd, bx = blockIdx.x;
  const int x = bx < q * kXcd ? (bx % kXcd) * q + bx / kXcd : bx;
  const int T = n / grp_cnt;
  const int g0 = x / (kSuperGrp * T) * kSuperGrp;  // first group of x's band
  const int wdt = min(kSuperGrp, grp_cnt - g0);    // groups in that band
  const int l = x - g0 * T;
  const int r = l % (wdt * kSuperTile);
  return {l / (wdt * kSuperTile) * kSuperTile + r / wdt, g0 + r % wdt};
}

// A launch may cover only the G' groups [grp_lo, grp_lo + grp_cnt) (one
// chunk of a pipelined host->device copy).  Its blocks take the index, and
// so the record, that the same tile has in the launch over all groups
// (b = ig_grp * ngrp + igp_grp), so the fixed-order reduction gives the same
// bits however the columns were split.
__device__ __forceinline__ int full_grid_block(int NG, int grp_lo,
                                               int grp_cnt) {
  const int ngrp = (NG + kWave - 1) / kWave;
  const TileIdx t = tile_order(grp_cnt);
  return t.tile * ngrp + grp_lo + t.grp;
}

// ---------------------------------------------------------------------------
// Non-occupied bands, mixed precision.
// Same decomposition as the FP64 kernel; the per-term work is FP32:
//   dr = (xh - wrh) + (xl - wrl)      wx (split once per band by the
//                                     prepass) and wtilde.re as FP32 hi/lo
//                                     pairs, so the cancellation near
//                                     resonance is exact
//   d  = dr^2 + wi^2, r = v_rcp_f32(d), B = aqsn * A, S += B*conj(wdiff)*r
// with the FP32 run sum S over the lane's K G-vectors (Kahan-compensated
// when KAHAN, the compensation folded in when S is widened); conj(aqsm)*S
// and all later sums are FP64.  On gfx950 an FP32 VALU instruction issues in half
// the cycles of an FP64 one when two waves per SIMD are ready, so the band
// loop keeps its next band's operands (the aqsn run and the split
// frequencies, wave-uniform: scalar loads) in flight a whole band ahead.
//
// A lane's run for a band is redone in FP64 when one of its terms is near
// resonance (|wdiff|^2 below thresh: FP32 underflow, exact resonance, or
// the large terms whose FP32 error would dominate the sum) or, for the bgw
// mask, when a term's |wdiff|^2 lies within 4e-6 (relative) of the mask's
// threshold, where FP32 could decide it differently from FP64.  Such runs
// are not redone in place (that would stall the whole wave on one lane's
// FP64 work): the lane drops its FP32 run sum and appends (band, lane) to a
// per-wave work list in LDS; whenever 64 entries are waiting, the wave
// evaluates 64 runs in FP64, one per lane, and adds each one's
// wgt*conj(aqsm)*S to the evaluating lane's correction sums (the owner's
// weight is shared through LDS).  The list is drained at the end; with the
// per-band log it is drained every band, so each band's log holds its own
// corrections.  The order of the list is fixed by the data, so results are
// reproducible run to run.
constexpr int kMixQueue = 128;  // entries per wave: < 64 left + 64 new
constexpr float kMaskBand = 4e-6f;  // relative band around the bgw threshold

// The FP64 run of one work-list entry, weighted: wgt * conj(aqsm) * S64.
template <int K, bool MASK, typename Prob>
__device__ __forceinline__ void mixed_fp64_run(const Prob &p, int n1, int my_igp,
                                               int ig0, double wo,
                                               double (&cr)[3], double (&ci)[3]) {
  const int NC = p.NC;
  double sr[3] = {0.0, 0.0, 0.0}, si[3] = {0.0, 0.0, 0.0};
  const double x[3] = {p.wx[n1 * 3 + 0], p.wx[n1 * 3 + 1], p.wx[n1 * 3 + 2]};
  const auto an_row = p.aqsn + (size_t)n1 * NC;
  // unrolled by 4 so that the loads of several G-vectors are in flight at
  // once (the run's operands come from HBM or L2: 25.2 -> 25.1 ms for the
  // 4e-3 window, 34.2 -> 29.9 ms for the bgw-masked kernel at 512/2/32768/20)
#pragma unroll 4
  for (int k = 0; k < K; ++k) {
    const int ig = ig0 + k;
    if (ig >= NC) break;
    const size_t off = (size_t)my_igp * NC + ig;
    const double2 w = ld(p.wtilde + off);
    const double2 e = ld(p.ieps + off);
    const double2 a = ld(an_row + ig);
    const double A_r = w.x * e.x - w.y * e.y;
    const double A_i = w.x * e.y + w.y * e.x;
    const double B_r = a.x * A_r - a.y * A_i;
    const double B_i = a.x * A_i + a.y * A_r;
    const double wy2 = w.y * w.y;
    double ddr[3], dd[3], rr[3];
#pragma unroll
    for (int iw = 0; iw < 3; ++iw) {
      ddr[iw] = x[iw] - w.x;
      dd[iw] = fma(ddr[iw], ddr[iw], wy2);
    }
    // one v_rcp_f64 for the three frequencies (IEEE divisions when their
    // product is not a positive normal number, e.g. an exact resonance)
    recip3(dd[0], dd[1], dd[2], rr[0], rr[1], rr[2]);
#pragma unroll
    for (int iw = 0; iw < 3; ++iw) {
      // the mask skips the term outright (a NaN term it drops adds nothing)
      if (MASK &&
          !(dd[iw] > kLimitTwo && (w.x * w.x + wy2) * rr[iw] < kLimitOne))
        continue;
      sr[iw] += (B_r * ddr[iw] - B_i * w.y) * rr[iw];
      si[iw] += (B_r * w.y + B_i * ddr[iw]) * rr[iw];
    }
  }
  const double2 am = ld(p.aqsmG + (size_t)n1 * p.NG + my_igp);
#pragma unroll
  for (int iw = 0; iw < 3; ++iw) {
    cr[iw] += wo * fma(am.x, sr[iw], am.y * si[iw]);
    ci[iw] += wo * fma(am.x, si[iw], -am.y * sr[iw]);
  }
}

template <int K, bool MASK, bool N1LOG, bool KAHAN, typename Prob>
__device__ __forceinline__ void unocc_mixed_block(
    const Prob &p, int n1_begin, int n1_end, int b,
    double *__restrict__ rec, double *__restrict__ n1log, float thresh,
    unsigned long long *__restrict__ fallback_count) {
  __shared__ unsigned queue_all[4][kMixQueue];
  __shared__ double wgt_all[4][kWave];
  // per-band log: with one wave per block (the default) four bands per
  // wave-level sum through LDS (4 KB), otherwise a butterfly per band
  __shared__ double log_slot[N1LOG ? 2 : 1][kBandLogChunk][4][2];
  __shared__ double2 log_stage[N1LOG ? 4 : 1][kWave];
  const BandLog<1> blog{log_slot, nullptr, n1log,
                        (size_t)b * (n1_end - n1_begin), n1_end - n1_begin};
  const BandLog<4> blog4{log_slot, log_stage, n1log,
                         (size_t)b * (n1_end - n1_begin), n1_end - n1_begin};
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = uniform(threadIdx.x / kWave);
  const int W = blockDim.x / kWave;
  const int NC = p.NC, NG = p.NG;
  const int ngrp = (NG + kWave - 1) / kWave;
  const int igp_grp = b % ngrp;
  const int ig_grp = b / ngrp;
  const int my_igp = igp_grp * kWave + lane;
  const int ig0 = uniform((ig_grp * W + wave) * K);
  unsigned *queue = queue_all[wave];
  double *wgts = wgt_all[wave];

  int igp = (my_igp < NG) ? p.igp_of[my_igp] : -1;
  const bool valid = igp >= 0;
  const double wgt = valid ? 0.5 * p.vcoul[igp] : 0.0;
  wgts[lane] = wgt;  // read by the lanes that evaluate this lane's entries

  float wrh[K], wrl[K], wi[K], wi2[K], Ar[K], Ai[K];
  float idm[K];  // MASK: 1 / the mask threshold max(limittwo, |wtilde|^2/limitone)
  bool lane_nf = false;  // a wtilde of the run is not finite: runs in FP64
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int ig = ig0 + k;
    if (valid && ig < NC) {
      const size_t off = (size_t)my_igp * NC + ig;
      const double2 w = ld(p.wtilde + off);
      const double2 e = ld(p.ieps + off);
      lane_nf |= !(isfinite(w.x) && isfinite(w.y));
      wrh[k] = (float)w.x;
      wrl[k] = (float)(w.x - (double)wrh[k]);
      wi[k] = (float)w.y;
      Ar[k] = (float)(w.x * e.x - w.y * e.y);
      Ai[k] = (float)(w.x * e.y + w.y * e.x);
      if (MASK)
        idm[k] = (float)(1.0 / fmax(kLimitTwo, (w.x * w.x + w.y * w.y) / kLimitOne));
    } else {
      // Inert slot: A = 0 and a far-away wtilde: |wdiff|^2 overflows to
      // +inf, r = 0, contribution 0, never near a threshold.
      wrh[k] = 1e30f;
      wrl[k] = 0.0f;
      wi[k] = 0.0f;
      Ar[k] = 0.0f;
      Ai[k] = 0.0f;
      if (MASK) idm[k] = 0.0f;
    }
    wi2[k] = wi[k] * wi[k];
  }

  // weighted sums: wgt * conj(aqsm) * S of the lane's FP32 runs plus the
  // weighted FP64 runs this lane evaluated from the work list
  double acc_r[3] = {0.0, 0.0, 0.0}, acc_i[3] = {0.0, 0.0, 0.0};
  unsigned long long nfb = 0;
  int qn = 0;  // entries waiting in this wave's list (wave-uniform)
  const auto aqsm_col = p.aqsmG + min(my_igp, NG - 1);
  // the prepass's FP32 split of the frequencies, per band
  // [xh0 xl0 xh1 xl1 xh2 xl2 f 0] (f nonzero: a frequency is not finite);
  // the zero-padded FP32 aqsn rows
  const auto xsplit = (decltype(p.aqsn32))(p.wx + 6 * (size_t)p.NB);
  const int ncp = padded_nc(NC);
  double2 am_next = (n1_begin < n1_end)
                        ? ld(aqsm_col + (size_t)n1_begin * NG)
                        : make_double2(0.0, 0.0);
  float2 an_cur[K], xs_cur[4];  // xs_cur[3].x: the band's non-finite flag
  if (n1_begin < n1_end) {
#pragma unroll
    for (int k = 0; k < K; ++k)
      an_cur[k] = ld(p.aqsn32 + ((size_t)n1_begin * ncp + ig0 + k));
#pragma unroll
    for (int iw = 0; iw < 4; ++iw) xs_cur[iw] = ld(xsplit + n1_begin * 4 + iw);
  }

  // Evaluates `cnt` (<= 64) entries from the top of the list in FP64, one
  // per lane, into the evaluating lanes' correction sums.
  // The lanes of a wave execute their LDS accesses in order and every read
  // here may alias the writes before it, so no fence is needed (a fence
  // would also stop the compiler from turning the wave-uniform aqsn loads
  // of the band loop into scalar loads).
  auto drain = [&](int cnt) {
    if (lane < cnt) {
      const unsigned item = queue[qn - cnt + lane];
      const int owner = (int)(item & (kWave - 1));
      const int n1 = (int)(item >> 6);
      mixed_fp64_run<K, MASK>(p, n1, igp_grp * kWave + owner, ig0,
                              wgts[owner], acc_r, acc_i);
      nfb += 3 * K;
    }
    qn -= cnt;
  };

  for (int n1 = n1_begin; n1 < n1_end; ++n1) {
    const double2 am = am_next;
    const int n1n = min(n1 + 1, n1_end - 1);
    am_next = ld(aqsm_col + (size_t)n1n * NG);
    float2 an_nxt[K], xs_nxt[4];
#pragma unroll
    for (int k = 0; k < K; ++k)
      an_nxt[k] = ld(p.aqsn32 + ((size_t)n1n * ncp + ig0 + k));
#pragma unroll
    for (int iw = 0; iw < 4; ++iw) xs_nxt[iw] = ld(xsplit + n1n * 4 + iw);

    float s_r[3] = {0.f, 0.f, 0.f}, s_i[3] = {0.f, 0.f, 0.f};
    float c_r[3] = {0.f, 0.f, 0.f}, c_i[3] = {0.f, 0.f, 0.f};  // KAHAN
    float dmin = 3.0e38f, bmin = 3.0e38f;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const float2 an = an_cur[k];
      const float Br = fmaf(an.x, Ar[k], -an.y * Ai[k]);
      const float Bi = fmaf(an.x, Ai[k], an.y * Ar[k]);
      const float nBiwi = -Bi * wi[k];
      const float Brwi = Br * wi[k];
      float d[3];
#pragma unroll
      for (int iw = 0; iw < 3; ++iw) {
        const float dr = (xs_cur[iw].x - wrh[k]) + (xs_cur[iw].y - wrl[k]);
        d[iw] = fmaf(dr, dr, wi2[k]);
        float r = rcp_f32(d[iw]);
        if (MASK) {
          // keep the term iff |wdiff|^2 > max(limittwo, |wtilde|^2 /
          // limitone) (the f90 mask, :364-366); track the distance to
          // that threshold so that runs near it are decided in FP64
          const float rel = fmaf(d[iw], idm[k], -1.0f);
          r = rel > 0.0f ? r : 0.0f;
          bmin = fminf(bmin, fabsf(rel));
        }
        if (KAHAN) {
          // compensated run sum: c holds minus the low part lost so far
          const float tr = fmaf(Br, dr, nBiwi) * r - c_r[iw];
          const float ti = fmaf(Bi, dr, Brwi) * r - c_i[iw];
          const float ur = s_r[iw] + tr, ui = s_i[iw] + ti;
          c_r[iw] = (ur - s_r[iw]) - tr;
          c_i[iw] = (ui - s_i[iw]) - ti;
          s_r[iw] = ur;
          s_i[iw] = ui;
        } else {
          s_r[iw] = fmaf(fmaf(Br, dr, nBiwi), r, s_r[iw]);
          s_i[iw] = fmaf(fmaf(Bi, dr, Brwi), r, s_i[iw]);
        }
      }
      dmin = fminf(dmin, fminf(d[0], fminf(d[1], d[2])));
    }
    double sr[3], si[3];
#pragma unroll
    for (int iw = 0; iw < 3; ++iw) {
      sr[iw] = (double)s_r[iw];
      si[iw] = (double)s_i[iw];
      if (KAHAN) {
        sr[iw] -= (double)c_r[iw];
        si[iw] -= (double)c_i[iw];
      }
    }
    // A run with a non-finite wtilde or frequency goes to FP64 too, whose
    // per-term form skips the terms the bgw mask drops (in FP32 a dropped
    // NaN term would still add 0 * NaN).  Only valid lanes: an invalid
    // lane's FP64 run would read past row NG.
    // (the flag tested as an integer: a scalar compare)
    const bool band_nf = __float_as_int(xs_cur[3].x) != 0;
    const bool near =
        valid && (lane_nf || band_nf || !(dmin >= thresh) ||
                  (MASK && !(bmin >= kMaskBand)));
    const unsigned long long m = __ballot(near);
    if (m) {  // wave-uniform
      if (near) {
        sr[0] = sr[1] = sr[2] = si[0] = si[1] = si[2] = 0.0;
        const int pos = qn + (int)__builtin_amdgcn_mbcnt_hi(
                                  (unsigned)(m >> 32),
                                  __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
        queue[pos] = ((unsigned)n1 << 6) | (unsigned)lane;
      }
      qn += __popcll(m);
    }
    const double wr = wgt * am.x, wi_ = wgt * am.y;  // wgt * aqsm
#pragma unroll
    for (int iw = 0; iw < 3; ++iw) {
      acc_r[iw] = fma(wr, sr[iw], fma(wi_, si[iw], acc_r[iw]));
      acc_i[iw] = fma(wr, si[iw], fma(-wi_, sr[iw], acc_i[iw]));
    }
    if (N1LOG) {
      // this band's log: the lane's FP32 run plus the FP64 runs of this
      // band, drained now (each evaluating lane adds what it evaluated)
      const double c2r = acc_r[2], c2i = acc_i[2];
      while (qn > 0) drain(min(qn, kWave));
      const double lr = fma(wr, sr[2], wi_ * si[2]) + (acc_r[2] - c2r);
      const double li = fma(wr, si[2], -wi_ * sr[2]) + (acc_i[2] - c2i);
      if (W == 1)  // block-uniform
        blog4.put(n1 - n1_begin, lr, li);
      else
        blog.put(n1 - n1_begin, lr, li);
    } else if (qn >= kWave) {
      drain(kWave);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) an_cur[k] = an_nxt[k];
#pragma unroll
    for (int iw = 0; iw < 4; ++iw) xs_cur[iw] = xs_nxt[iw];
  }
  while (qn > 0) drain(min(qn, kWave));

  double v[6];
#pragma unroll
  for (int iw = 0; iw < 3; ++iw) {
    v[2 * iw] = acc_r[iw];
    v[2 * iw + 1] = acc_i[iw];
  }
  block_store<6>(v, rec);
  if (threadIdx.x >= 6 && threadIdx.x < kSlots) rec[threadIdx.x] = 0.0;
  if (fallback_count) {
    // Integer count: the order of the atomics does not change the value.
    unsigned long long c = nfb;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) c += __shfl_xor(c, m, kWave);
    if (lane == 0 && c) atomicAdd(fallback_count, c);
  }
}


// The log-free variants with K <= 8 are held to at least 3 waves per SIMD
// (168 VGPRs): the ILP-first scheduler would otherwise take the masked one
// to 190 VGPRs and 2 waves.  The other variants need more registers.
template <int K, bool MASK, bool N1LOG, bool KAHAN>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(N1LOG || K > 8 ? 1 : 3)))
void unocc_mixed_kernel(
    DevProblem p, int n1_begin, int n1_end, int grp_lo, int grp_cnt,
    double *__restrict__ partials, double *__restrict__ n1log, float thresh,
    unsigned long long *__restrict__ fallback_count) {
  const int b = full_grid_block(p.NG, grp_lo, grp_cnt);
  unocc_mixed_block<K, MASK, N1LOG, KAHAN>(p, n1_begin, n1_end, b,
                                           partials + (size_t)b * kSlots,
                                           n1log, thresh, fallback_count);
}

}  // namespace
}  // namespace gpp
