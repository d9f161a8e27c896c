// gpp_device.hip — gfx950 kernels of the GPP sigma sum.
//
// Work decomposition (non-occupied bands, the ~99.6% hot loop of
// gppKer.cpp:316-345):
//   * a wavefront owns 64 G' columns (one my_igp per lane) and a run of
//     K consecutive G-vectors; every lane keeps, for its my_igp and those
//     K G-vectors, wtilde and A = wtilde * I_eps in registers for the
//     whole band loop (they do not depend on n1);
//   * the band loop n1 runs inside the wave.  aqsn[n1][ig-run] is the same
//     for all 64 lanes (scalar loads, SGPR operands); conj(aqsm[n1][igp])
//     differs per lane but is constant over the ig run, so it multiplies
//     the run's partial sum once per band instead of once per term;
//   * the three frequencies share one reciprocal: 1/(d0 d1 d2) by
//     v_rcp_f64 plus one Newton step, then six multiplies give 1/d_iw
//     (batch inversion).  A zero/denormal/non-finite product falls back to
//     three plain divisions so that IEEE inf/NaN behave as in the
//     reference;
//   * lane partials are weighted by 0.5*vcoul[igp], reduced across the wave
//     with cross-lane shuffles, across the block through LDS, and written
//     to one record per block; a second kernel sums the records in a fixed
//     order.  No atomics touch the sums, so results are bit-reproducible.
// The occupied bands (n1 < nvband, gppKer.cpp:122-146 / 16-62) run in a
// separate FP64 kernel with one block per (my_igp, ig chunk).
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

// ---------------------------------------------------------------------------
// Prepass: resolve igp per my_igp and expand/clamp the frequency grid.
// Index rule of gppKer.cpp:287-293: indigp = inv_igp_index[my_igp];
// igp = indinv[indigp], except indigp == NC -> igp = NC-1.  Rows whose igp
// falls outside [0, NC) are skipped (the reference's guard admits igp == NC
// and then reads one element past the row; here that row is skipped).
// Clamp of gppKer.cpp:279: |wx| < to1 -> to1.
__global__ void prepass_kernel(const int *__restrict__ inv_igp_index,
                               const int *__restrict__ indinv,
                               const double *__restrict__ wx_in,
                               int wx_per_band, int NB, int NC, int NG,
                               int *__restrict__ igp_of,
                               double *__restrict__ wx_out,
                               unsigned long long *__restrict__ counter) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  // Cleared by a kernel store, not a memset, so the sequence stays valid
  // inside a captured graph.
  if (i == 0) *counter = 0ull;
  if (i < NG) {
    const int indigp = inv_igp_index[i];
    int igp;
    if (indigp == NC)
      igp = NC - 1;
    else if (indigp < 0 || indigp > NC)
      igp = -1;
    else
      igp = indinv[indigp];
    if (igp < 0 || igp >= NC) igp = -1;
    igp_of[i] = igp;
  }
  if (i < NB * 3) {
    double w = wx_per_band ? wx_in[i] : wx_in[i % 3];
    if (fabs(w) < kTo1) w = kTo1;
    wx_out[i] = w;
  }
}

__global__ void to_fp32_kernel(const double2 *__restrict__ in,
                               float2 *__restrict__ out, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    const double2 v = in[i];
    out[i] = make_float2((float)v.x, (float)v.y);
  }
}

// ---------------------------------------------------------------------------
// Non-occupied bands, FP64.
//
// Block b: igp group g = b % ngrp (64 my_igp), ig group = b / ngrp; wave w
// of the block takes G-vectors [ (ig_group*W + w)*K, +K ).
// Per (n1, my_igp, ig) with wdiff = wx - wtilde = dr - i*wi:
//   B   = aqsn[n1][ig] * A,          A = wtilde * I_eps
//   Y   = B * conj(wdiff) = (Br*dr - Bi*wi) + i(Br*wi + Bi*dr)
//   S  += Y / |wdiff|^2              (per iw, over the K G-vectors)
// and after the run: acc += conj(aqsm[n1][igp]) * S.
// Equals conj(aqsm)*aqsn*delw*I_eps of gppKer.cpp:328-337 term by term.
template <int K, bool MASK>
__global__ __launch_bounds__(256) void unocc_fp64_kernel(
    DevProblem p, int n1_begin, int n1_end, double *__restrict__ partials,
    int rec_base, int block_offset) {
  const int b = blockIdx.x + block_offset;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = uniform(threadIdx.x / kWave);
  const int W = blockDim.x / kWave;
  const int NC = p.NC, NG = p.NG;
  const int ngrp = (NG + kWave - 1) / kWave;
  const int igp_grp = b % ngrp;
  const int ig_grp = b / ngrp;
  const int my_igp = igp_grp * kWave + lane;
  const int ig0 = uniform((ig_grp * W + wave) * K);

  int igp = (my_igp < NG) ? p.igp_of[my_igp] : -1;
  const bool valid = igp >= 0;
  const int igp_c = valid ? igp : 0;

  double wr[K], wi[K], wi2[K], Ar[K], Ai[K];
  double wt2[K];  // |wtilde|^2, only for the mask
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int ig = ig0 + k;
    if (valid && ig < NC) {
      const size_t off = (size_t)my_igp * NC + ig;
      const double2 w = p.wtilde[off];
      const double2 e = p.ieps[off];
      wr[k] = w.x;
      wi[k] = w.y;
      Ar[k] = w.x * e.x - w.y * e.y;
      Ai[k] = w.x * e.y + w.y * e.x;
    } else {
      // Inert slot: A = 0 and |wdiff|^2 = wx^2 >= to1^2 > 0.
      wr[k] = 0.0;
      wi[k] = 0.0;
      Ar[k] = 0.0;
      Ai[k] = 0.0;
    }
    wi2[k] = wi[k] * wi[k];
    wt2[k] = wr[k] * wr[k] + wi2[k];
  }

  double acc_r[3] = {0.0, 0.0, 0.0}, acc_i[3] = {0.0, 0.0, 0.0};
  const double2 *aqsm_col = p.aqsm + igp_c;
  double2 am_next = (n1_begin < n1_end) ? aqsm_col[(size_t)n1_begin * NC]
                                        : make_double2(0.0, 0.0);

  for (int n1 = n1_begin; n1 < n1_end; ++n1) {
    const double2 am = am_next;
    // Prefetch next band's aqsm (index clamped: no branch in the loop).
    am_next = aqsm_col[(size_t)min(n1 + 1, n1_end - 1) * NC];
    const double x0 = p.wx[n1 * 3 + 0];
    const double x1 = p.wx[n1 * 3 + 1];
    const double x2 = p.wx[n1 * 3 + 2];
    const double2 *an_row = p.aqsn + (size_t)n1 * NC;

    double s_r0 = 0.0, s_i0 = 0.0, s_r1 = 0.0, s_i1 = 0.0, s_r2 = 0.0,
           s_i2 = 0.0;
    bool bad = false;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int igk = min(ig0 + k, NC - 1);
      const double2 an = an_row[igk];
      const double Br = an.x * Ar[k] - an.y * Ai[k];
      const double Bi = an.x * Ai[k] + an.y * Ar[k];
      const double nBiwi = -Bi * wi[k];
      const double Brwi = Br * wi[k];
      const double dr0 = x0 - wr[k];
      const double dr1 = x1 - wr[k];
      const double dr2 = x2 - wr[k];
      const double d0 = fma(dr0, dr0, wi2[k]);
      const double d1 = fma(dr1, dr1, wi2[k]);
      const double d2 = fma(dr2, dr2, wi2[k]);
      const double p01 = d0 * d1;
      const double pp = p01 * d2;
      bad |= !is_pos_normal(pp);
      double q = rcp_f64(pp);
      const double e = fma(-pp, q, 1.0);
      q = fma(q, e, q);
      double r2 = q * p01;
      const double t = q * d2;
      double r0 = t * d1;
      double r1 = t * d0;
      if (MASK) {
        // gppkernel.deslippe.09.30.2014.f90:364-366 /
        // BGW-Kernels/gppKer.cpp:406: keep the term only if
        // |wdiff|^2 > limittwo and |delw|^2 = |wtilde|^2/|wdiff|^2 < limitone.
        r0 = (d0 > kLimitTwo && wt2[k] * r0 < kLimitOne) ? r0 : 0.0;
        r1 = (d1 > kLimitTwo && wt2[k] * r1 < kLimitOne) ? r1 : 0.0;
        r2 = (d2 > kLimitTwo && wt2[k] * r2 < kLimitOne) ? r2 : 0.0;
      }
      s_r0 = fma(fma(Br, dr0, nBiwi), r0, s_r0);
      s_i0 = fma(fma(Bi, dr0, Brwi), r0, s_i0);
      s_r1 = fma(fma(Br, dr1, nBiwi), r1, s_r1);
      s_i1 = fma(fma(Bi, dr1, Brwi), r1, s_i1);
      s_r2 = fma(fma(Br, dr2, nBiwi), r2, s_r2);
      s_i2 = fma(fma(Bi, dr2, Brwi), r2, s_i2);
    }
    if (__builtin_expect(bad, 0)) {
      // Some product d0*d1*d2 was zero, denormal, infinite or NaN: redo this
      // lane's run with one division per term so that exact resonances give
      // the reference's inf/NaN in their own frequency only.
      s_r0 = s_i0 = s_r1 = s_i1 = s_r2 = s_i2 = 0.0;
      for (int k = 0; k < K; ++k) {
        const int igk = min(ig0 + k, NC - 1);
        const double2 an = an_row[igk];
        const double Br = an.x * Ar[k] - an.y * Ai[k];
        const double Bi = an.x * Ai[k] + an.y * Ar[k];
        const double xs[3] = {x0, x1, x2};
        double sr[3], si[3];
        for (int iw = 0; iw < 3; ++iw) {
          const double dr = xs[iw] - wr[k];
          const double d = fma(dr, dr, wi2[k]);
          double r = 1.0 / d;
          if (MASK) r = (d > kLimitTwo && wt2[k] * r < kLimitOne) ? r : 0.0;
          sr[iw] = (Br * dr - Bi * wi[k]) * r;
          si[iw] = (Br * wi[k] + Bi * dr) * r;
        }
        s_r0 += sr[0];
        s_i0 += si[0];
        s_r1 += sr[1];
        s_i1 += si[1];
        s_r2 += sr[2];
        s_i2 += si[2];
      }
    }
    // acc += conj(am) * S
    acc_r[0] = fma(am.x, s_r0, fma(am.y, s_i0, acc_r[0]));
    acc_i[0] = fma(am.x, s_i0, fma(-am.y, s_r0, acc_i[0]));
    acc_r[1] = fma(am.x, s_r1, fma(am.y, s_i1, acc_r[1]));
    acc_i[1] = fma(am.x, s_i1, fma(-am.y, s_r1, acc_i[1]));
    acc_r[2] = fma(am.x, s_r2, fma(am.y, s_i2, acc_r[2]));
    acc_i[2] = fma(am.x, s_i2, fma(-am.y, s_r2, acc_i[2]));
  }

  // sch_array[iw] = 0.5 * sum; achtemp[iw] += sch_array[iw] * vcoul[igp]
  const double wgt = valid ? 0.5 * p.vcoul[igp] : 0.0;
  double v[6];
#pragma unroll
  for (int iw = 0; iw < 3; ++iw) {
    v[2 * iw] = wgt * acc_r[iw];
    v[2 * iw + 1] = wgt * acc_i[iw];
  }
  double *rec = partials + (size_t)(rec_base + b) * kSlots;
  block_store<6>(v, rec);
  if (threadIdx.x >= 6 && threadIdx.x < kSlots) rec[threadIdx.x] = 0.0;
}

// ---------------------------------------------------------------------------
// Non-occupied bands, mixed precision.
// Same decomposition.  B, Y, |wdiff|^2 and the reciprocal are FP32; the
// difference dr = wx - wr is taken in FP64 and then rounded, so the
// catastrophic cancellation near resonance stays in FP64.  Terms with
// |wdiff|^2 < thresh are re-evaluated fully in FP64 (from the FP64 arrays)
// and accumulated in FP64.  The per-band run sum S is FP32 over K terms;
// everything from conj(aqsm)*S on is FP64.
template <int K>
__global__ __launch_bounds__(256) void unocc_mixed_kernel(
    DevProblem p, int n1_begin, int n1_end, double *__restrict__ partials,
    int rec_base, int block_offset, float thresh,
    unsigned long long *__restrict__ fallback_count) {
  const int b = blockIdx.x + block_offset;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = uniform(threadIdx.x / kWave);
  const int W = blockDim.x / kWave;
  const int NC = p.NC, NG = p.NG;
  const int ngrp = (NG + kWave - 1) / kWave;
  const int igp_grp = b % ngrp;
  const int ig_grp = b / ngrp;
  const int my_igp = igp_grp * kWave + lane;
  const int ig0 = uniform((ig_grp * W + wave) * K);

  int igp = (my_igp < NG) ? p.igp_of[my_igp] : -1;
  const bool valid = igp >= 0;
  const int igp_c = valid ? igp : 0;

  double wr[K];
  float wi[K], wi2[K], Ar[K], Ai[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int ig = ig0 + k;
    if (valid && ig < NC) {
      const size_t off = (size_t)my_igp * NC + ig;
      const double2 w = p.wtilde[off];
      const double2 e = p.ieps[off];
      wr[k] = w.x;
      wi[k] = (float)w.y;
      Ar[k] = (float)(w.x * e.x - w.y * e.y);
      Ai[k] = (float)(w.x * e.y + w.y * e.x);
    } else {
      wr[k] = 0.0;
      wi[k] = 0.0f;
      Ar[k] = 0.0f;
      Ai[k] = 0.0f;
    }
    wi2[k] = wi[k] * wi[k];
  }

  double acc_r[3] = {0.0, 0.0, 0.0}, acc_i[3] = {0.0, 0.0, 0.0};
  unsigned long long nfb = 0;
  const double2 *aqsm_col = p.aqsm + igp_c;
  double2 am_next = (n1_begin < n1_end) ? aqsm_col[(size_t)n1_begin * NC]
                                        : make_double2(0.0, 0.0);

  for (int n1 = n1_begin; n1 < n1_end; ++n1) {
    const double2 am = am_next;
    if (n1 + 1 < n1_end) am_next = aqsm_col[(size_t)(n1 + 1) * NC];
    const double x[3] = {p.wx[n1 * 3 + 0], p.wx[n1 * 3 + 1],
                         p.wx[n1 * 3 + 2]};
    const float2 *an_row = p.aqsn32 + (size_t)n1 * NC;

    float s_r[3] = {0.f, 0.f, 0.f}, s_i[3] = {0.f, 0.f, 0.f};
    double f_r[3] = {0.0, 0.0, 0.0}, f_i[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int igk = min(ig0 + k, NC - 1);
      const float2 an = an_row[igk];
      const float Br = an.x * Ar[k] - an.y * Ai[k];
      const float Bi = an.x * Ai[k] + an.y * Ar[k];
      const float nBiwi = -Bi * wi[k];
      const float Brwi = Br * wi[k];
#pragma unroll
      for (int iw = 0; iw < 3; ++iw) {
        const float dr = (float)(x[iw] - wr[k]);
        const float d = fmaf(dr, dr, wi2[k]);
        float r = rcp_f32(d);
        if (__builtin_expect(!(d >= thresh), 0)) {
          // Near resonance (or NaN): redo this term in FP64.
          r = 0.0f;
          const size_t off = (size_t)my_igp * NC + (ig0 + k);
          const double2 w = p.wtilde[off];
          const double2 e = p.ieps[off];
          const double2 a = p.aqsn[(size_t)n1 * NC + igk];
          const double A_r = w.x * e.x - w.y * e.y;
          const double A_i = w.x * e.y + w.y * e.x;
          const double B_r = a.x * A_r - a.y * A_i;
          const double B_i = a.x * A_i + a.y * A_r;
          const double ddr = x[iw] - w.x;
          const double dd = ddr * ddr + w.y * w.y;
          const double rr = 1.0 / dd;
          f_r[iw] = fma((B_r * ddr - B_i * w.y), rr, f_r[iw]);
          f_i[iw] = fma((B_r * w.y + B_i * ddr), rr, f_i[iw]);
          ++nfb;
        }
        s_r[iw] = fmaf(fmaf(Br, dr, nBiwi), r, s_r[iw]);
        s_i[iw] = fmaf(fmaf(Bi, dr, Brwi), r, s_i[iw]);
      }
    }
#pragma unroll
    for (int iw = 0; iw < 3; ++iw) {
      const double sr = (double)s_r[iw] + f_r[iw];
      const double si = (double)s_i[iw] + f_i[iw];
      acc_r[iw] = fma(am.x, sr, fma(am.y, si, acc_r[iw]));
      acc_i[iw] = fma(am.x, si, fma(-am.y, sr, acc_i[iw]));
    }
  }

  const double wgt = valid ? 0.5 * p.vcoul[igp] : 0.0;
  double v[6];
#pragma unroll
  for (int iw = 0; iw < 3; ++iw) {
    v[2 * iw] = wgt * acc_r[iw];
    v[2 * iw + 1] = wgt * acc_i[iw];
  }
  double *rec = partials + (size_t)(rec_base + b) * kSlots;
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

// ---------------------------------------------------------------------------
// Occupied bands (n1 < nvband): ssxt_scht_solver / flagOCC_solver of
// gppKer.cpp:16-62,122-146 (SEM gppker, gppker-strict) or the occupied
// branch of gppkernel.deslippe.09.30.2014.f90:271-334 (SEM bgw).
constexpr int kOccThreads = 256;
constexpr int kOccChunk = 2048;  // G-vectors per block

template <int SEM>
__global__ __launch_bounds__(kOccThreads) void occ_kernel(
    DevProblem p, int n1_end, double *__restrict__ partials, int rec_base) {
  const int NC = p.NC, NG = p.NG;
  const int nch = (NC + kOccChunk - 1) / kOccChunk;
  const int my_igp = blockIdx.x / nch;
  const int ch = blockIdx.x % nch;
  const int igp = p.igp_of[my_igp];
  double *rec = partials + (size_t)(rec_base + blockIdx.x) * kSlots;
  if (igp < 0) {  // skipped row (block-uniform)
    if (threadIdx.x < kSlots) rec[threadIdx.x] = 0.0;
    return;
  }
  double sch_r[3] = {0, 0, 0}, sch_i[3] = {0, 0, 0};
  double ssx_r[3] = {0, 0, 0}, ssx_i[3] = {0, 0, 0};
  const int ig_end = min(NC, (ch + 1) * kOccChunk);
  for (int ig = ch * kOccChunk + threadIdx.x; ig < ig_end;
       ig += kOccThreads) {
    const size_t off = (size_t)my_igp * NC + ig;
    const double2 w = p.wtilde[off];
    const double2 e = p.ieps[off];
    // I_eps used for sch and for the cutoff.  gppKer.cpp:139 passes
    // I_eps_array[my_igp*ngpown+ig]; the strict mode keeps that index.
    const double2 ex =
        (SEM == kSemGppkerStrict) ? p.ieps[(size_t)my_igp * NG + ig] : e;
    // wtilde2 = wtilde^2, Omega2 = wtilde2 * I_eps (gppKer.cpp:132-133)
    const double w2r = w.x * w.x - w.y * w.y;
    const double w2i = 2.0 * w.x * w.y;
    const double omr = w2r * e.x - w2i * e.y;
    const double omi = w2r * e.y + w2i * e.x;
    const double cut2 = kSexcut * kSexcut * (ex.x * ex.x + ex.y * ex.y);
    for (int n1 = 0; n1 < n1_end; ++n1) {
      const double2 am = p.aqsm[(size_t)n1 * NC + igp];
      const double2 an = p.aqsn[(size_t)n1 * NC + ig];
      // matngmatmgp = aqsn[n1][ig] * conj(aqsm[n1][igp])
      const double mr = an.x * am.x + an.y * am.y;
      const double mi = an.y * am.x - an.x * am.y;
#pragma unroll
      for (int iw = 0; iw < 3; ++iw) {
        const double wxt = p.wx[n1 * 3 + iw];
        // wdiff = wxt - wtilde = (dr, -w.y); conj(wdiff) = (dr, w.y)
        const double dr = wxt - w.x;
        const double wdiffr = dr * dr + w.y * w.y;
        const double rden = 1.0 / wdiffr;
        const double dlr = (w.x * dr - w.y * w.y) * rden;
        const double dli = (w.x * w.y + w.y * dr) * rden;
        const double delwr = dlr * dlr + dli * dli;
        double schr = 0.0, schi = 0.0, ssr = 0.0, ssi = 0.0;
        if (wdiffr > kLimitTwo && delwr < kLimitOne) {
          schr = dlr * ex.x - dli * ex.y;
          schi = dlr * ex.y + dli * ex.x;
          if (SEM == kSemBgw) {
            // cden = wxt^2 - wtilde2 (f90:303)
            const double cr = wxt * wxt - w2r;
            const double ci = -w2i;
            const double rd = 1.0 / (cr * cr + ci * ci);
            ssr = (omr * cr + omi * ci) * rd;
            ssi = (omi * cr - omr * ci) * rd;
          } else {
            // cden = wxt^2 (real), ssx = Omega2 * conj(cden) / |cden|^2
            const double c = wxt * wxt;
            const double rd = 1.0 / (c * c);
            ssr = omr * c * rd;
            ssi = omi * c * rd;
          }
        } else if (delwr > kTo1) {
          // cden = 4 * wtilde2 * (delw + 0.5)
          const double ar = dlr + 0.5, ai = dli;
          const double cr = 4.0 * (w2r * ar - w2i * ai);
          const double ci = 4.0 * (w2r * ai + w2i * ar);
          const double rd = 1.0 / (cr * cr + ci * ci);
          // ssx = -Omega2 * conj(cden) * rden * delw
          const double tr = -(omr * cr + omi * ci) * rd;
          const double ti = -(omi * cr - omr * ci) * rd;
          ssr = tr * dlr - ti * dli;
          ssi = tr * dli + ti * dlr;
        }
        if (SEM == kSemBgw) {
          // f90:321: |ssx| > sexcut*|I_eps| and wxt < 0 -> ssx = 0
          if (wxt < 0.0 && (ssr * ssr + ssi * ssi) > cut2) {
            ssr = 0.0;
            ssi = 0.0;
          }
        }
        // (gppKer.cpp:58 tests abs(wxt) < 0, which never holds.)
        ssx_r[iw] += mr * ssr - mi * ssi;
        ssx_i[iw] += mr * ssi + mi * ssr;
        sch_r[iw] += mr * schr - mi * schi;
        sch_i[iw] += mr * schi + mi * schr;
      }
    }
  }
  const double vc = p.vcoul[igp];
  const double wsch = 0.5 * vc;                                // :310,355
  const double wssx = (SEM == kSemBgw ? -kOcc : kOcc) * vc;    // :352 / f90:393
  double v[12];
#pragma unroll
  for (int iw = 0; iw < 3; ++iw) {
    v[2 * iw] = wsch * sch_r[iw];
    v[2 * iw + 1] = wsch * sch_i[iw];
    v[6 + 2 * iw] = wssx * ssx_r[iw];
    v[6 + 2 * iw + 1] = wssx * ssx_i[iw];
  }
  block_store<12>(v, rec);
}

// ---------------------------------------------------------------------------
// Fixed-order sum of `nrec` records of kSlots doubles -> out[kSlots].
// One block: thread t sums records t, t+256, ... of one slot group in
// order; then a fixed pairwise tree over the 256 partials.
constexpr int kRedThreads = 256;
__global__ __launch_bounds__(kRedThreads) void reduce_kernel(
    const double *__restrict__ partials, int nrec, double *__restrict__ out) {
  __shared__ double buf[kRedThreads];
  for (int s = 0; s < kSlots; ++s) {
    double acc = 0.0, c = 0.0;  // Neumaier compensation
    for (int r = threadIdx.x; r < nrec; r += kRedThreads) {
      const double x = partials[(size_t)r * kSlots + s];
      const double t = acc + x;
      c += (fabs(acc) >= fabs(x)) ? ((acc - t) + x) : ((x - t) + acc);
      acc = t;
    }
    buf[threadIdx.x] = acc + c;
    __syncthreads();
    for (int h = kRedThreads / 2; h > 0; h >>= 1) {
      if (threadIdx.x < h) buf[threadIdx.x] += buf[threadIdx.x + h];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[s] = buf[0];
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Static remainder achstemp, gppker form (gppKer.cpp:64-120 with the
// shared-accumulator race removed): per band and my_igp, the igp term
//   0.5 * vcoul[igp] * aqsn[n1][igp] * conj(aqsm[n1][igp]) * I_eps[my_igp][igp]
// taken when |I_eps[my_igp][igp]| > to1 (:96-97).  Record slot 0/1.
constexpr int kAchThreads = 256;
__global__ __launch_bounds__(kAchThreads) void achstemp_gppker_kernel(
    DevProblem p, double *__restrict__ partials, int rec_base) {
  const int my_igp = blockIdx.x * kAchThreads + threadIdx.x;
  double vr = 0.0, vi = 0.0;
  if (my_igp < p.NG) {
    const int igp = p.igp_of[my_igp];
    if (igp >= 0) {
      const double2 s = p.ieps[(size_t)my_igp * p.NC + igp];
      if (sqrt(s.x * s.x + s.y * s.y) > kTo1) {
        double tr = 0.0, ti = 0.0;
        for (int n1 = 0; n1 < p.NB; ++n1) {
          const double2 am = p.aqsm[(size_t)n1 * p.NC + igp];
          const double2 an = p.aqsn[(size_t)n1 * p.NC + igp];
          tr += an.x * am.x + an.y * am.y;
          ti += an.y * am.x - an.x * am.y;
        }
        const double w = 0.5 * p.vcoul[igp];
        vr = w * (tr * s.x - ti * s.y);
        vi = w * (tr * s.y + ti * s.x);
      }
    }
  }
  double v[2] = {vr, vi};
  double *rec = partials + (size_t)(rec_base + blockIdx.x) * kSlots;
  block_store<2>(v, rec);
  if (threadIdx.x >= 2 && threadIdx.x < kSlots) rec[threadIdx.x] = 0.0;
}

// Static remainder, F90 form (gppkernel.deslippe.09.30.2014.f90:195-229):
//   -0.5 * vcoul[igp] * conj(aqsm[n1][igp]) * sum_ig aqsn[n1][ig] *
//   I_eps[my_igp][ig]
// One block per (my_igp, band chunk); lanes over ig.
constexpr int kAchBands = 8;
__global__ __launch_bounds__(kAchThreads) void achstemp_bgw_kernel(
    DevProblem p, double *__restrict__ partials, int rec_base) {
  const int nbc = (p.NB + kAchBands - 1) / kAchBands;
  const int my_igp = blockIdx.x / nbc;
  const int n1b = (blockIdx.x % nbc) * kAchBands;
  const int n1e = min(p.NB, n1b + kAchBands);
  const int igp = p.igp_of[my_igp];
  double *rec = partials + (size_t)(rec_base + blockIdx.x) * kSlots;
  if (igp < 0) {
    if (threadIdx.x < kSlots) rec[threadIdx.x] = 0.0;
    return;
  }
  double vr = 0.0, vi = 0.0;
  for (int n1 = n1b; n1 < n1e; ++n1) {
    double sr = 0.0, si = 0.0;
    for (int ig = threadIdx.x; ig < p.NC; ig += kAchThreads) {
      const double2 an = p.aqsn[(size_t)n1 * p.NC + ig];
      const double2 e = p.ieps[(size_t)my_igp * p.NC + ig];
      sr += an.x * e.x - an.y * e.y;
      si += an.x * e.y + an.y * e.x;
    }
    const double2 am = p.aqsm[(size_t)n1 * p.NC + igp];
    // -(conj(am) * s)
    vr -= am.x * sr + am.y * si;
    vi -= am.x * si - am.y * sr;
  }
  const double w = 0.5 * p.vcoul[igp];
  double v[2] = {w * vr, w * vi};
  block_store<2>(v, rec);
  if (threadIdx.x >= 2 && threadIdx.x < kSlots) rec[threadIdx.x] = 0.0;
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers.

hipError_t launch_prepass(const int *inv_igp_index, const int *indinv,
                          const double *wx_in, int wx_per_band, int NB,
                          int NC, int NG, int *igp_of, double *wx_out,
                          unsigned long long *counter, hipStream_t s) {
  int n = NG > NB * 3 ? NG : NB * 3;
  if (n < 1) n = 1;
  prepass_kernel<<<(n + 255) / 256, 256, 0, s>>>(inv_igp_index, indinv, wx_in,
                                                 wx_per_band, NB, NC, NG,
                                                 igp_of, wx_out, counter);
  return hipGetLastError();
}

hipError_t launch_aqsn_to_fp32(const double2 *aqsn, float2 *out, size_t n,
                               hipStream_t s) {
  if (n == 0) return hipSuccess;
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  to_fp32_kernel<<<(unsigned)blocks, 256, 0, s>>>(aqsn, out, n);
  return hipGetLastError();
}

int unocc_grid(const DevProblem &p, const TileCfg &t) {
  if (p.NG <= 0 || p.NC <= 0) return 0;
  const int ngrp = (p.NG + kWave - 1) / kWave;
  const int per_blk = t.k_ig * t.waves;
  const int nig = (p.NC + per_blk - 1) / per_blk;
  return ngrp * nig;
}

template <int K>
static hipError_t launch_unocc_k(const DevProblem &p, int n1b, int n1e,
                                 int precision, int sem, int waves,
                                 double thresh, double *partials,
                                 unsigned long long *fb, int grid,
                                 hipStream_t s) {
  const dim3 blk(kWave * waves);
  if (precision == 0) {
    if (sem == kSemBgw)
      unocc_fp64_kernel<K, true><<<grid, blk, 0, s>>>(p, n1b, n1e, partials,
                                                       0, 0);
    else
      unocc_fp64_kernel<K, false><<<grid, blk, 0, s>>>(p, n1b, n1e, partials,
                                                        0, 0);
  } else {
    if (sem == kSemBgw) return hipErrorNotSupported;
    unocc_mixed_kernel<K><<<grid, blk, 0, s>>>(p, n1b, n1e, partials, 0, 0,
                                               (float)thresh, fb);
  }
  return hipGetLastError();
}

hipError_t launch_unocc(const DevProblem &p, int n1_begin, int n1_end,
                        int precision, int sem, const TileCfg &t,
                        double res_thresh, double *partials,
                        unsigned long long *fallback_count, hipStream_t s) {
  const int grid = unocc_grid(p, t);
  if (grid == 0) return hipSuccess;
  if (t.waves < 1 || t.waves > 16) return hipErrorInvalidValue;
  switch (t.k_ig) {
    case 4:
      return launch_unocc_k<4>(p, n1_begin, n1_end, precision, sem, t.waves,
                               res_thresh, partials, fallback_count, grid, s);
    case 8:
      return launch_unocc_k<8>(p, n1_begin, n1_end, precision, sem, t.waves,
                               res_thresh, partials, fallback_count, grid, s);
    case 16:
      return launch_unocc_k<16>(p, n1_begin, n1_end, precision, sem, t.waves,
                                res_thresh, partials, fallback_count, grid, s);
    case 32:
      return launch_unocc_k<32>(p, n1_begin, n1_end, precision, sem, t.waves,
                                res_thresh, partials, fallback_count, grid, s);
    default:
      return hipErrorInvalidValue;
  }
}

int occ_grid(const DevProblem &p) {
  if (p.NG <= 0 || p.NC <= 0) return 0;
  return p.NG * ((p.NC + kOccChunk - 1) / kOccChunk);
}

hipError_t launch_occ(const DevProblem &p, int n1_end, int sem,
                      double *partials, hipStream_t s) {
  const int grid = occ_grid(p);
  if (grid == 0) return hipSuccess;
  switch (sem) {
    case kSemGppker:
      occ_kernel<kSemGppker><<<grid, kOccThreads, 0, s>>>(p, n1_end, partials,
                                                          0);
      break;
    case kSemGppkerStrict:
      occ_kernel<kSemGppkerStrict><<<grid, kOccThreads, 0, s>>>(
          p, n1_end, partials, 0);
      break;
    case kSemBgw:
      occ_kernel<kSemBgw><<<grid, kOccThreads, 0, s>>>(p, n1_end, partials, 0);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_reduce(const double *partials, int nrec, double *out,
                         hipStream_t s) {
  reduce_kernel<<<1, kRedThreads, 0, s>>>(partials, nrec, out);
  return hipGetLastError();
}

int achstemp_records(const DevProblem &p, int sem) {
  if (p.NG <= 0 || p.NC <= 0 || p.NB <= 0) return 0;
  if (sem == kSemBgw)
    return p.NG * ((p.NB + kAchBands - 1) / kAchBands);
  return (p.NG + kAchThreads - 1) / kAchThreads;
}

hipError_t launch_achstemp(const DevProblem &p, int sem, double *partials,
                           hipStream_t s) {
  const int n = achstemp_records(p, sem);
  if (n == 0) return hipSuccess;
  if (sem == kSemBgw)
    achstemp_bgw_kernel<<<n, kAchThreads, 0, s>>>(p, partials, 0);
  else
    achstemp_gppker_kernel<<<n, kAchThreads, 0, s>>>(p, partials, 0);
  return hipGetLastError();
}

}  // namespace gpp
