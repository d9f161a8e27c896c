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
//   * the three frequencies of two G-vectors share one reciprocal: the
//     product of the six |wdiff|^2 by v_rcp_f64 plus one Newton step, then
//     multiplies give each 1/|wdiff|^2 (batch inversion).  A zero/denormal/
//     non-finite product falls back to plain divisions so that IEEE
//     inf/NaN behave as in the reference;
//   * lane partials are weighted by 0.5*vcoul[igp], reduced across the wave
//     with cross-lane shuffles, across the block through LDS, and written
//     to one record per block; a second kernel sums the records in a fixed
//     order.  No atomics touch the sums, so results are bit-reproducible.
// The occupied bands (n1 < nvband, gppKer.cpp:122-146 / 16-62) run in a
// separate FP64 kernel with one block per (my_igp, ig chunk).
// When every band shares the frequency grid, the non-occupied sum can also
// be taken band sum first: a complex product over the bands on the FP64
// matrix cores, contracted with delw * I_eps (unocc_gemm_kernel; its
// per-band log by a second product over the G-vectors, n1log_gemm_kernel).
#include "gpp_device_common.h"

namespace gpp {
namespace {

// ---------------------------------------------------------------------------
// Prepass: resolve igp per my_igp and expand/clamp the frequency grid.
// Index rule of gppKer.cpp:287-293: indigp = inv_igp_index[my_igp];
// igp = indinv[indigp], except indigp == NC -> igp = NC-1.  Rows whose igp
// falls outside [0, NC) are skipped (the reference's guard admits igp == NC
// and then reads one element past the row; here that row is skipped).
// Clamp of gppKer.cpp:279: |wx| < to1 -> to1.
__device__ __forceinline__ void prepass_one(int i, const int *__restrict__ inv_igp_index,
                                            const int *__restrict__ indinv,
                                            const double *__restrict__ wx_in,
                                            int wx_per_band, int NB, int NC,
                                            int NG, int *__restrict__ igp_of,
                                            double *__restrict__ wx_out) {
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
    // the occupied solver's first-branch factor cden / |cden|^2 with
    // cden = wxt^2 (gppKer.cpp:38-39), once per (band, frequency)
    const double c = w * w;
    wx_out[NB * 3 + i] = c * recip(c * c);
    // FP32 hi/lo split of the frequency for the mixed-precision kernel,
    // 8 floats (32 bytes) per band
    float *split = (float *)(wx_out + NB * 6) + (i / 3) * 8 + (i % 3) * 2;
    const float hi = (float)w;
    split[0] = hi;
    split[1] = (float)(w - (double)hi);
    if (i % 3 == 2) {
      // slot 6 of the row: nonzero when a frequency of the band is not
      // finite (the mixed kernel then runs the band in FP64); slot 7: 0
      const double w0 = wx_per_band ? wx_in[i - 2] : wx_in[0];
      const double w1 = wx_per_band ? wx_in[i - 1] : wx_in[1];
      const bool fin = isfinite(w0) && isfinite(w1) && isfinite(w);
      split[2] = fin ? 0.0f : 1.0f;
      split[3] = 0.0f;
    }
  }
}

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
  prepass_one(i, inv_igp_index, indinv, wx_in, wx_per_band, NB, NC, NG,
              igp_of, wx_out);
}

// aqsmG[n1][my_igp] = aqsm[n1][igp_of[my_igp]] (0 for skipped rows): the
// per-lane aqsm operand of the non-occupied kernels becomes one coalesced
// row per band instead of a gather with stride nodes_per_group.
__global__ void gather_aqsm_kernel(const double2 *__restrict__ aqsm,
                                   const int *__restrict__ igp_of, int NB,
                                   int NC, int NG,
                                   double2 *__restrict__ out) {
  const size_t n = (size_t)NB * NG;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    const int n1 = (int)(i / NG);
    const int g = (int)(i % NG);
    const int igp = igp_of[g];
    out[i] = igp >= 0 ? aqsm[(size_t)n1 * NC + igp] : make_double2(0.0, 0.0);
  }
}

// out[n1][ig] = (float2)in[n1][ig] with rows padded to padded_nc(NC) by
// zeros.
__device__ __forceinline__ void to_fp32_rows(const double2 *__restrict__ in,
                                             float2 *__restrict__ out, int NB,
                                             int NC, size_t i0, size_t stride) {
  const int ncp = padded_nc(NC);
  const size_t n = (size_t)NB * ncp;
  for (size_t i = i0; i < n; i += stride) {
    const int ig = (int)(i % ncp);
    const size_t n1 = i / ncp;
    float2 o = make_float2(0.0f, 0.0f);
    if (ig < NC) {
      const double2 v = in[n1 * NC + ig];
      o = make_float2((float)v.x, (float)v.y);
    }
    out[i] = o;
  }
}

__global__ void to_fp32_kernel(const double2 *__restrict__ in,
                               float2 *__restrict__ out, int NB, int NC) {
  to_fp32_rows(in, out, NB, NC, (size_t)blockIdx.x * blockDim.x + threadIdx.x,
               (size_t)gridDim.x * blockDim.x);
}

// The same three steps for every problem of a batch in one launch each
// (problem = blockIdx.y), so a batch of many small problems is not a chain
// of thousands of tiny launches.
__global__ void prepass_batch_kernel(const DevProblem *__restrict__ probs,
                                     const PrepItem *__restrict__ preps,
                                     unsigned long long *__restrict__ counter) {
  const DevProblem &p = probs[blockIdx.y];
  const PrepItem &q = preps[blockIdx.y];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (blockIdx.y == 0 && i == 0) *counter = 0ull;
  prepass_one(i, q.inv, q.indinv, q.wx_in, q.wx_per_band, p.NB, p.NC, p.NG,
              (int *)p.igp_of, (double *)p.wx);
}

__global__ void gather_aqsm_batch_kernel(const DevProblem *__restrict__ probs) {
  const DevProblem &p = probs[blockIdx.y];
  const size_t n = (size_t)p.NB * p.NG;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  double2 *out = (double2 *)p.aqsmG;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += stride) {
    const int n1 = (int)(i / p.NG);
    const int g = (int)(i % p.NG);
    const int igp = p.igp_of[g];
    out[i] = igp >= 0 ? p.aqsm[(size_t)n1 * p.NC + igp] : make_double2(0.0, 0.0);
  }
}

__global__ void to_fp32_batch_kernel(const DevProblem *__restrict__ probs) {
  const DevProblem &p = probs[blockIdx.y];
  if (p.NB <= p.NV || !p.aqsn32) return;
  to_fp32_rows(p.aqsn, (float2 *)p.aqsn32, p.NB, p.NC,
               (size_t)blockIdx.x * blockDim.x + threadIdx.x,
               (size_t)gridDim.x * blockDim.x);
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
// n1log (optional): [block][n1 - n1_begin] complex, the block's share of
// acht_n1_loc (gppKer.cpp:357).
template <int K, bool MASK, bool N1LOG, typename Prob>
__device__ __forceinline__ void unocc_fp64_block(const Prob &p,
                                                 int n1_begin, int n1_end,
                                                 int b, double *__restrict__ rec,
                                                 double *__restrict__ n1log) {
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

  double wr[K], wi[K], wi2[K], Ar[K], Ai[K];
  // the mask keeps a term iff |wdiff|^2 > max(limittwo, |wtilde|^2 / limitone)
  double mthr[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int ig = ig0 + k;
    if (valid && ig < NC) {
      const size_t off = (size_t)my_igp * NC + ig;
      const double2 w = ld(p.wtilde + off);
      const double2 e = ld(p.ieps + off);
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
    mthr[k] = fmax(kLimitTwo, (wr[k] * wr[k] + wi2[k]) / kLimitOne);
  }

  // sch_array[iw] = 0.5 * sum; achtemp[iw] += sch_array[iw] * vcoul[igp]
  const double wgt = valid ? 0.5 * p.vcoul[igp] : 0.0;
  double acc_r[3] = {0.0, 0.0, 0.0}, acc_i[3] = {0.0, 0.0, 0.0};
  // aqsm gathered per band into [n1][my_igp] (coalesced across lanes).
  const auto aqsm_col = p.aqsmG + min(my_igp, NG - 1);
  double2 am_next = (n1_begin < n1_end)
                        ? ld(aqsm_col + (size_t)n1_begin * NG)
                        : make_double2(0.0, 0.0);

  // For K <= 8 the next band's aqsn run and frequencies (wave-uniform, in
  // SGPRs) are loaded a whole band ahead, so no scalar-load latency sits
  // between the bands; for larger K (or the per-band log's extra live
  // values) the registers do not hold two runs.
  constexpr bool kAhead = K <= 8 && !N1LOG;
  // per-band log: four bands per wave-level sum (16 KB of LDS per block)
  __shared__ double log_slot[N1LOG ? 2 : 1][kBandLogChunk][4][2];
  __shared__ double2 log_stage[N1LOG ? 4 : 1][N1LOG ? 4 : 1][kWave];
  const BandLog<N1LOG ? 4 : 1> blog{log_slot, log_stage[N1LOG ? wave : 0],
                                    n1log, (size_t)b * (n1_end - n1_begin),
                                    n1_end - n1_begin};
  double2 an_cur[K];
  double x_cur[3] = {0.0, 0.0, 0.0};
  if (kAhead && n1_begin < n1_end) {
#pragma unroll
    for (int k = 0; k < K; ++k)
      an_cur[k] =
          ld(p.aqsn + (size_t)n1_begin * NC + min(ig0 + k, NC - 1));
#pragma unroll
    for (int iw = 0; iw < 3; ++iw) x_cur[iw] = p.wx[n1_begin * 3 + iw];
  }

  for (int n1 = n1_begin; n1 < n1_end; ++n1) {
    const double2 am = am_next;
    const int n1n = min(n1 + 1, n1_end - 1);
    // Prefetch next band's aqsm (index clamped: no branch in the loop).
    am_next = ld(aqsm_col + (size_t)n1n * NG);
    double2 an_nxt[K];
    double x_nxt[3];
    if (kAhead) {
#pragma unroll
      for (int k = 0; k < K; ++k)
        an_nxt[k] = ld(p.aqsn + (size_t)n1n * NC + min(ig0 + k, NC - 1));
#pragma unroll
      for (int iw = 0; iw < 3; ++iw) x_nxt[iw] = p.wx[n1n * 3 + iw];
    } else {
#pragma unroll
      for (int iw = 0; iw < 3; ++iw) x_cur[iw] = p.wx[n1 * 3 + iw];
    }
    const double x0 = x_cur[0];
    const double x1 = x_cur[1];
    const double x2 = x_cur[2];
    const auto an_row = p.aqsn + (size_t)n1 * NC;

    double s_r0 = 0.0, s_i0 = 0.0, s_r1 = 0.0, s_i1 = 0.0, s_r2 = 0.0,
           s_i2 = 0.0;
    bool bad = false;
    // G-vectors in pairs: the six |wdiff|^2 of a pair share one reciprocal
    // (batch inversion: a product tree, one v_rcp_f64 + Newton step, then
    // back-substitution).  That takes 15 multiplies for six reciprocals
    // instead of 12 for two groups of three, but one v_rcp_f64 (a
    // quarter-rate instruction) instead of two.
    constexpr int G = 2;
    static_assert(K % G == 0, "K must be even");
#pragma unroll
    for (int k = 0; k < K; k += G) {
      double Br[G], Bi[G], nBiwi[G], Brwi[G], dr[G][3], d[G][3], r[G][3];
#pragma unroll
      for (int j = 0; j < G; ++j) {
        const double2 an =
            kAhead ? an_cur[k + j] : ld(an_row + min(ig0 + k + j, NC - 1));
        Br[j] = an.x * Ar[k + j] - an.y * Ai[k + j];
        Bi[j] = an.x * Ai[k + j] + an.y * Ar[k + j];
        nBiwi[j] = -Bi[j] * wi[k + j];
        Brwi[j] = Br[j] * wi[k + j];
        dr[j][0] = x0 - wr[k + j];
        dr[j][1] = x1 - wr[k + j];
        dr[j][2] = x2 - wr[k + j];
#pragma unroll
        for (int iw = 0; iw < 3; ++iw)
          d[j][iw] = fma(dr[j][iw], dr[j][iw], wi2[k + j]);
      }
      // Tree (depth 3 each way) rather than a running product, so the
      // dependent chain around the reciprocal stays short.
      const double a = d[0][0] * d[0][1];
      const double b = d[0][2] * d[1][0];
      const double c = d[1][1] * d[1][2];
      const double ab = a * b;
      const double pp = ab * c;
      bad |= !is_pos_normal(pp);
      double q = rcp_f64(pp);
      const double e = fma(-pp, q, 1.0);
      q = fma(q, e, q);           // 1 / (a b c)
      const double qab = q * c;   // 1 / (a b)
      const double qc = q * ab;   // 1 / c
      const double qa = qab * b;  // 1 / a
      const double qb = qab * a;  // 1 / b
      r[0][0] = qa * d[0][1];
      r[0][1] = qa * d[0][0];
      r[0][2] = qb * d[1][0];
      r[1][0] = qb * d[0][2];
      r[1][1] = qc * d[1][2];
      r[1][2] = qc * d[1][1];
#pragma unroll
      for (int j = 0; j < G; ++j) {
        if (MASK) {
          // gppkernel.deslippe.09.30.2014.f90:364-366 /
          // BGW-Kernels/gppKer.cpp:406: keep the term only if |wdiff|^2 >
          // limittwo and |delw|^2 = |wtilde|^2/|wdiff|^2 < limitone, i.e.
          // (|wdiff|^2 > 0) |wdiff|^2 > max(limittwo, |wtilde|^2/limitone):
          // one compare per term against the G-vector's threshold.
#pragma unroll
          for (int iw = 0; iw < 3; ++iw)
            r[j][iw] = d[j][iw] > mthr[k + j] ? r[j][iw] : 0.0;
        }
        s_r0 = fma(fma(Br[j], dr[j][0], nBiwi[j]), r[j][0], s_r0);
        s_i0 = fma(fma(Bi[j], dr[j][0], Brwi[j]), r[j][0], s_i0);
        s_r1 = fma(fma(Br[j], dr[j][1], nBiwi[j]), r[j][1], s_r1);
        s_i1 = fma(fma(Bi[j], dr[j][1], Brwi[j]), r[j][1], s_i1);
        s_r2 = fma(fma(Br[j], dr[j][2], nBiwi[j]), r[j][2], s_r2);
        s_i2 = fma(fma(Bi[j], dr[j][2], Brwi[j]), r[j][2], s_i2);
      }
    }
    if (__builtin_expect(bad, 0)) {
      // Some pair's product of |wdiff|^2 was zero, denormal, infinite or
      // NaN: redo this lane's run with one division per term so that exact
      // resonances give the reference's inf/NaN in their own frequency only.
      // The operands are reloaded (this path is rare) and the loop is kept
      // rolled, so that it adds no register pressure to the hot loop.
      s_r0 = s_i0 = s_r1 = s_i1 = s_r2 = s_i2 = 0.0;
#pragma unroll 1
      for (int k = 0; k < K; ++k) {
        const int ig = ig0 + k;
        if (!valid || ig >= NC) continue;  // inert slot: contributes 0
        const size_t off = (size_t)my_igp * NC + ig;
        const double2 w = ld(p.wtilde + off);
        const double2 e = ld(p.ieps + off);
        const double A_r = w.x * e.x - w.y * e.y;
        const double A_i = w.x * e.y + w.y * e.x;
        const double2 an = ld(an_row + ig);
        const double Br = an.x * A_r - an.y * A_i;
        const double Bi = an.x * A_i + an.y * A_r;
        const double wy2 = w.y * w.y;
        const double xs[3] = {x0, x1, x2};
        double sr[3], si[3];
#pragma unroll
        for (int iw = 0; iw < 3; ++iw) {
          const double dr = xs[iw] - w.x;
          const double d = fma(dr, dr, wy2);
          const double r = 1.0 / d;
          // the mask skips the term outright (a NaN term it drops adds
          // nothing, as in the f90 loop)
          const bool keep =
              !MASK || (d > kLimitTwo && (w.x * w.x + wy2) * r < kLimitOne);
          sr[iw] = keep ? (Br * dr - Bi * w.y) * r : 0.0;
          si[iw] = keep ? (Br * w.y + Bi * dr) * r : 0.0;
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
    if (N1LOG)  // acht_n1_loc[n1] += sch_array[2] * vcoul[igp]
      blog.put(n1 - n1_begin, wgt * fma(am.x, s_r2, am.y * s_i2),
               wgt * fma(am.x, s_i2, -am.y * s_r2));
    if (kAhead) {
#pragma unroll
      for (int k = 0; k < K; ++k) an_cur[k] = an_nxt[k];
#pragma unroll
      for (int iw = 0; iw < 3; ++iw) x_cur[iw] = x_nxt[iw];
    }
  }

  double v[6];
#pragma unroll
  for (int iw = 0; iw < 3; ++iw) {
    v[2 * iw] = wgt * acc_r[iw];
    v[2 * iw + 1] = wgt * acc_i[iw];
  }
  block_store<6>(v, rec);
  if (threadIdx.x >= 6 && threadIdx.x < kSlots) rec[threadIdx.x] = 0.0;
}

// For K <= 8 the register budget is capped at 3 waves per SIMD (168 VGPRs):
// left alone the compiler takes a few more for scheduling and falls to 2
// waves; the cap costs only spills outside the band loop (or one reload per
// band in the masked variant).
template <int K, bool MASK, bool N1LOG>
__global__ __launch_bounds__(256) __attribute__((
    amdgpu_waves_per_eu(K <= 8 ? 3 : 1))) void unocc_fp64_kernel(
    DevProblem p, int n1_begin, int n1_end, int grp_lo, int grp_cnt,
    double *__restrict__ partials, double *__restrict__ n1log) {
  const int b = full_grid_block(p.NG, grp_lo, grp_cnt);
  unocc_fp64_block<K, MASK, N1LOG>(p, n1_begin, n1_end, b,
                                   partials + (size_t)b * kSlots, n1log);
}

// ---------------------------------------------------------------------------
// Occupied bands (n1 < nvband): ssxt_scht_solver / flagOCC_solver of
// gppKer.cpp:16-62,122-146 (SEM gppker, gppker-strict) or the occupied
// branch of gppkernel.deslippe.09.30.2014.f90:271-334 (SEM bgw).
// Block b: my_igp = b / nch, G-vector chunk b % nch; threads stride over the
// chunk.  Each thread reads wtilde and I_eps of its G-vector once and runs
// the occupied bands inside, so those rows cross HBM once, not once per band
// (the per-band log variant keeps bands outer: it needs one block sum per
// band).
constexpr int kOccThreads = 256;
constexpr int kOccChunk = 2048;  // G-vectors per block

// The band-independent part of one (my_igp, ig) of the occupied solver.
struct OccG {
  double wr, wi, wy2;  // wtilde, wtilde.im^2
  double wiwr;         // wtilde.im * wtilde.re
  double w2r, w2i;     // wtilde2 = wtilde^2 (gppKer.cpp:132)
  double omr, omi;     // Omega2 = wtilde2 * I_eps (gppKer.cpp:133)
  double er, ei;       // I_eps[my_igp][ig], Omega2's entry
  double exr, exi;     // the I_eps entry used for sch and the cutoff
  double cut2;         // (sexcut * |I_eps|)^2
};

// From wtilde, I_eps[my_igp][ig] and ex, the I_eps entry the solver reads.
__device__ __forceinline__ OccG occ_g_of(double2 w, double2 e, double2 ex) {
  OccG g;
  g.wr = w.x;
  g.wi = w.y;
  g.wy2 = w.y * w.y;
  g.wiwr = w.y * w.x;
  g.w2r = w.x * w.x - g.wy2;
  g.w2i = 2.0 * w.x * w.y;
  g.omr = g.w2r * e.x - g.w2i * e.y;
  g.omi = g.w2r * e.y + g.w2i * e.x;
  g.er = e.x;
  g.ei = e.y;
  g.exr = ex.x;
  g.exi = ex.y;
  g.cut2 = kSexcut * kSexcut * (ex.x * ex.x + ex.y * ex.y);
  return g;
}

template <int SEM, typename Prob>
__device__ __forceinline__ OccG occ_g(const Prob &p, int my_igp, int ig) {
  const int NC = p.NC, NG = p.NG;
  const size_t off = (size_t)my_igp * NC + ig;
  const double2 w = ld(p.wtilde + off);
  const double2 e = ld(p.ieps + off);
  // gppKer.cpp:139 passes I_eps_array[my_igp*ngpown+ig] to the solver; the
  // strict mode keeps that index, the others read [my_igp][ig].
  const double2 ex =
      (SEM == kSemGppkerStrict) ? ld(p.ieps + (size_t)my_igp * NG + ig) : e;
  return occ_g_of(w, e, ex);
}

// The band's frequencies and, for the gppker first branch, the factor
// cden / |cden|^2 with cden = wxt^2 (real, the same for every G-vector).
struct OccBand {
  double wxt[3], fssx[3];
};

template <typename Prob>
__device__ __forceinline__ OccBand occ_band(const Prob &p, int n1) {
  OccBand B;
#pragma unroll
  for (int iw = 0; iw < 3; ++iw) {
    B.wxt[iw] = p.wx[n1 * 3 + iw];
    B.fssx[iw] = p.wx[p.NB * 3 + n1 * 3 + iw];  // formed by prepass_kernel
  }
  return B;
}

// One (n1, my_igp, ig), three frequencies: adds matngmatmgp * sch to ts and
// matngmatmgp * ssx to tx (gppKer.cpp:27-61, 141-142; f90:290-329).
//   first branch  (|wdiff|^2 > limittwo, |delw|^2 < limitone):
//     sch = delw * I_eps_x, ssx = Omega2 * conj(cden) / |cden|^2 with
//     cden = wxt^2 (gppker) or wxt^2 - wtilde2 (bgw);
//   second branch (|delw|^2 > to1): sch = 0 and, with cden =
//     4 wtilde2 (delw + 1/2), ssx = -Omega2 conj(cden) delw / |cden|^2 =
//     -Omega2 delw / cden = -I_eps delw / (4 delw + 2): wtilde2 cancels
//     (it is nonzero there: delw = 0 when wtilde = 0), which spares the
//     product with wtilde2 and Omega2;
//   else 0.
// The products with matngmatmgp m are formed once per (n1, ig) for the three
// frequencies (m * I_eps_x for sch, m * Omega2 for the gppker ssx).
template <int SEM>
__device__ __forceinline__ void occ_term(const OccG &g, const OccBand &B,
                                         double mr, double mi,
                                         double (&tsr)[3], double (&tsi)[3],
                                         double (&txr)[3], double (&txi)[3]) {
  constexpr bool BGW = SEM == kSemBgw;
  // wdiff = wxt - wtilde = (dr, -wi); conj(wdiff) = (dr, wi)
  double dr[3], wdiffr[3], rden[3];
#pragma unroll
  for (int iw = 0; iw < 3; ++iw) {
    dr[iw] = B.wxt[iw] - g.wr;
    wdiffr[iw] = fma(dr[iw], dr[iw], g.wy2);
  }
  recip3(wdiffr[0], wdiffr[1], wdiffr[2], rden[0], rden[1], rden[2]);
  // bgw first branch: 1 / |wxt^2 - wtilde2|^2 for the three frequencies
  double cdr[3], rcd[3];
  if (BGW) {
    double cmag[3];
#pragma unroll
    for (int iw = 0; iw < 3; ++iw) {
      cdr[iw] = fma(B.wxt[iw], B.wxt[iw], -g.w2r);
      cmag[iw] = fma(cdr[iw], cdr[iw], g.w2i * g.w2i);
    }
    recip3(cmag[0], cmag[1], cmag[2], rcd[0], rcd[1], rcd[2]);
  }
  const double mxr = mr * g.exr - mi * g.exi;  // m * I_eps_x
  const double mxi = mr * g.exi + mi * g.exr;
  const double mor = mr * g.omr - mi * g.omi;  // m * Omega2
  const double moi = mr * g.omi + mi * g.omr;
#pragma unroll
  for (int iw = 0; iw < 3; ++iw) {
    // delw = wtilde * conj(wdiff) * rden
    const double dlr = fma(g.wr, dr[iw], -g.wy2) * rden[iw];
    const double dli = fma(g.wi, dr[iw], g.wiwr) * rden[iw];
    const double delwr = fma(dlr, dlr, dli * dli);
    double ssr = 0.0, ssi = 0.0;  // ssx (bgw, and the second branch)
    bool have_ssx = false;
    if (wdiffr[iw] > kLimitTwo && delwr < kLimitOne) {
      tsr[iw] = fma(dlr, mxr, fma(-dli, mxi, tsr[iw]));
      tsi[iw] = fma(dlr, mxi, fma(dli, mxr, tsi[iw]));
      if (BGW) {
        // cden = wxt^2 - wtilde2 (f90:303): Omega2 * conj(cden) / |cden|^2
        const double ci = -g.w2i;
        ssr = fma(g.omr, cdr[iw], g.omi * ci) * rcd[iw];
        ssi = fma(g.omi, cdr[iw], -g.omr * ci) * rcd[iw];
        have_ssx = true;
      } else {
        txr[iw] = fma(mor, B.fssx[iw], txr[iw]);
        txi[iw] = fma(moi, B.fssx[iw], txi[iw]);
      }
    } else if (delwr > kTo1) {
      // q = 4 delw + 2; ssx = -I_eps * delw * conj(q) / |q|^2
      const double qr = fma(4.0, dlr, 2.0), qi = 4.0 * dli;
      const double rq = recip(fma(qr, qr, qi * qi));
      const double nr = g.er * dlr - g.ei * dli;  // I_eps * delw
      const double ni = g.er * dli + g.ei * dlr;
      ssr = -fma(nr, qr, ni * qi) * rq;
      ssi = -fma(ni, qr, -nr * qi) * rq;
      have_ssx = true;
    }
    if (have_ssx) {
      if (BGW) {
        // f90:321: |ssx| > sexcut*|I_eps| and wxt < 0 -> ssx = 0
        if (B.wxt[iw] < 0.0 && fma(ssr, ssr, ssi * ssi) > g.cut2) {
          ssr = 0.0;
          ssi = 0.0;
        }
      }
      // (gppKer.cpp:58 tests abs(wxt) < 0, which never holds.)
      txr[iw] = fma(mr, ssr, fma(-mi, ssi, txr[iw]));
      txi[iw] = fma(mr, ssi, fma(mi, ssr, txi[iw]));
    }
  }
}

// FACT: every band has the same frequencies (wx shared), so the solver's
// sch and ssx do not depend on n1; the bands' matngmatmgp are summed first
// and the solver runs once per G-vector instead of once per band.
template <int SEM, bool N1LOG, typename Prob, bool FACT = false>
__device__ __forceinline__ void occ_block(const Prob &p, int n1_end,
                                          int b, double *__restrict__ rec,
                                          double *__restrict__ n1log) {
  const int NC = p.NC;
  const int nch = (NC + kOccChunk - 1) / kOccChunk;
  const int my_igp = b / nch;
  const int ch = b % nch;
  const int igp = p.igp_of[my_igp];
  if (igp < 0) {  // skipped row (block-uniform)
    if (threadIdx.x < kSlots) rec[threadIdx.x] = 0.0;
    if (N1LOG)
      for (int n1 = 0; n1 < n1_end; ++n1)
        n1log_store(0.0, 0.0, n1log, (size_t)b * n1_end + n1);
    return;
  }
  const double vc = p.vcoul[igp];
  const double wsch = 0.5 * vc;                              // :310,355
  const double wssx = (SEM == kSemBgw ? -kOcc : kOcc) * vc;  // :352 / f90:393
  double tsr[3] = {0, 0, 0}, tsi[3] = {0, 0, 0};
  double txr[3] = {0, 0, 0}, txi[3] = {0, 0, 0};
  const int ig_begin = ch * kOccChunk + threadIdx.x;
  const int ig_end = min(NC, (ch + 1) * kOccChunk);
  if (N1LOG) {
    for (int n1 = 0; n1 < n1_end; ++n1) {
      const double2 am = ld(p.aqsm + (size_t)n1 * NC + igp);
      const OccBand B = occ_band(p, n1);
      double bsr[3] = {0, 0, 0}, bsi[3] = {0, 0, 0};
      for (int ig = ig_begin; ig < ig_end; ig += blockDim.x) {
        const OccG g = occ_g<SEM>(p, my_igp, ig);
        const double2 an = ld(p.aqsn + (size_t)n1 * NC + ig);
        // matngmatmgp = aqsn[n1][ig] * conj(aqsm[n1][igp])
        const double mr = an.x * am.x + an.y * am.y;
        const double mi = an.y * am.x - an.x * am.y;
        occ_term<SEM>(g, B, mr, mi, bsr, bsi, txr, txi);
      }
#pragma unroll
      for (int iw = 0; iw < 3; ++iw) {
        tsr[iw] += bsr[iw];
        tsi[iw] += bsi[iw];
      }
      // acht_n1_loc[n1] += sch_array[2] * vcoul[igp]
      n1log_store(wsch * bsr[2], wsch * bsi[2], n1log,
                  (size_t)b * n1_end + n1);
    }
  } else if (FACT) {
    const OccBand B = occ_band(p, 0);
    for (int ig = ig_begin; ig < ig_end; ig += blockDim.x) {
      const OccG g = occ_g<SEM>(p, my_igp, ig);
      double mr = 0.0, mi = 0.0;
      for (int n1 = 0; n1 < n1_end; ++n1) {
        const double2 am = ld(p.aqsm + (size_t)n1 * NC + igp);
        const double2 an = ld(p.aqsn + (size_t)n1 * NC + ig);
        mr += an.x * am.x + an.y * am.y;
        mi += an.y * am.x - an.x * am.y;
      }
      occ_term<SEM>(g, B, mr, mi, tsr, tsi, txr, txi);
    }
  } else {
    // wtilde and I_eps of the thread's next G-vector are loaded before the
    // current one's bands are evaluated, so that with few occupied bands
    // (2 at the canonical shape) the loads overlap the solver.
    const size_t row = (size_t)my_igp * NC;
    const size_t rowx = (SEM == kSemGppkerStrict) ? (size_t)my_igp * p.NG : row;
    double2 w_nxt = make_double2(0.0, 0.0), e_nxt = w_nxt, x_nxt = w_nxt;
    if (ig_begin < ig_end) {
      w_nxt = ld(p.wtilde + row + ig_begin);
      e_nxt = ld(p.ieps + row + ig_begin);
      if (SEM == kSemGppkerStrict) x_nxt = ld(p.ieps + rowx + ig_begin);
    }
    for (int ig = ig_begin; ig < ig_end; ig += blockDim.x) {
      const OccG g = occ_g_of(w_nxt, e_nxt,
                              SEM == kSemGppkerStrict ? x_nxt : e_nxt);
      const int ign = min(ig + (int)blockDim.x, ig_end - 1);
      w_nxt = ld(p.wtilde + row + ign);
      e_nxt = ld(p.ieps + row + ign);
      if (SEM == kSemGppkerStrict) x_nxt = ld(p.ieps + rowx + ign);
      for (int n1 = 0; n1 < n1_end; ++n1) {
        const double2 am = ld(p.aqsm + (size_t)n1 * NC + igp);
        const OccBand B = occ_band(p, n1);
        const double2 an = ld(p.aqsn + (size_t)n1 * NC + ig);
        const double mr = an.x * am.x + an.y * am.y;
        const double mi = an.y * am.x - an.x * am.y;
        occ_term<SEM>(g, B, mr, mi, tsr, tsi, txr, txi);
      }
    }
  }
  double v[12];
#pragma unroll
  for (int iw = 0; iw < 3; ++iw) {
    v[2 * iw] = wsch * tsr[iw];
    v[2 * iw + 1] = wsch * tsi[iw];
    v[6 + 2 * iw] = wssx * txr[iw];
    v[6 + 2 * iw + 1] = wssx * txi[iw];
  }
  block_store<12>(v, rec);
}

template <int SEM, bool N1LOG, bool FACT = false>
__global__ __launch_bounds__(kOccThreads) void occ_kernel(
    DevProblem p, int n1_end, double *__restrict__ partials,
    double *__restrict__ n1log) {
  occ_block<SEM, N1LOG, DevProblem, FACT>(
      p, n1_end, blockIdx.x, partials + (size_t)blockIdx.x * kSlots, n1log);
}

// ---------------------------------------------------------------------------
// Non-occupied bands, band sum first (GPP_PREC_FP64_FACTORED).
// When wx does not depend on the band, neither does delw, and the sum of
// gppKer.cpp:316-345 factors:
//   achtemp[iw] = sum_{my_igp} 0.5 vcoul[igp] sum_ig P[my_igp][ig] *
//                 wtilde*I_eps * conj(wdiff_iw) / |wdiff_iw|^2
//   P[my_igp][ig] = sum_{n1} conj(aqsm[n1][igp]) * aqsn[n1][ig]
// P is a complex matrix product over the bands (K = n1_end - n1_begin),
// formed on the FP64 matrix cores as three real products (the 3M method):
//   T1 = Ar^T Br,  T2 = Ai^T Bi,  T3 = (Ar - Ai)^T (Br + Bi)
//   Re P = T1 + T2,  Im P = T3 - T1 + T2
// (A = aqsmG [n1][my_igp], B = aqsn [n1][ig]): a quarter fewer MFMAs than
// the four products Ar^T Br + Ai^T Bi, Ar^T Bi - Ai^T Br (2.68 against
// 3.42 ms at 512/2/32768/20), the two operand sums formed in registers.  A block owns 64 G' columns
// (one group, as the other kernels) by 64 G-vectors; each of its four waves
// a 32 x 32 quarter: 2 x 2 tiles of v_mfma_f64_16x16x4 (16 x 16 outputs, 4
// bands per step).  That leaves room for 3 waves per SIMD (2 waves of
// 32 x 64 quarters were 15% slower).  The band rows of A and B go through
// LDS 8 bands at a time, double-buffered, the next stage's global loads in
// flight during this stage's MFMAs (reading the operands straight from
// global memory instead was 10% slower at 512/2/32768/20 on MI355X).  P
// stays in registers and the epilogue contracts it with delw * I_eps of
// the same (my_igp, ig) read once from HBM; P is never stored.
typedef double dbl4 __attribute__((ext_vector_type(4)));
constexpr int kGemmM = 64;   // G' columns per block
constexpr int kGemmNB = 2;   // 16-wide column tiles per wave
constexpr int kGemmN = 2 * 16 * kGemmNB;  // G-vectors per block
constexpr int kGemmThreads = 256;
constexpr int kGemmWaves = 3;  // waves per SIMD the registers are sized for
constexpr int kGemmStage = 8;  // bands per LDS stage (a multiple of 4)
static_assert(kGemmN == kWave && kGemmM == kWave,
              "one lane per column of the staged A and B rows");

__device__ __forceinline__ dbl4 mfma4(double a, double b, dbl4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// One step of 4 bands: a[mb] = A[n1][row tile mb], b[nb] = B[n1][col tile nb]
// for this lane's (band, row/column) slot.  Each of the three products runs
// over all tiles before the next, so no MFMA waits on the previous one's
// result.  The third product takes (Ar - Ai)(Br + Bi) for conj(A) B (CONJ)
// and (Ar + Ai)(Br + Bi) for A B.
template <int NB, bool CONJ>
__device__ __forceinline__ void mfma3_step(const double2 (&a)[2],
                                           const double2 (&b)[NB],
                                           dbl4 (&t1)[2][NB], dbl4 (&t2)[2][NB],
                                           dbl4 (&t3)[2][NB]) {
  double as[2], bs[NB];
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
    as[mb] = CONJ ? a[mb].x - a[mb].y : a[mb].x + a[mb].y;
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) bs[nb] = b[nb].x + b[nb].y;
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
      t1[mb][nb] = mfma4(a[mb].x, b[nb].x, t1[mb][nb]);
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
      t2[mb][nb] = mfma4(a[mb].y, b[nb].y, t2[mb][nb]);
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
      t3[mb][nb] = mfma4(as[mb], bs[nb], t3[mb][nb]);
}

template <bool MASK>
__global__ __launch_bounds__(kGemmThreads) __attribute__((
    amdgpu_waves_per_eu(kGemmWaves))) void unocc_gemm_kernel(
    DevProblem p, int n1_begin, int n1_end, int grp_lo, int grp_cnt,
    double *__restrict__ partials) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = uniform(threadIdx.x / kWave);
  const int NC = p.NC, NG = p.NG;
  const int ngrp = (NG + kGemmM - 1) / kGemmM;
  // Tile of this block, numbered as in the launch over all groups (so a
  // pipelined launch over a range of groups writes the same records).
  const TileIdx t = tile_order(grp_cnt);
  const int b = t.tile * ngrp + grp_lo + t.grp;
  const int igp_grp = b % ngrp, ig_tile = b / ngrp;
  const int m0 = igp_grp * kGemmM + (wave & 1) * 32;
  const int c0 = ig_tile * kGemmN + (wave >> 1) * 16 * kGemmNB;
  const int li = lane & 15, lk = lane >> 4;

  dbl4 t1[2][kGemmNB], t2[2][kGemmNB], t3[2][kGemmNB];
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int nb = 0; nb < kGemmNB; ++nb) {
      t1[mb][nb] = dbl4{0.0, 0.0, 0.0, 0.0};
      t2[mb][nb] = dbl4{0.0, 0.0, 0.0, 0.0};
      t3[mb][nb] = dbl4{0.0, 0.0, 0.0, 0.0};
    }

  // Bands go through LDS kGemmStage at a time, double-buffered.  Thread
  // (wave, lane) copies rows wave*kRows .. +kRows of the stage: one complex
  // of aqsmG (G' column lane of the block) and one of aqsn (G-vector lane),
  // so each row is 1 KB read by one wave instruction.  The next stage's
  // loads are issued into registers before this stage's MFMAs and written
  // to the other buffer after them.  Rows past n1_end are stored as zeros;
  // columns past NG / NC read clamped (finite) data and get zero weight in
  // the epilogue.
  constexpr int kRows = kGemmStage / 4;
  __shared__ double2 As[2][kGemmStage][kGemmM];
  __shared__ double2 Bs[2][kGemmStage][kGemmN];
  const int la = min(igp_grp * kGemmM + lane, NG - 1);
  const int lb = min(ig_tile * kGemmN + lane, NC - 1);
  const int nst = (n1_end - n1_begin + kGemmStage - 1) / kGemmStage;
  double2 ga[kRows], gb[kRows];
  auto gload = [&](int st) {
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const int n1 = n1_begin + st * kGemmStage + wave * kRows + r;
      const int n1c = min(n1, n1_end - 1);
      const double2 va = p.aqsmG[(size_t)n1c * NG + la];
      const double2 vb = p.aqsn[(size_t)n1c * NC + lb];
      const bool in = n1 < n1_end;
      ga[r] = in ? va : make_double2(0.0, 0.0);
      gb[r] = in ? vb : make_double2(0.0, 0.0);
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      As[buf][wave * kRows + r][lane] = ga[r];
      Bs[buf][wave * kRows + r][lane] = gb[r];
    }
  };
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 16 * kGemmNB;
  if (nst > 0) {
    gload(0);
    lstore(0);
  }
  __syncthreads();
  for (int st = 0; st < nst; ++st) {
    const int buf = st & 1;
    const bool more = st + 1 < nst;
    if (more) gload(st + 1);
    // Keeps the next stage's loads issued ahead of this stage's MFMAs
    // (otherwise the compiler moves them down to their LDS stores).
    asm volatile("" ::: "memory");
#pragma unroll
    for (int kk = 0; kk < kGemmStage / 4; ++kk) {
      const int kr = 4 * kk + lk;
      double2 a[2], bv[kGemmNB];
#pragma unroll
      for (int mb = 0; mb < 2; ++mb) a[mb] = As[buf][kr][wm + 16 * mb + li];
#pragma unroll
      for (int nb = 0; nb < kGemmNB; ++nb)
        bv[nb] = Bs[buf][kr][wn + 16 * nb + li];
      mfma3_step<kGemmNB, true>(a, bv, t1, t2, t3);
    }
    if (more) lstore(buf ^ 1);
    __syncthreads();
  }
  // P = conj(A)^T B from the three products; the registers of T1..T3 are
  // free again for the epilogue.
  dbl4 pr[2][kGemmNB], pi[2][kGemmNB];
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int nb = 0; nb < kGemmNB; ++nb) {
      pr[mb][nb] = t1[mb][nb] + t2[mb][nb];
      pi[mb][nb] = (t3[mb][nb] - t1[mb][nb]) + t2[mb][nb];
    }

  // Epilogue.  Output (i, j) of a 16 x 16 tile sits in lane 16*(i%4) + j,
  // element i/4 (operands: A[i][k] in lane 16k + i, B[k][j] in lane
  // 16k + j).
  const double x0 = p.wx[n1_begin * 3 + 0];
  const double x1 = p.wx[n1_begin * 3 + 1];
  const double x2 = p.wx[n1_begin * 3 + 2];
  // The element loop stays rolled (its index picks the accumulator element
  // by selects): unrolled, the compiler hoists all 64 complex loads of the
  // epilogue and spills.
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
  for (int e = 0; e < 4; ++e)
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
      const int my_igp = m0 + 16 * mb + lk + 4 * e;
      const int igp = my_igp < NG ? p.igp_of[my_igp] : -1;
      const bool valid = igp >= 0;
      const double wgt = valid ? 0.5 * p.vcoul[igp] : 0.0;
      double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int nb = 0; nb < kGemmNB; ++nb) {
        const int ig = c0 + 16 * nb + li;
        double wr = 0.0, wi = 0.0, Ar = 0.0, Ai = 0.0;
        if (valid && ig < NC) {
          const size_t off = (size_t)my_igp * NC + ig;
          const double2 w = p.wtilde[off];
          const double2 ep = p.ieps[off];
          wr = w.x;
          wi = w.y;
          Ar = w.x * ep.x - w.y * ep.y;
          Ai = w.x * ep.y + w.y * ep.x;
        }
        const double Pr = pr[mb][nb][e], Pi = pi[mb][nb][e];
        const double Br = Pr * Ar - Pi * Ai;
        const double Bi = Pr * Ai + Pi * Ar;
        const double wi2 = wi * wi;
        const double dr0 = x0 - wr, dr1 = x1 - wr, dr2 = x2 - wr;
        const double d0 = fma(dr0, dr0, wi2), d1 = fma(dr1, dr1, wi2),
                     d2 = fma(dr2, dr2, wi2);
        double r0, r1, r2;
        recip3(d0, d1, d2, r0, r1, r2);
        if (MASK) {
          // gppkernel.deslippe.09.30.2014.f90:364-366: keep the term only if
          // |wdiff|^2 > limittwo and |wtilde|^2/|wdiff|^2 < limitone.
          const double wt2 = wr * wr + wi2;
          r0 = (d0 > kLimitTwo && wt2 * r0 < kLimitOne) ? r0 : 0.0;
          r1 = (d1 > kLimitTwo && wt2 * r1 < kLimitOne) ? r1 : 0.0;
          r2 = (d2 > kLimitTwo && wt2 * r2 < kLimitOne) ? r2 : 0.0;
        }
        const double nBiwi = -Bi * wi, Brwi = Br * wi;
        s[0] = fma(fma(Br, dr0, nBiwi), r0, s[0]);
        s[1] = fma(fma(Bi, dr0, Brwi), r0, s[1]);
        s[2] = fma(fma(Br, dr1, nBiwi), r1, s[2]);
        s[3] = fma(fma(Bi, dr1, Brwi), r1, s[3]);
        s[4] = fma(fma(Br, dr2, nBiwi), r2, s[4]);
        s[5] = fma(fma(Bi, dr2, Brwi), r2, s[5]);
      }
#pragma unroll
      for (int i = 0; i < 6; ++i) acc[i] = fma(wgt, s[i], acc[i]);
    }
  double *rec = partials + (size_t)b * kSlots;
  block_store<6>(acc, rec);
  if (threadIdx.x >= 6 && threadIdx.x < kSlots) rec[threadIdx.x] = 0.0;

}

// Per-band log (acht_n1_loc, gppKer.cpp:357) of the band-sum-first mode.
// For a non-occupied band n1 the log is
//   acht_n1_loc[n1] = sum_{my_igp} 0.5 vcoul[igp] conj(aqsm[n1][igp]) Q[n1][my_igp]
//   Q[n1][my_igp]   = sum_ig aqsn[n1][ig] * S2[my_igp][ig]
// with S2 = wtilde*I_eps*conj(wdiff_2)/|wdiff_2|^2, the third frequency's
// factor of the sum (masked as the bgw semantics masks it).  Q is a complex
// matrix product over the G-vectors, formed on the FP64 matrix cores like P
// with three real products (T1 = Ar Br, T2 = Ai Bi, T3 = (Ar + Ai)(Br + Bi);
// Re Q = T1 - T2, Im Q = T3 - T1 - T2).  A block owns 64 bands by
// 64 G' columns (one group) and one chunk of kLogChunk G-vectors; its four
// waves hold 32 x 32 quarters.  Rows of aqsn and of S2 (computed from
// wtilde and I_eps while staging) go through LDS kLogStage G-vectors at a
// time.  The epilogue contracts Q with the weights and conj(aqsm) over the
// block's G' columns and stores one complex per band into the record of
// (group, chunk): [record][n1 - n1_begin], the layout launch_reduce_n1 reads.
// ACH: the same product with I_eps in place of S2 over all bands gives the
// bgw static remainder (F90 :195-229, the full form of reduce_achstemp,
// gppKer.cpp:64-120)
//   achstemp = -sum_{n1, my_igp} 0.5 vcoul[igp] conj(aqsm[n1][igp]) *
//              sum_ig aqsn[n1][ig] * I_eps[my_igp][ig]
// and each block adds its bands (in order) into one kSlots record of
// -(its share), records numbered over the grid, for launch_reduce.
constexpr int kLogM = 64;        // bands per block
constexpr int kLogStage = 8;     // G-vectors per LDS stage
constexpr int kLogChunk = 2048;  // G-vectors per block

template <bool MASK, bool ACH = false>
__global__ __launch_bounds__(kGemmThreads) __attribute__((
    amdgpu_waves_per_eu(kGemmWaves))) void n1log_gemm_kernel(
    DevProblem p, int n1_begin, int n1_end, int grp_lo,
    double *__restrict__ n1log) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = uniform(threadIdx.x / kWave);
  const int NC = p.NC, NG = p.NG;
  const int nbu = n1_end - n1_begin;
  const int nsplit = (NC + kLogChunk - 1) / kLogChunk;
  const int bt = blockIdx.x;                 // band tile
  const int grp = grp_lo + (int)blockIdx.y;  // G' group
  const int ch = blockIdx.z;                 // G-vector chunk
  const int i0 = bt * kLogM, m0 = grp * kGemmM;
  const int k_begin = ch * kLogChunk, k_end = min(NC, k_begin + kLogChunk);
  const int li = lane & 15, lk = lane >> 4;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;

  dbl4 t1[2][2], t2[2][2], t3[2][2];
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int nb = 0; nb < 2; ++nb) {
      t1[mb][nb] = dbl4{0.0, 0.0, 0.0, 0.0};
      t2[mb][nb] = dbl4{0.0, 0.0, 0.0, 0.0};
      t3[mb][nb] = dbl4{0.0, 0.0, 0.0, 0.0};
    }

  // Staging: thread t copies row t/4, G-vectors 2*(t%4) and 2*(t%4)+1 of
  // the stage, of aqsn (bands) and of S2 (G' columns).  Rows are padded by
  // one complex against LDS bank conflicts.
  __shared__ double2 As[2][kLogM][kLogStage + 1];
  __shared__ double2 Bs[2][kGemmM][kLogStage + 1];
  const int sr = threadIdx.x >> 2, sc = (threadIdx.x & 3) * 2;
  const int n1s = n1_begin + i0 + sr;
  const bool a_row = i0 + sr < nbu;
  const double2 *arow = p.aqsn + (size_t)min(n1s, n1_end - 1) * NC;
  const int my_igp_s = m0 + sr;
  const int igp_s = my_igp_s < NG ? p.igp_of[my_igp_s] : -1;
  const double2 *wrow = p.wtilde + (size_t)min(my_igp_s, NG - 1) * NC;
  const double2 *erow = p.ieps + (size_t)min(my_igp_s, NG - 1) * NC;
  const double x2 = p.wx[n1_begin * 3 + 2];
  // The loads of the next stage are issued before this stage's MFMAs; S2 is
  // formed from them after the MFMAs, when they are stored to LDS.
  double2 ga[2], gw[2], ge[2];
  bool gin[2];
  auto gload = [&](int kb) {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int ig = kb + sc + c;
      gin[c] = ig < k_end;
      const int igc = min(ig, NC - 1);
      ga[c] = arow[igc];
      if (!ACH) gw[c] = wrow[igc];
      ge[c] = erow[igc];
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const bool live = gin[c] && igp_s >= 0;
      As[buf][sr][sc + c] =
          (gin[c] && a_row) ? ga[c] : make_double2(0.0, 0.0);
      if (ACH) {  // I_eps itself (0 for inert slots)
        Bs[buf][sr][sc + c] = live ? ge[c] : make_double2(0.0, 0.0);
        continue;
      }
      // S2 = A conj(wdiff_2) r_2, A = wtilde I_eps (0 for inert slots)
      const double2 w = gw[c], e = ge[c];
      const double wr = live ? w.x : 0.0, wi = live ? w.y : 0.0;
      const double Ar = live ? w.x * e.x - w.y * e.y : 0.0;
      const double Ai = live ? w.x * e.y + w.y * e.x : 0.0;
      const double dr = x2 - wr;
      const double wi2 = wi * wi;
      const double d = fma(dr, dr, wi2);
      double r = recip(d);
      if (MASK)
        r = (d > kLimitTwo && (wr * wr + wi2) * r < kLimitOne) ? r : 0.0;
      Bs[buf][sr][sc + c] =
          make_double2(fma(Ar, dr, -Ai * wi) * r, fma(Ai, dr, Ar * wi) * r);
    }
  };
  const int nst = (k_end - k_begin + kLogStage - 1) / kLogStage;
  if (nst > 0) {
    gload(k_begin);
    lstore(0);
  }
  __syncthreads();
  for (int st = 0; st < nst; ++st) {
    const int buf = st & 1;
    const bool more = st + 1 < nst;
    if (more) gload(k_begin + (st + 1) * kLogStage);
    asm volatile("" ::: "memory");
#pragma unroll
    for (int kk = 0; kk < kLogStage / 4; ++kk) {
      const int kr = 4 * kk + lk;
      double2 a[2], b[2];
#pragma unroll
      for (int mb = 0; mb < 2; ++mb) a[mb] = As[buf][wm + 16 * mb + li][kr];
#pragma unroll
      for (int nb = 0; nb < 2; ++nb) b[nb] = Bs[buf][wn + 16 * nb + li][kr];
      mfma3_step<2, false>(a, b, t1, t2, t3);
    }
    if (more) lstore(buf ^ 1);
    __syncthreads();
  }
  dbl4 qr[2][2], qi[2][2];
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int nb = 0; nb < 2; ++nb) {
      qr[mb][nb] = t1[mb][nb] - t2[mb][nb];
      qi[mb][nb] = (t3[mb][nb] - t1[mb][nb]) - t2[mb][nb];
    }

  // Epilogue: row i (band) of a tile sits in lanes 16*(i%4) + j, element
  // i/4; the 16 lanes of a row hold its 16 G' columns.
  __shared__ double red[2][kLogM][2];
  double v[2][4][2];
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int i = wm + 16 * mb + lk + 4 * e;  // band within the tile
      const int n1 = n1_begin + i0 + i;
      const bool brow = i0 + i < nbu;
      double sr2 = 0.0, si2 = 0.0;
#pragma unroll
      for (int nb = 0; nb < 2; ++nb) {
        const int my_igp = m0 + wn + 16 * nb + li;
        const int igp = my_igp < NG ? p.igp_of[my_igp] : -1;
        if (brow && igp >= 0) {
          const double wgt = 0.5 * p.vcoul[igp];
          const double2 am = p.aqsmG[(size_t)n1 * NG + my_igp];
          const double Qr = qr[mb][nb][e], Qi = qi[mb][nb][e];
          // conj(am) * Q
          sr2 = fma(wgt, fma(am.x, Qr, am.y * Qi), sr2);
          si2 = fma(wgt, fma(am.x, Qi, -am.y * Qr), si2);
        }
      }
#pragma unroll
      for (int m = 1; m < 16; m <<= 1) {
        sr2 += __shfl_xor(sr2, m, kWave);
        si2 += __shfl_xor(si2, m, kWave);
      }
      v[mb][e][0] = sr2;
      v[mb][e][1] = si2;
    }
  // The two waves of a row half (wn = 0, 32) add in that order.
  if (li == 0 && wn == 32) {
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int i = wm + 16 * mb + lk + 4 * e;
        red[1][i][0] = v[mb][e][0];
        red[1][i][1] = v[mb][e][1];
      }
  }
  __syncthreads();
  if (li == 0 && wn == 0) {
    const size_t rec = (size_t)grp * nsplit + ch;
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int i = wm + 16 * mb + lk + 4 * e;
        const bool in = i0 + i < nbu;
        const double br = v[mb][e][0] + red[1][i][0];
        const double bi = v[mb][e][1] + red[1][i][1];
        if (ACH) {
          red[0][i][0] = in ? br : 0.0;
          red[0][i][1] = in ? bi : 0.0;
        } else if (in) {
          double *o = n1log + 2 * (rec * nbu + i0 + i);
          o[0] = br;
          o[1] = bi;
        }
      }
  }
  if (ACH) {
    __syncthreads();
    double *o = n1log + (((size_t)blockIdx.z * gridDim.y + blockIdx.y) *
                             gridDim.x + blockIdx.x) * kSlots;
    if (threadIdx.x == 0) {
      double sr = 0.0, si = 0.0;
      for (int i = 0; i < kLogM; ++i) {
        sr += red[0][i][0];
        si += red[0][i][1];
      }
      o[0] = -sr;
      o[1] = -si;
    } else if (threadIdx.x >= 2 && threadIdx.x < kSlots) {
      o[threadIdx.x] = 0.0;
    }
  }
}

// Per-band log reduction: out[n1] = sum over the unoccupied-kernel blocks
// (records [rec][n1 - u_begin]) plus sum over the occupied-kernel blocks
// (records [rec][n1]), each in record order.  One block per band.
__global__ __launch_bounds__(256) void reduce_n1_kernel(
    const double *__restrict__ src_u, int nrec_u, int u_begin, int nb_u,
    const double *__restrict__ src_o, int nrec_o, int nb_o,
    double *__restrict__ out) {
  __shared__ double buf[2][256];
  const int n1 = blockIdx.x;
  double ar = 0.0, ai = 0.0;
  if (n1 >= u_begin && n1 - u_begin < nb_u) {
    const int j = n1 - u_begin;
    for (int r = threadIdx.x; r < nrec_u; r += 256) {
      ar += src_u[2 * ((size_t)r * nb_u + j)];
      ai += src_u[2 * ((size_t)r * nb_u + j) + 1];
    }
  }
  if (n1 < nb_o) {
    for (int r = threadIdx.x; r < nrec_o; r += 256) {
      ar += src_o[2 * ((size_t)r * nb_o + n1)];
      ai += src_o[2 * ((size_t)r * nb_o + n1) + 1];
    }
  }
  buf[0][threadIdx.x] = ar;
  buf[1][threadIdx.x] = ai;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      buf[0][threadIdx.x] += buf[0][threadIdx.x + h];
      buf[1][threadIdx.x] += buf[1][threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * n1] = buf[0][0];
    out[2 * n1 + 1] = buf[1][0];
  }
}

// ---------------------------------------------------------------------------
// Fixed-order sum of `nrec` records of kSlots doubles -> out[kSlots].
// One block per slot: thread t sums records t, t+256, ... in order
// (Neumaier-compensated), then a fixed pairwise tree over the 256 partials.
constexpr int kRedThreads = 256;

// Neumaier-compensated acc += x.
__device__ __forceinline__ void neumaier_add(double &acc, double &c, double x) {
  const double t = acc + x;
  c += (fabs(acc) >= fabs(x)) ? ((acc - t) + x) : ((x - t) + acc);
  acc = t;
}

// This thread's records r = threadIdx.x + k * kRedThreads, k = 0, 1, ...
// in that order, slot s, compensated.  Eight loads are issued before their
// eight adds, so a thread waits for memory once per eight records (the
// order of the adds, and so the result, is that of the plain loop).
__device__ __forceinline__ double strided_sum(const double *__restrict__ part,
                                              int nrec, int s) {
  double acc = 0.0, c = 0.0;
  int r = threadIdx.x;
  constexpr int U = 8;
  for (; r + (U - 1) * kRedThreads < nrec; r += U * kRedThreads) {
    double x[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
      x[u] = part[(size_t)(r + u * kRedThreads) * kSlots + s];
#pragma unroll
    for (int u = 0; u < U; ++u) neumaier_add(acc, c, x[u]);
  }
  for (; r < nrec; r += kRedThreads)
    neumaier_add(acc, c, part[(size_t)r * kSlots + s]);
  return acc + c;
}

__global__ __launch_bounds__(kRedThreads) void reduce_kernel(
    const double *__restrict__ partials, int nrec, double *__restrict__ out) {
  __shared__ double buf[kRedThreads];
  const int s = blockIdx.x;
  buf[threadIdx.x] = strided_sum(partials, nrec, s);
  __syncthreads();
  for (int h = kRedThreads / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) buf[threadIdx.x] += buf[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[s] = buf[0];
}

// ---------------------------------------------------------------------------
// Static remainder achstemp, gppker form (gppKer.cpp:64-120 with the
// shared-accumulator race removed): per band and my_igp, the igp term
//   0.5 * vcoul[igp] * aqsn[n1][igp] * conj(aqsm[n1][igp]) * I_eps[my_igp][igp]
// taken when |I_eps[my_igp][igp]| > to1 (:96-97).  One block per my_igp,
// threads over bands; record slot 0/1.
constexpr int kAchThreads = 256;
__device__ __forceinline__ void achstemp_gppker_block(const DevProblem &p,
                                                      int my_igp,
                                                      double *__restrict__ rec) {
  const int igp = p.igp_of[my_igp];
  double vr = 0.0, vi = 0.0;
  if (igp >= 0) {  // block-uniform
    const double2 s = p.ieps[(size_t)my_igp * p.NC + igp];
    if (sqrt(s.x * s.x + s.y * s.y) > kTo1) {
      double tr = 0.0, ti = 0.0;
      for (int n1 = threadIdx.x; n1 < p.NB; n1 += kAchThreads) {
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
  double v[2] = {vr, vi};
  block_store<2>(v, rec);
  if (threadIdx.x >= 2 && threadIdx.x < kSlots) rec[threadIdx.x] = 0.0;
}

__global__ __launch_bounds__(kAchThreads) void achstemp_gppker_kernel(
    DevProblem p, double *__restrict__ partials, int rec_base) {
  achstemp_gppker_block(p, blockIdx.x,
                        partials + (size_t)(rec_base + blockIdx.x) * kSlots);
}

// The F90 form of the static remainder (gppkernel.deslippe.09.30.2014.f90:
// 195-229, every G-vector) runs on the matrix cores: n1log_gemm_kernel<false,
// true>, launched per problem by launch_achstemp (batches included).
__host__ __device__ inline int achstemp_nrec(const DevProblem &p) {
  if (p.NG <= 0 || p.NC <= 0 || p.NB <= 0) return 0;
  return p.NG;
}

// All problems of a batch, gppKer.cpp form: problem blockIdx.y, record
// rec_begin[prob] + blockIdx.x (blocks past a problem's record count return
// at once).
__global__ __launch_bounds__(kAchThreads) void achstemp_batch_kernel(
    const DevProblem *__restrict__ probs, const int *__restrict__ rec_begin,
    double *__restrict__ partials) {
  const DevProblem p = probs[blockIdx.y];
  if ((int)blockIdx.x >= achstemp_nrec(p)) return;  // block-uniform
  double *rec = partials + (size_t)(rec_begin[blockIdx.y] + blockIdx.x) * kSlots;
  achstemp_gppker_block(p, blockIdx.x, rec);
}

// ---------------------------------------------------------------------------
// DevProblem with its arrays typed as global memory (address space 1).
struct DevProblemG {
  GlobalPtr<double2> aqsm, aqsn, ieps, wtilde;
  GlobalPtr<double> vcoul;
  GlobalPtr<int> igp_of;
  GlobalPtr<double> wx;
  GlobalPtr<float2> aqsn32;
  GlobalPtr<double2> aqsmG;
  int NB, NV, NC, NG;
};

__device__ __forceinline__ DevProblemG global_view(const DevProblem &p) {
  DevProblemG g;
  g.aqsm = (GlobalPtr<double2>)p.aqsm;
  g.aqsn = (GlobalPtr<double2>)p.aqsn;
  g.ieps = (GlobalPtr<double2>)p.ieps;
  g.wtilde = (GlobalPtr<double2>)p.wtilde;
  g.vcoul = (GlobalPtr<double>)p.vcoul;
  g.igp_of = (GlobalPtr<int>)p.igp_of;
  g.wx = (GlobalPtr<double>)p.wx;
  g.aqsn32 = (GlobalPtr<float2>)p.aqsn32;
  g.aqsmG = (GlobalPtr<double2>)p.aqsmG;
  g.NB = p.NB;
  g.NV = p.NV;
  g.NC = p.NC;
  g.NG = p.NG;
  return g;
}

// Batched launch: one grid over a work list of (problem, block) items of
// all problems, sorted largest-first on the host, so the hardware
// dispatcher hands the longest blocks out first and the short ones fill the
// tail.  Block b runs item b and writes that item's own record, so the sums
// stay in a fixed order.  (A persistent loop that pulls items from an
// atomic counter was measured too: the loop made the compiler keep 250
// VGPRs instead of 164, one wave per SIMD less, and it ran slower.)
// Non-occupied and occupied items are two launches (KIND 0 / 1) so each
// keeps its own register budget (capped at 3 waves per SIMD for K <= 8, as
// for unocc_fp64_kernel).
template <int K, int PREC, int SEM, int KIND>
__global__ __launch_bounds__(256) __attribute__((
    amdgpu_waves_per_eu(K <= 8 ? 3 : 1))) void batch_kernel(
    const DevProblem *__restrict__ probs, const BatchItem *__restrict__ items,
    double *__restrict__ partials, float thresh,
    unsigned long long *__restrict__ fallback_count) {
  const BatchItem w = items[blockIdx.x];
  // Pointers read from memory are generic (flat) to the compiler; the
  // global-address-space view lets uniform loads become scalar loads, as
  // they are for the kernel-argument problem of the direct launches.
  const DevProblemG p = global_view(probs[w.prob]);
  double *rec = partials + (size_t)w.rec * kSlots;
  if (KIND == 0) {
    if (PREC == 0)
      unocc_fp64_block<K, SEM == kSemBgw, false>(p, p.NV, p.NB, w.block, rec,
                                                 nullptr);
    else
      unocc_mixed_block<K, SEM == kSemBgw, false, PREC == kPrecMixedKahan>(
          p, p.NV, p.NB, w.block, rec, nullptr, thresh, fallback_count);
  } else {
    occ_block<SEM, false>(p, p.NV, w.block, rec, nullptr);
  }
}

// Clears the fallback counter with a kernel store (capturable in a graph).
__global__ void batch_reset_kernel(unsigned long long *fallback_count) {
  if (threadIdx.x == 0) *fallback_count = 0ull;
}

// One block per problem: fixed-order sum of its records (as reduce_kernel).
__global__ __launch_bounds__(kRedThreads) void reduce_batch_kernel(
    const double *__restrict__ partials, const int *__restrict__ rec_begin,
    const int *__restrict__ rec_count, double *__restrict__ out) {
  __shared__ double buf[kRedThreads];
  const double *part = partials + (size_t)rec_begin[blockIdx.x] * kSlots;
  const int nrec = rec_count[blockIdx.x];
  for (int s = 0; s < kSlots; ++s) {
    buf[threadIdx.x] = strided_sum(part, nrec, s);
    __syncthreads();
    for (int h = kRedThreads / 2; h > 0; h >>= 1) {
      if (threadIdx.x < h) buf[threadIdx.x] += buf[threadIdx.x + h];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[(size_t)blockIdx.x * kSlots + s] = buf[0];
    __syncthreads();
  }
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

hipError_t launch_gather_aqsm(const double2 *aqsm, const int *igp_of, int NB,
                              int NC, int NG, double2 *out, hipStream_t s) {
  const size_t n = (size_t)NB * NG;
  if (n == 0) return hipSuccess;
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  gather_aqsm_kernel<<<(unsigned)blocks, 256, 0, s>>>(aqsm, igp_of, NB, NC, NG,
                                                      out);
  return hipGetLastError();
}

hipError_t launch_aqsn_to_fp32(const double2 *aqsn, float2 *out, int NB,
                               int NC, hipStream_t s) {
  const size_t n = (size_t)NB * padded_nc(NC);
  if (n == 0) return hipSuccess;
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  to_fp32_kernel<<<(unsigned)blocks, 256, 0, s>>>(aqsn, out, NB, NC);
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
                                 double *n1log, unsigned long long *fb,
                                 int glo, int gcnt, int grid, hipStream_t s) {
  const dim3 blk(kWave * waves);
  const bool lg = n1log != nullptr;
  if (precision == 0) {
    if (sem == kSemBgw) {
      if (lg)
        unocc_fp64_kernel<K, true, true>
            <<<grid, blk, 0, s>>>(p, n1b, n1e, glo, gcnt, partials, n1log);
      else
        unocc_fp64_kernel<K, true, false>
            <<<grid, blk, 0, s>>>(p, n1b, n1e, glo, gcnt, partials, n1log);
    } else {
      if (lg)
        unocc_fp64_kernel<K, false, true>
            <<<grid, blk, 0, s>>>(p, n1b, n1e, glo, gcnt, partials, n1log);
      else
        unocc_fp64_kernel<K, false, false>
            <<<grid, blk, 0, s>>>(p, n1b, n1e, glo, gcnt, partials, n1log);
    }
  } else if (sem != kSemBgw) {
    // the unmasked mixed kernel is built in gpp_mixed.hip (its own scheduler)
    return launch_unocc_mixed(p, n1b, n1e, K, precision == kPrecMixedKahan,
                              waves, thresh, partials, n1log, fb, glo, gcnt,
                              grid, s);
  } else {
    const float th = (float)thresh;
#define GPP_MIX(L, KA)                                                      \
  unocc_mixed_kernel<K, true, L, KA><<<grid, blk, 0, s>>>(p, n1b, n1e, glo, \
                                                          gcnt, partials,   \
                                                          n1log, th, fb)
    if (precision == kPrecMixedKahan) {
      if (lg) GPP_MIX(true, true); else GPP_MIX(false, true);
    } else {
      if (lg) GPP_MIX(true, false); else GPP_MIX(false, false);
    }
#undef GPP_MIX
  }
  return hipGetLastError();
}

int unocc_groups(const DevProblem &p) {
  return p.NG > 0 ? (p.NG + kWave - 1) / kWave : 0;
}

hipError_t launch_unocc(const DevProblem &p, int n1_begin, int n1_end,
                        int precision, int sem, const TileCfg &t,
                        double res_thresh, double *partials, double *n1log,
                        unsigned long long *fallback_count, hipStream_t s,
                        int grp_lo, int grp_cnt) {
  const int ngrp = unocc_groups(p);
  if (grp_cnt < 0) {
    grp_lo = 0;
    grp_cnt = ngrp;
  }
  if (grp_lo < 0 || grp_lo + grp_cnt > ngrp) return hipErrorInvalidValue;
  const int full = unocc_grid(p, t);
  if (full == 0 || grp_cnt == 0 || n1_begin >= n1_end) return hipSuccess;
  const int grid = full / ngrp * grp_cnt;
  // __launch_bounds__(256): at most 4 waves per block.
  if (t.waves < 1 || t.waves > 4) return hipErrorInvalidValue;
  switch (t.k_ig) {
#define GPP_CASE(KK)                                                        \
  case KK:                                                                  \
    return launch_unocc_k<KK>(p, n1_begin, n1_end, precision, sem, t.waves, \
                              res_thresh, partials, n1log, fallback_count,  \
                              grp_lo, grp_cnt, grid, s);
    GPP_CASE(4)
    GPP_CASE(8)
    GPP_CASE(16)
    GPP_CASE(32)
#undef GPP_CASE
    default:
      return hipErrorInvalidValue;
  }
}

int n1log_gemm_records(const DevProblem &p) {
  if (p.NG <= 0 || p.NC <= 0) return 0;
  return unocc_groups(p) * ((p.NC + kLogChunk - 1) / kLogChunk);
}

hipError_t launch_n1log_gemm(const DevProblem &p, int n1_begin, int n1_end,
                             int sem, double *n1log, hipStream_t s,
                             int grp_lo, int grp_cnt) {
  const int ngrp = unocc_groups(p);
  if (grp_cnt < 0) {
    grp_lo = 0;
    grp_cnt = ngrp;
  }
  if (grp_lo < 0 || grp_lo + grp_cnt > ngrp) return hipErrorInvalidValue;
  if (ngrp == 0 || grp_cnt == 0 || n1_begin >= n1_end) return hipSuccess;
  const dim3 grid((n1_end - n1_begin + kLogM - 1) / kLogM, grp_cnt,
                  (p.NC + kLogChunk - 1) / kLogChunk);
  if (grid.x > 65535u * 1024u || grid.y > 65535u || grid.z > 65535u)
    return hipErrorInvalidValue;
  if (sem == kSemBgw)
    n1log_gemm_kernel<true><<<grid, kGemmThreads, 0, s>>>(p, n1_begin, n1_end,
                                                         grp_lo, n1log);
  else
    n1log_gemm_kernel<false><<<grid, kGemmThreads, 0, s>>>(
        p, n1_begin, n1_end, grp_lo, n1log);
  return hipGetLastError();
}

int unocc_gemm_grid(const DevProblem &p) {
  if (p.NG <= 0 || p.NC <= 0) return 0;
  return unocc_groups(p) * ((p.NC + kGemmN - 1) / kGemmN);
}

hipError_t launch_unocc_gemm(const DevProblem &p, int n1_begin, int n1_end,
                             int sem, double *partials, hipStream_t s,
                             int grp_lo, int grp_cnt) {
  static_assert(kGemmM == kWave, "a block covers one group of G' columns");
  const int ngrp = unocc_groups(p);
  if (grp_cnt < 0) {
    grp_lo = 0;
    grp_cnt = ngrp;
  }
  if (grp_lo < 0 || grp_lo + grp_cnt > ngrp) return hipErrorInvalidValue;
  const int full = unocc_gemm_grid(p);
  if (full == 0 || grp_cnt == 0 || n1_begin >= n1_end) return hipSuccess;
  const int grid = full / ngrp * grp_cnt;
  if (sem == kSemBgw)
    unocc_gemm_kernel<true><<<grid, kGemmThreads, 0, s>>>(
        p, n1_begin, n1_end, grp_lo, grp_cnt, partials);
  else
    unocc_gemm_kernel<false><<<grid, kGemmThreads, 0, s>>>(
        p, n1_begin, n1_end, grp_lo, grp_cnt, partials);
  return hipGetLastError();
}

int occ_grid(const DevProblem &p) {
  if (p.NG <= 0 || p.NC <= 0) return 0;
  return p.NG * ((p.NC + kOccChunk - 1) / kOccChunk);
}

template <int SEM>
static void launch_occ_sem(const DevProblem &p, int n1_end, double *partials,
                           double *n1log, bool fact, int grid, hipStream_t s) {
  if (n1log)
    occ_kernel<SEM, true><<<grid, kOccThreads, 0, s>>>(p, n1_end, partials,
                                                       n1log);
  else if (fact)
    occ_kernel<SEM, false, true><<<grid, kOccThreads, 0, s>>>(
        p, n1_end, partials, n1log);
  else
    occ_kernel<SEM, false><<<grid, kOccThreads, 0, s>>>(p, n1_end, partials,
                                                        n1log);
}

hipError_t launch_occ(const DevProblem &p, int n1_end, int sem,
                      double *partials, double *n1log, hipStream_t s,
                      bool shared_wx) {
  const int grid = occ_grid(p);
  if (grid == 0 || n1_end <= 0) return hipSuccess;
  switch (sem) {
    case kSemGppker:
      launch_occ_sem<kSemGppker>(p, n1_end, partials, n1log, shared_wx, grid,
                                 s);
      break;
    case kSemGppkerStrict:
      launch_occ_sem<kSemGppkerStrict>(p, n1_end, partials, n1log, shared_wx,
                                       grid, s);
      break;
    case kSemBgw:
      launch_occ_sem<kSemBgw>(p, n1_end, partials, n1log, shared_wx, grid, s);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_reduce(const double *partials, int nrec, double *out,
                         hipStream_t s) {
  reduce_kernel<<<kSlots, kRedThreads, 0, s>>>(partials, nrec, out);
  return hipGetLastError();
}

hipError_t launch_reduce_n1(const double *src_u, int nrec_u, int u_begin,
                            int nb_u, const double *src_o, int nrec_o,
                            int nb_o, int NB, double *out, hipStream_t s) {
  if (NB <= 0) return hipSuccess;
  reduce_n1_kernel<<<NB, 256, 0, s>>>(src_u, nrec_u, u_begin, nb_u, src_o,
                                      nrec_o, nb_o, out);
  return hipGetLastError();
}

// The bgw static remainder of one problem runs on the matrix cores
// (n1log_gemm_kernel<false, true>): one record per block of that grid.
static dim3 achstemp_gemm_grid(const DevProblem &p) {
  if (p.NB <= 0 || p.NG <= 0 || p.NC <= 0) return dim3(0, 0, 0);
  return dim3((p.NB + kLogM - 1) / kLogM, unocc_groups(p),
              (p.NC + kLogChunk - 1) / kLogChunk);
}

int achstemp_records(const DevProblem &p, int sem) {
  if (sem == kSemBgw) {
    const dim3 g = achstemp_gemm_grid(p);
    return (int)(g.x * g.y * g.z);
  }
  return achstemp_nrec(p);
}

int achstemp_batch_records(const DevProblem &p, int sem) {
  return sem == kSemBgw ? 0 : achstemp_nrec(p);
}

hipError_t launch_prepare_batch(const DevProblem *probs, const PrepItem *preps,
                                int n, int max_ng_nb3, size_t max_nbng,
                                size_t max_nbnc, bool mixed,
                                unsigned long long *counter, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (n > 65535) return hipErrorInvalidValue;
  auto blocks = [](size_t work) {
    size_t b = (work + 255) / 256;
    return (unsigned)std::max<size_t>(1, std::min<size_t>(b, 4096));
  };
  // prepass_one covers one element per thread: no cap on this grid
  prepass_batch_kernel<<<dim3((unsigned)((max_ng_nb3 + 255) / 256), n), 256, 0,
                          s>>>(probs, preps, counter);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  gather_aqsm_batch_kernel<<<dim3(blocks(max_nbng), n), 256, 0, s>>>(probs);
  e = hipGetLastError();
  if (e != hipSuccess || !mixed) return e;
  to_fp32_batch_kernel<<<dim3(blocks(max_nbnc), n), 256, 0, s>>>(probs);
  return hipGetLastError();
}

hipError_t launch_achstemp_batch(const DevProblem *probs, const int *rec_begin,
                                 int n, int max_rec, int sem,
                                 double *partials, hipStream_t s) {
  if (n <= 0 || max_rec <= 0) return hipSuccess;
  if (n > 65535) return hipErrorInvalidValue;
  // bgw batches launch the matrix-core form per problem (launch_achstemp)
  if (sem == kSemBgw) return hipErrorInvalidValue;
  const dim3 grid((unsigned)max_rec, n);
  achstemp_batch_kernel<<<grid, kAchThreads, 0, s>>>(probs, rec_begin,
                                                     partials);
  return hipGetLastError();
}

hipError_t launch_achstemp(const DevProblem &p, int sem, double *partials,
                           hipStream_t s) {
  const int n = achstemp_records(p, sem);
  if (n == 0) return hipSuccess;
  if (sem == kSemBgw) {
    const dim3 g = achstemp_gemm_grid(p);
    if (g.x > 65535u * 1024u || g.y > 65535u || g.z > 65535u)
      return hipErrorInvalidValue;
    n1log_gemm_kernel<false, true><<<g, kGemmThreads, 0, s>>>(p, 0, p.NB, 0,
                                                             partials);
  } else
    achstemp_gppker_kernel<<<n, kAchThreads, 0, s>>>(p, partials, 0);
  return hipGetLastError();
}

// ---- batch

hipError_t launch_reduce_batch(const double *partials, const int *rec_begin,
                               const int *rec_count, int nprob, double *out,
                               hipStream_t s) {
  if (nprob <= 0) return hipSuccess;
  reduce_batch_kernel<<<nprob, kRedThreads, 0, s>>>(partials, rec_begin,
                                                    rec_count, out);
  return hipGetLastError();
}

hipError_t launch_batch(const DevProblem *probs, const BatchItem *items_u,
                        int n_u, const BatchItem *items_o, int n_o, int k_ig,
                        int precision, int sem, double res_thresh,
                        double *partials, unsigned long long *fallback_count,
                        int unocc_threads, const int *rec_begin,
                        const int *rec_count, int nprob, double *out,
                        hipStream_t s) {
  if (unocc_threads < 64 || unocc_threads > 256 || unocc_threads % 64)
    return hipErrorInvalidValue;
  batch_reset_kernel<<<1, 64, 0, s>>>(fallback_count);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const float thr = (float)res_thresh;
  if (n_u > 0) {
#define GPP_BL(KK)                                                          \
  case KK:                                                                  \
    if (precision != 0 && sem == kSemBgw) {                                 \
      if (precision == kPrecMixedKahan)                                     \
        batch_kernel<KK, kPrecMixedKahan, kSemBgw, 0>                       \
            <<<n_u, unocc_threads, 0, s>>>(probs, items_u, partials, thr,   \
                                           fallback_count);                 \
      else                                                                  \
        batch_kernel<KK, 1, kSemBgw, 0><<<n_u, unocc_threads, 0, s>>>(      \
            probs, items_u, partials, thr, fallback_count);                 \
    } else if (precision == kPrecMixedKahan)                                \
      batch_kernel<KK, kPrecMixedKahan, kSemGppker, 0>                      \
          <<<n_u, unocc_threads, 0, s>>>(probs, items_u, partials, thr,     \
                                         fallback_count);                   \
    else if (precision == 1)                                                \
      batch_kernel<KK, 1, kSemGppker, 0><<<n_u, unocc_threads, 0, s>>>(     \
          probs, items_u, partials, thr, fallback_count);                   \
    else if (sem == kSemGppkerStrict)                                       \
      batch_kernel<KK, 0, kSemGppkerStrict, 0><<<n_u, unocc_threads, 0, s>>>(         \
          probs, items_u, partials, thr, fallback_count);                   \
    else if (sem == kSemBgw)                                                \
      batch_kernel<KK, 0, kSemBgw, 0><<<n_u, unocc_threads, 0, s>>>(                  \
          probs, items_u, partials, thr, fallback_count);                   \
    else                                                                    \
      batch_kernel<KK, 0, kSemGppker, 0><<<n_u, unocc_threads, 0, s>>>(               \
          probs, items_u, partials, thr, fallback_count);                   \
    break;
    switch (k_ig) {
      GPP_BL(4)
      GPP_BL(8)
      GPP_BL(16)
      GPP_BL(32)
      default:
        return hipErrorInvalidValue;
    }
#undef GPP_BL
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (n_o > 0) {
    if (sem == kSemGppkerStrict)
      batch_kernel<4, 0, kSemGppkerStrict, 1><<<n_o, 256, 0, s>>>(
          probs, items_o, partials, thr, fallback_count);
    else if (sem == kSemBgw)
      batch_kernel<4, 0, kSemBgw, 1><<<n_o, 256, 0, s>>>(
          probs, items_o, partials, thr, fallback_count);
    else
      batch_kernel<4, 0, kSemGppker, 1><<<n_o, 256, 0, s>>>(
          probs, items_o, partials, thr, fallback_count);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (nprob > 0) {
    reduce_batch_kernel<<<nprob, kRedThreads, 0, s>>>(partials, rec_begin,
                                                      rec_count, out);
    e = hipGetLastError();
  }
  return e;
}

}  // namespace gpp
