// gpp_internal.h — shared between the HIP kernels (gpp_device.hip) and the
// host-side C ABI (gpp_api.cpp).  Nothing here is part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace gpp {

// Physics constants of the reference (gppKer.cpp:182-188, 234).
constexpr double kTo1 = 1e-6;                    // to1
constexpr double kSexcut = 4.0;                  // sexcut
constexpr double kLimitOne = 1.0 / (kTo1 * 4.0); // limitone = 2.5e5
constexpr double kLimitTwo = 0.25;               // limittwo = 0.5^2
constexpr double kOcc = 1.0;                     // occ

// Semantics (mirrors GPP_SEM_* in include/gpp_sigma.h).
enum Sem : int { kSemGppker = 0, kSemGppkerStrict = 1, kSemBgw = 2 };

// The precision argument of launch_unocc / launch_batch: 0 FP64, 1 mixed,
// kPrecMixedKahan mixed with a compensated FP32 run sum.
constexpr int kPrecMixedKahan = 3;

// Doubles per band of the prepass's frequency table DevProblem::wx.
constexpr int kWxDoubles = 10;

// Row length of the FP32 copy of aqsn (mixed precision): ncouls rounded up
// to 32 G-vectors with zeros, so a lane run of up to 32 G-vectors is one
// contiguous, in-bounds load.
__host__ __device__ constexpr int padded_nc(int nc) { return (nc + 31) & ~31; }

// Output slot layout of every partial-sum record (doubles):
//   [0..5]  achtemp re/im for iw = 0,1,2
//   [6..11] asxtemp re/im for iw = 0,1,2
constexpr int kSlots = 12;

// Device view of one sigma evaluation, after the prepass.
struct DevProblem {
  const double2 *aqsm;    // [NB][NC]
  const double2 *aqsn;    // [NB][NC]
  const double2 *ieps;    // [NG][NC]
  const double2 *wtilde;  // [NG][NC]
  const double *vcoul;    // [NC]
  const int *igp_of;      // [NG]: resolved igp in [0, NC) or -1 (skipped)
  const double *wx;       // [NB][3], clamped, band-resolved; then [NB][3]
                          // wx^2 / |wx^2|^2 (occupied solver, gppker); then
                          // [NB][8] floats: wx split into FP32 hi + lo per
                          // frequency, a flag (nonzero when a frequency of
                          // the band is not finite) and a zero (mixed
                          // precision).
                          // kWxDoubles per band.
  const float2 *aqsn32;   // [NB][padded_nc(NC)] FP32 copy, zero padded
                          // (mixed precision only)
  const double2 *aqsmG;   // [NB][NG]: aqsm[n1][igp_of[my_igp]], 0 if skipped
  int NB, NV, NC, NG;
};

// The per-problem inputs of the prepass that DevProblem does not carry.
struct PrepItem {
  const int *inv;         // inv_igp_index [NG]
  const int *indinv;      // [NC]
  const double *wx_in;    // [3] or [NB][3]
  int wx_per_band;
  int pad;
};

// Tile shape of the non-occupied kernel.
struct TileCfg {
  int k_ig = 16;          // G-vectors each lane keeps in registers
  int waves = 4;          // waves per block (each takes its own k_ig run)
};

// One work item of the batched launch (block b of the grid runs item b).
struct BatchItem {
  int prob;   // index into the problem table
  int kind;   // 0 = non-occupied block, 1 = occupied block
  int block;  // block index within that kernel's grid for the problem
  int rec;    // global record index of its partial sums
};

// Launchers (gpp_device.hip).  All enqueue on `s` and return hipError_t.
hipError_t launch_prepass(const int *inv_igp_index, const int *indinv,
                          const double *wx_in, int wx_per_band, int NB,
                          int NC, int NG, int *igp_of, double *wx_out,
                          unsigned long long *counter, hipStream_t s);
hipError_t launch_aqsn_to_fp32(const double2 *aqsn, float2 *out, int NB,
                               int NC, hipStream_t s);
hipError_t launch_gather_aqsm(const double2 *aqsm, const int *igp_of, int NB,
                              int NC, int NG, double2 *out, hipStream_t s);
// Blocks of the non-occupied launch over all G' columns (= records).
int unocc_grid(const DevProblem &p, const TileCfg &t);
// G' groups of 64 columns (a launch may cover a contiguous range of them).
int unocc_groups(const DevProblem &p);
// n1log: NULL, or [grid][n1_end - n1_begin] complex per-band records.
// grp_cnt < 0: all G' groups; else groups [grp_lo, grp_lo + grp_cnt), whose
// records land where the launch over all groups puts them.
hipError_t launch_unocc(const DevProblem &p, int n1_begin, int n1_end,
                        int precision, int sem, const TileCfg &t,
                        double res_thresh, double *partials, double *n1log,
                        unsigned long long *fallback_count, hipStream_t s,
                        int grp_lo = 0, int grp_cnt = -1);
// The unmasked mixed-precision kernel of launch_unocc (gpp_mixed.hip): tile
// k_ig, `waves` waves per block, a grid of `grid` blocks over G' groups
// [grp_lo, grp_lo + grp_cnt).
hipError_t launch_unocc_mixed(const DevProblem &p, int n1_begin, int n1_end,
                              int k_ig, bool kahan, int waves,
                              double res_thresh, double *partials,
                              double *n1log,
                              unsigned long long *fallback_count, int grp_lo,
                              int grp_cnt, int grid, hipStream_t s);
// Band-sum-first variant (wx shared by all bands): one block per 64 G'
// columns x 64 G-vectors; same record and group conventions as
// launch_unocc.
int unocc_gemm_grid(const DevProblem &p);
hipError_t launch_unocc_gemm(const DevProblem &p, int n1_begin, int n1_end,
                             int sem, double *partials, hipStream_t s,
                             int grp_lo = 0, int grp_cnt = -1);
// Per-band log of the band-sum-first mode, non-occupied bands
// [n1_begin, n1_end): records (group, G-vector chunk), each
// [n1_end - n1_begin] complex, as launch_reduce_n1 reads them; a launch
// over groups [grp_lo, grp_lo + grp_cnt) writes their records only.
int n1log_gemm_records(const DevProblem &p);
hipError_t launch_n1log_gemm(const DevProblem &p, int n1_begin, int n1_end,
                             int sem, double *n1log, hipStream_t s,
                             int grp_lo = 0, int grp_cnt = -1);
int occ_grid(const DevProblem &p);
// n1log: NULL, or [grid][n1_end] complex per-band records.
// shared_wx: every band has the same frequencies; the bands' products are
// then summed before the solver, which runs once per (my_igp, ig) (no n1log).
hipError_t launch_occ(const DevProblem &p, int n1_end, int sem,
                      double *partials, double *n1log, hipStream_t s,
                      bool shared_wx = false);
hipError_t launch_reduce(const double *partials, int nrec, double *out,
                         hipStream_t s);
hipError_t launch_reduce_n1(const double *src_u, int nrec_u, int u_begin,
                            int nb_u, const double *src_o, int nrec_o,
                            int nb_o, int NB, double *out, hipStream_t s);
int achstemp_records(const DevProblem &p, int sem);
// Records of the batched gppKer.cpp form (achstemp_batch_kernel's layout;
// 0 for bgw, whose batches launch launch_achstemp per problem).
int achstemp_batch_records(const DevProblem &p, int sem);
// Prepass, aqsm gather and (mixed) FP32 copy of aqsn for problems
// [0, n) of a device table, one launch each (problem = blockIdx.y);
// max_ng_nb3 = max over problems of max(NG, 3 NB), max_nbng / max_nbnc the
// largest NB*NG / NB*padded_nc(NC).
hipError_t launch_prepare_batch(const DevProblem *probs, const PrepItem *preps,
                                int n, int max_ng_nb3, size_t max_nbng,
                                size_t max_nbnc, bool mixed,
                                unsigned long long *counter, hipStream_t s);
// achstemp records of problems [0, n): problem i writes records
// rec_begin[i] + [0, achstemp_records(i)); max_rec is the largest count.
hipError_t launch_achstemp_batch(const DevProblem *probs, const int *rec_begin,
                                 int n, int max_rec, int sem,
                                 double *partials, hipStream_t s);
hipError_t launch_achstemp(const DevProblem &p, int sem, double *partials,
                           hipStream_t s);
// One reduction block per problem: records [rec_begin, +rec_count) ->
// out[prob][12].
hipError_t launch_reduce_batch(const double *partials, const int *rec_begin,
                               const int *rec_count, int nprob, double *out,
                               hipStream_t s);
// Two flat launches (non-occupied items, then occupied items), one block per
// item of each list, then launch_reduce_batch into out.
hipError_t launch_batch(const DevProblem *probs, const BatchItem *items_u,
                        int n_u, const BatchItem *items_o, int n_o, int k_ig,
                        int precision, int sem, double res_thresh,
                        double *partials, unsigned long long *fallback_count,
                        int unocc_threads, const int *rec_begin,
                        const int *rec_count, int nprob, double *out,
                        hipStream_t s);

}  // namespace gpp
