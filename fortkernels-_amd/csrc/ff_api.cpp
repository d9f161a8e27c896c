// ff_api.cpp — the C ABI of include/ff_sigma.h: the full-frequency (FF)
// self-energy terms of FFKer_cori on the GPU.
//
// Per call:
//   1. device, on a stream of the handle's own: the frequency-grid terms
//      (they need no plan, so they start first and run through 2-4);
//   2. plan on the host: for every (band, evaluation frequency) the grid
//      search of the chosen variant (compute_fact, ffKernel.cpp:28-78, or
//      ffKernel.f90:231-245,391-430) gives two frequency slices and their
//      weights; the distinct (array, slice, band) triples are collected
//      and packed into passes of up to 8 bands per slice;
//   3. device, on the caller's stream: per pass, the coefficients
//      vcoul * conj(aqsm) of its bands, then one HBM pass over the slice
//      computes the per-slice sums T(band, slice);
//   4. host: the outputs are the weighted combinations of the T values
//      (the interpolations are linear) plus the grid terms.
// Host inputs are copied through the handle's pinned ring, and only the
// I_eps slices the evaluation frequencies touch are copied.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "gpp_guard.h"
#include "../../include/ff_sigma.h"
#include "ff_internal.h"
#include "gpp_handle.h"
#include "host_pool.h"

using gpp::Arena;
using gpp::ensure;
using gpp::ensure_host_out;
using gpp::set_err;
using gpp::ff::SliceJob;

namespace {

enum Kind { kR = 0, kA = 1 };

// Interpolation between slices f and f+1 with weights f1, f2; f < 0: the
// variant computes nothing for this (band, frequency).
struct Fact {
  int f = -1;
  double f1 = 0.0, f2 = 0.0;
};

// compute_fact's search (ffKernel.cpp:30-35): the last k with
// g[k] < x < g[k+1], else 0.
int search_open(double x, int nF, const double *g) {
  int f = 0;
  for (int k = 0; k < nF - 1; ++k)
    if (x > g[k] && x < g[k + 1]) f = k;
  return f;
}

// The F90 search (ffKernel.f90:231-245): the last k with
// g[k] <= x < g[k+1], else the last interval.  For a strictly increasing
// grid at most one interval matches, so a binary search finds it.
int search_half_open(double x, int nF, const double *g, bool sorted) {
  if (sorted) {
    if (!(x >= g[0] && x < g[nF - 1])) return nF - 2;
    int lo = 0, hi = nF - 1;  // g[lo] <= x < g[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) / 2;
      if (g[mid] <= x)
        lo = mid;
      else
        hi = mid;
    }
    return lo;
  }
  int f = -1;
  for (int k = 0; k < nF - 1; ++k)
    if (x >= g[k] && x < g[k + 1]) f = k;
  return f < 0 ? nF - 2 : f;
}

// The same search for a strictly increasing grid when consecutive queries
// move little (the evaluation frequencies of one band step by
// freqevalstep): walk from the last interval instead of bisecting, which
// keeps the host planning of 10^5-10^6 (band, frequency) pairs cheap.
struct Walker {
  int k = 0;
  int find(double x, int nF, const double *g) {
    if (!(x >= g[0] && x < g[nF - 1])) return nF - 2;
    while (x < g[k]) --k;       // stops at k = 0 since x >= g[0]
    while (x >= g[k + 1]) ++k;  // stops at k = nF-2 since x < g[nF-1]
    return k;
  }
};

Fact weights(int f, double x, double scale, const double *g) {
  Fact r;
  r.f = f;
  const double d = g[f + 1] - g[f];
  r.f1 = scale * (g[f + 1] - x) / d;
  r.f2 = scale * (x - g[f]) / d;
  return r;
}

// asx (loop 1) and cor (loop 2) weights of the two variants.  `w` (sorted
// grids only) carries the F90 search position between calls.
Fact fact(int sem, int loop, bool occ, double wx, int nF, const double *g,
          bool sorted, Walker *w) {
  Fact r;
  auto half_open = [&](double x) {
    return (sorted && w) ? w->find(x, nF, g) : search_half_open(x, nF, g, sorted);
  };
  if (sem == FF_SEM_CPP) {
    if (loop == 1) {  // called with flag_occ = 0 (ffKernel.cpp:146)
      const bool pos = wx > 0.0;
      int f = search_open(pos ? wx : -wx, nF, g);
      if (f == 0) f = nF - 2;
      if (pos) return weights(f, wx, 1.0, g);
      r = weights(f, -wx, 1.0, g);  // (g[f+1] + wx)/d, (-g[f] - wx)/d
      return r;
    }
    if (wx > 0.0) {
      if (occ) return r;  // achDtemp_cor_Kernel calls nothing (:228)
      return weights(search_open(wx, nF, g), wx, -0.5, g);  // no fallback
    }
    if (!occ) return r;
    int f = search_open(-wx, nF, g);
    if (f == 0) f = nF - 2;
    return weights(f, -wx, 1.0, g);
  }
  if (loop == 1) {  // occupied bands only
    if (!occ) return r;
    const double x = wx >= 0.0 ? wx : -wx;
    return weights(half_open(x), x, 1.0, g);
  }
  if (wx >= 0.0)
    return weights(half_open(wx), wx, -0.5, g);
  if (!occ) return r;
  return weights(half_open(-wx), -wx, 1.0, g);
}

struct Cplx {
  double re = 0.0, im = 0.0;
  void add(double r, double i) {
    re += r;
    im += i;
  }
};

}  // namespace

// Host planning buffers of one handle, kept between calls so that a
// repeated evaluation neither allocates nor touches fresh pages.
struct gpp::ff::PlanCache {
  gpp::HostPool pool;
  std::vector<int> need;  // [(kind*nF + f)*NB + band] -> job*8 + slot, or -1
  std::vector<Fact> fasx, fcor;
  std::vector<signed char> kasx, pcor;
  std::vector<int> row_jobs, job_off, slice_off;  // per (kind, f) row
  std::vector<SliceJob> jobs;
  std::vector<int> job_slice;
  std::vector<std::pair<int, int>> slices;  // (kind, f)
};

void gpp::ff::free_plan_cache(PlanCache *c) { delete c; }

namespace {

// Sizes the handle's planning buffers (need filled with -1); GPP_ENOMEM
// when the host allocation fails.
int plan_buffers(gpp_handle h, size_t nneed, size_t nasx, size_t ncor,
                 size_t nrows, gpp::ff::PlanCache **out) {
  try {
    if (!h->ff_plan) h->ff_plan = new gpp::ff::PlanCache;
    gpp::ff::PlanCache &c = *h->ff_plan;
    c.need.assign(nneed, -1);
    c.fasx.resize(nasx);
    c.kasx.resize(nasx);
    c.fcor.resize(ncor);
    c.pcor.resize(ncor);
    c.row_jobs.resize(nrows);
    c.job_off.resize(nrows + 1);
    c.slice_off.resize(nrows + 1);
  } catch (...) {
    return set_err(h, GPP_ENOMEM, "host planning buffers (%zu entries)", nneed);
  }
  *out = h->ff_plan;
  return GPP_OK;
}

int validate(gpp_handle h, const ff_problem *p, const ff_options *o,
             int parts) {
  if (!p || !o) return set_err(h, GPP_EINVAL, "null problem or options");
  if (p->number_bands < 0 || p->nvband < 0 || p->ncouls <= 0 ||
      p->ngpown < 0 || p->nFreq < 2 || p->nfreqeval < 0)
    return set_err(h, GPP_EINVAL,
                   "bad shape: number_bands=%d nvband=%d ncouls=%d ngpown=%d "
                   "nFreq=%d nfreqeval=%d (nFreq must be >= 2)",
                   p->number_bands, p->nvband, p->ncouls, p->ngpown, p->nFreq,
                   p->nfreqeval);
  if (o->semantics != FF_SEM_CPP && o->semantics != FF_SEM_F90)
    return set_err(h, GPP_EINVAL, "bad semantics %d", o->semantics);
  if (o->memory != GPP_MEM_HOST && o->memory != GPP_MEM_DEVICE)
    return set_err(h, GPP_EINVAL, "bad memory space %d", o->memory);
  if (parts & ~(FF_PART_ACHS | FF_PART_ASX | FF_PART_ACH | FF_PART_COR))
    return set_err(h, GPP_EINVAL, "bad parts mask %d", parts);
  if (p->number_bands > 65535 || (long long)p->nFreq * p->number_bands > (1LL << 31))
    return set_err(h, GPP_EINVAL, "number_bands or nFreq too large");
  const bool nb = p->number_bands > 0, ng = p->ngpown > 0;
  const bool sums = parts & (FF_PART_ACHS | FF_PART_ASX | FF_PART_COR);
  if (sums && nb && ng &&
      (!p->aqsmtemp || !p->aqsntemp || !p->I_epsR_array || !p->vcoul ||
       !p->inv_igp_index || !p->indinv))
    return set_err(h, GPP_EINVAL, "null input array");
  if ((parts & FF_PART_COR) && nb && ng && !p->I_epsA_array)
    return set_err(h, GPP_EINVAL, "null I_epsA_array");
  if ((parts & FF_PART_ASX) && o->semantics == FF_SEM_F90 && nb && ng &&
      !p->I_epsA_array)
    return set_err(h, GPP_EINVAL, "null I_epsA_array");
  if ((parts & (FF_PART_ASX | FF_PART_ACH | FF_PART_COR)) && nb &&
      (!p->ekq || !p->dFreqGrid))
    return set_err(h, GPP_EINVAL, "null ekq or dFreqGrid");
  if ((parts & FF_PART_ACH) && nb &&
      (!p->pref || !p->dFreqBrd || !p->schDt_matrix))
    return set_err(h, GPP_EINVAL, "null pref, dFreqBrd or schDt_matrix");
  return GPP_OK;
}

// Frequency-grid terms (achDtemp, achDtemp_corb and the grid parts of
// ach2Dtemp/achDtemp_cor) on h->ff_stream, after h->ev[0] (recorded by the
// caller on its stream, behind the inputs): constants, grid kernel,
// reduction, copy of the NE x 8 results to h->host_grid; h->ev[3]/ev[4]
// time the kernels and h->ev[5] marks the results as arrived.  Uses
// h->ff_grid and h->host_grid only, so the slice sums may run meanwhile.
int launch_grid_terms(gpp_handle h, const ff_problem *p, bool host, int sem) {
  int rc = GPP_OK;
  const int NB = p->number_bands, NV = std::min(p->nvband, NB);
  const int nF = p->nFreq, NE = p->nfreqeval;
  if (!h->ff_stream)
    HIP_OR_RET(h, hipStreamCreateWithFlags(&h->ff_stream, hipStreamNonBlocking));
  hipStream_t g = h->ff_stream;
  const int fsplit = gpp::ff::freq_splits(NB, NE, nF);
  Arena a;
  const size_t o_part = a.take((size_t)fsplit * NB * NE * gpp::ff::kFreqSlots * 8);
  const size_t o_fout = a.take((size_t)NE * gpp::ff::kFreqSlots * 8);
  const size_t o_rows = a.take((size_t)NB * nF * gpp::ff::kRowDoubles * 8);
  size_t o_ek = 0, o_gr = 0, o_pr = 0, o_br = 0, o_sd = 0;
  if (host) {
    o_ek = a.take((size_t)NB * 8);
    o_gr = a.take((size_t)nF * 8);
    o_pr = a.take((size_t)nF * 8);
    o_br = a.take((size_t)nF * 16);
    o_sd = a.take((size_t)NB * nF * 16);
  }
  if ((rc = ensure(h, h->ff_grid, a.used))) return rc;
  const size_t nres = (size_t)NE * gpp::ff::kFreqSlots;
  if (h->host_grid_cap < nres) {
    if (h->host_grid) (void)hipHostFree(h->host_grid);
    h->host_grid = nullptr;
    h->host_grid_cap = 0;
    HIP_OR_RET(h, hipHostMalloc((void **)&h->host_grid, nres * sizeof(double),
                                hipHostMallocDefault));
    h->host_grid_cap = nres;
  }
  char *base = (char *)h->ff_grid.ptr;
  gpp::ff::FreqArgs fa{};
  fa.ekq = p->ekq;
  fa.grid = p->dFreqGrid;
  fa.pref = p->pref;
  fa.brd = (const double2 *)p->dFreqBrd;
  fa.schdt = (const double2 *)p->schDt_matrix;
  HIP_OR_RET(h, hipStreamWaitEvent(g, h->ev[0], 0));
  if (host) {
    gpp::Stager *st = gpp::stager(h, &rc);
    if (!st) return rc;
    const struct {
      size_t off;
      const void *src;
      size_t bytes;
    } put[] = {{o_ek, p->ekq, (size_t)NB * 8},
               {o_gr, p->dFreqGrid, (size_t)nF * 8},
               {o_pr, p->pref, (size_t)nF * 8},
               {o_br, p->dFreqBrd, (size_t)nF * 16},
               {o_sd, p->schDt_matrix, (size_t)NB * nF * 16}};
    for (const auto &c : put) HIP_OR_RET(h, st->copy(base + c.off, c.src, c.bytes));
    HIP_OR_RET(h, st->fence(g));
    fa.ekq = (const double *)(base + o_ek);
    fa.grid = (const double *)(base + o_gr);
    fa.pref = (const double *)(base + o_pr);
    fa.brd = (const double2 *)(base + o_br);
    fa.schdt = (const double2 *)(base + o_sd);
  }
  fa.NB = NB;
  fa.NV = NV;
  fa.nFreq = nF;
  fa.NE = NE;
  fa.fmin = p->freqevalmin;
  fa.fstep = p->freqevalstep;
  fa.pref_zb = p->pref_zb;
  HIP_OR_RET(h, hipEventRecord(h->ev[3], g));
  HIP_OR_RET(h, gpp::ff::launch_freq(fa, sem, fsplit, (double *)(base + o_rows),
                                     (double *)(base + o_part), g));
  HIP_OR_RET(h, gpp::ff::launch_freq_reduce((const double *)(base + o_part),
                                            fsplit * NB, NE,
                                            (double *)(base + o_fout), g));
  HIP_OR_RET(h, hipEventRecord(h->ev[4], g));
  HIP_OR_RET(h, hipMemcpyAsync(h->host_grid, base + o_fout, nres * 8,
                               hipMemcpyDeviceToHost, g));
  HIP_OR_RET(h, hipEventRecord(h->ev[5], g));
  return GPP_OK;
}

}  // namespace

extern "C" {

void ff_default_options(ff_options *o) {
  if (!o) return;
  std::memset(o, 0, sizeof *o);
  o->semantics = FF_SEM_CPP;
  o->memory = GPP_MEM_HOST;
}

int ff_sigma_ex(gpp_handle h, const ff_problem *p, const ff_options *o,
                ff_result *r) try {
  if (!h) return GPP_EINVAL;
  if (!r) return set_err(h, GPP_EINVAL, "null result");
  h->err.clear();
  const int parts =
      (o && o->parts) ? o->parts
                      : (FF_PART_ACHS | FF_PART_ASX | FF_PART_ACH | FF_PART_COR);
  int rc = validate(h, p, o, parts);
  if (rc) return rc;
  HIP_OR_RET(h, hipSetDevice(h->device));
  const bool host = o->memory == GPP_MEM_HOST;
  hipStream_t s = (!host && o->stream) ? (hipStream_t)o->stream : h->stream;
  const int sem = o->semantics;
  const int NB = p->number_bands, NV = std::min(p->nvband, NB);
  const int NC = p->ncouls, NG = p->ngpown, nF = p->nFreq, NE = p->nfreqeval;
  const size_t per_slice = (size_t)NG * NC;

  using clk = std::chrono::steady_clock;
  const auto t_plan = clk::now();
  // Host copies of the two small arrays the plan needs.
  std::vector<double> ekq(NB), grid(nF);
  const bool need_grid = parts & (FF_PART_ASX | FF_PART_ACH | FF_PART_COR);
  if (need_grid && NB > 0) {
    const hipMemcpyKind k = host ? hipMemcpyHostToHost : hipMemcpyDeviceToHost;
    HIP_OR_RET(h, hipMemcpy(ekq.data(), p->ekq, NB * sizeof(double), k));
    HIP_OR_RET(h, hipMemcpy(grid.data(), p->dFreqGrid, nF * sizeof(double), k));
  }
  bool sorted = true;
  for (int k = 0; k + 1 < nF && need_grid; ++k)
    sorted = sorted && grid[k] < grid[k + 1];

  // ---- frequency-grid terms: need no plan, so they start first, on a
  // stream of their own, and run while the host plans and the slice sums
  // go through the caller's stream.
  const bool do_ach = (parts & FF_PART_ACH) && NB > 0 && NE > 0;
  // On every return the grid stream has finished with the handle's buffers.
  struct GridJoin {
    gpp_handle h;
    bool active;
    ~GridJoin() {
      if (active && h->ff_stream) (void)hipStreamSynchronize(h->ff_stream);
    }
  } join{h, do_ach};
  HIP_OR_RET(h, hipEventRecord(h->ev[0], s));
  if (do_ach && (rc = launch_grid_terms(h, p, host, sem))) return rc;

  // ---- plan: which (kind, slice, band) sums are needed
  gpp::simulated_host_alloc_failure();
  gpp::ff::PlanCache *pc = nullptr;
  if ((rc = plan_buffers(h, (size_t)2 * nF * NB, (size_t)NV * NE,
                         (size_t)NB * NE, (size_t)2 * nF, &pc)))
    return rc;
  std::vector<int> &need = pc->need;
  // Written only when it changes: the bands of one (kind, f) share a cache
  // line, and the planning threads mark the same (kind, f) over and over
  // (every iw, in the ffKernel.cpp semantics); two tasks of one band may
  // mark one entry, hence the relaxed atomic accesses.
  auto mark = [&](int kind, int f, int band) {
    int *e = &need[((size_t)kind * nF + f) * NB + band];
    if (__atomic_load_n(e, __ATOMIC_RELAXED) != 0)
      __atomic_store_n(e, 0, __ATOMIC_RELAXED);
  };
  const bool sums_possible = NB > 0 && NG > 0;
  const int achs_slice = sem == FF_SEM_CPP ? 1 : 0;
  if ((parts & FF_PART_ACHS) && sums_possible)
    for (int b = 0; b < NB; ++b) mark(kR, achs_slice, b);
  // asx: [NV][NE] weights (C++: the same for every iw); cor: [NB][NE].
  std::vector<Fact> &fasx = pc->fasx, &fcor = pc->fcor;
  std::vector<signed char> &kasx = pc->kasx, &pcor = pc->pcor;
  const bool plan_asx = (parts & FF_PART_ASX) && sums_possible;
  const bool plan_cor = (parts & FF_PART_COR) && sums_possible;
  // Tasks of kPlanChunk evaluation frequencies of one band and one loop,
  // so that the occupied bands' second loop does not set the time; each
  // task starts its grid walk with a binary search.
  constexpr int kPlanChunk = 2048;
  const int nch = (NE + kPlanChunk - 1) / kPlanChunk;
  const int nt_asx = plan_asx ? NV * nch : 0, nt_cor = plan_cor ? NB * nch : 0;
  auto wx_of = [&](int b, int iw) {
    return sem == FF_SEM_CPP ? p->freqevalmin - ekq[b] + p->freqevalstep
                             : p->freqevalmin - ekq[b] + iw * p->freqevalstep;
  };
  // ffKernel.cpp semantics: wx does not depend on iw, so each band's two
  // facts are found once (compute_fact scans the whole grid).
  std::vector<Fact> cpp_fact;
  if (sem == FF_SEM_CPP) {
    try {
      cpp_fact.resize((size_t)2 * NB);
    } catch (...) {
      return set_err(h, GPP_ENOMEM, "host planning buffers");
    }
    for (int b = 0; b < NB; ++b) {
      if (plan_asx && b < NV)
        cpp_fact[2 * b] = fact(sem, 1, true, wx_of(b, 0), nF, grid.data(), sorted, nullptr);
      if (plan_cor)
        cpp_fact[2 * b + 1] = fact(sem, 2, b < NV, wx_of(b, 0), nF, grid.data(), sorted, nullptr);
    }
  }
  auto plan_task = [&](int t) {
    const bool asx_task = t < nt_asx;
    const int u = asx_task ? t : t - nt_asx;
    const int b = u / nch, i0 = (u % nch) * kPlanChunk;
    const int i1 = std::min(NE, i0 + kPlanChunk);
    Walker walk;
    if (sorted)
      walk.k = search_half_open(std::fabs(wx_of(b, i0)), nF, grid.data(), true);
    if (asx_task) {
      for (int iw = i0; iw < i1; ++iw) {
        const double wx = wx_of(b, iw);
        const Fact fc = sem == FF_SEM_CPP
                            ? cpp_fact[2 * b]
                            : fact(sem, 1, true, wx, nF, grid.data(), sorted, &walk);
        const int kind = (sem == FF_SEM_F90 && wx < 0.0) ? kA : kR;
        fasx[(size_t)b * NE + iw] = fc;
        kasx[(size_t)b * NE + iw] = (signed char)kind;
        if (fc.f >= 0) {
          mark(kind, fc.f, b);
          mark(kind, fc.f + 1, b);
        }
      }
    } else {
      const bool occ = b < NV;
      for (int iw = i0; iw < i1; ++iw) {
        const double wx = wx_of(b, iw);
        const Fact fc = sem == FF_SEM_CPP
                            ? cpp_fact[2 * b + 1]
                            : fact(sem, 2, occ, wx, nF, grid.data(), sorted, &walk);
        fcor[(size_t)b * NE + iw] = fc;
        pcor[(size_t)b * NE + iw] =
            (signed char)(sem == FF_SEM_CPP ? wx > 0.0 : wx >= 0.0);
        if (fc.f >= 0)
          for (int kind = kR; kind <= kA; ++kind) {
            mark(kind, fc.f, b);
            mark(kind, fc.f + 1, b);
          }
      }
    }
  };
  parallel_ranges(&pc->pool, nt_asx + nt_cor, (long long)(NV + NB) * NE, 20000,
                  plan_task);

  // ---- pack passes: up to 8 bands per (kind, slice), numbered in (kind,
  // f) order: count each row's passes, take the offsets, then fill the
  // rows on several host threads (each row writes its own range).
  constexpr int kRowBlock = 256;
  const int nrows = 2 * nF, nrb = (nrows + kRowBlock - 1) / kRowBlock;
  parallel_ranges(&pc->pool, nrb, (long long)nrows * NB, 20000, [&](int blk) {
    for (int r = blk * kRowBlock; r < std::min(nrows, (blk + 1) * kRowBlock); ++r) {
      const int *row = &need[(size_t)r * NB];
      int nb = 0;
      for (int b = 0; b < NB; ++b) nb += row[b] >= 0;
      pc->row_jobs[r] = (nb + gpp::ff::kMaxJobBands - 1) / gpp::ff::kMaxJobBands;
    }
  });
  pc->job_off[0] = pc->slice_off[0] = 0;
  for (int r = 0; r < nrows; ++r) {
    pc->job_off[r + 1] = pc->job_off[r] + pc->row_jobs[r];
    pc->slice_off[r + 1] = pc->slice_off[r] + (pc->row_jobs[r] > 0);
  }
  try {
    pc->jobs.assign(pc->job_off[nrows], SliceJob{});
    pc->job_slice.resize(pc->job_off[nrows]);
    pc->slices.resize(pc->slice_off[nrows]);
  } catch (...) {
    return set_err(h, GPP_ENOMEM, "host planning buffers (%d passes)",
                   pc->job_off[nrows]);
  }
  std::vector<SliceJob> &jobs = pc->jobs;
  std::vector<int> &job_slice = pc->job_slice;  // index into the staged slices
  std::vector<std::pair<int, int>> &slices = pc->slices;  // (kind, f)
  parallel_ranges(&pc->pool, nrb, (long long)nrows * NB, 20000, [&](int blk) {
    for (int r = blk * kRowBlock; r < std::min(nrows, (blk + 1) * kRowBlock); ++r) {
      if (pc->row_jobs[r] == 0) continue;
      const int si = pc->slice_off[r];
      slices[si] = {r / nF, r % nF};
      int *row = &need[(size_t)r * NB];
      int j = pc->job_off[r];
      for (int b = 0; b < NB; ++b) {
        if (row[b] < 0) continue;
        if (jobs[j].nb == gpp::ff::kMaxJobBands) ++j;
        row[b] = j * gpp::ff::kMaxJobBands + jobs[j].nb;
        jobs[j].band[jobs[j].nb++] = b;
        job_slice[j] = si;
      }
    }
  });
  const int njobs = (int)jobs.size();
  const double plan_ms =
      std::chrono::duration<double, std::milli>(clk::now() - t_plan).count();
  if (njobs > 65535)
    return set_err(h, GPP_EUNSUPPORTED, "%d slice passes (max 65535)", njobs);

  // ---- device layout (h->ff) and staging
  const int nchunk = gpp::ff::slice_chunks(NC);
  int ngroups = 1;
  if (njobs > 0) {
    const long long per = (long long)njobs * nchunk;
    ngroups = (int)std::min<long long>(
        std::max(1, NG / 4), std::max<long long>(1, (2048 + per - 1) / per));
  }
  const int per_job = ngroups * nchunk;
  Arena a;
  // coefficient table for up to jbatch passes at a time (at most 1 GiB)
  const int jbatch =
      njobs > 0 ? (int)std::max<size_t>(
                      1, std::min<size_t>((size_t)njobs,
                                          ((size_t)1 << 30) /
                                              ((size_t)NG * gpp::ff::kJobCgBytes)))
                : 0;
  const size_t o_cg = a.take((size_t)jbatch * NG * gpp::ff::kJobCgBytes);
  const size_t o_jobs = a.take(jobs.size() * sizeof(SliceJob));
  const size_t o_rec = a.take((size_t)njobs * per_job * gpp::ff::kRecSlots * 8);
  const size_t o_T = a.take((size_t)njobs * gpp::ff::kRecSlots * 8);
  // host inputs: aqsm/aqsn and the index arrays staged here too
  const bool need_aqs = njobs > 0;
  size_t o_aqsm = 0, o_aqsn = 0, o_vc = 0, o_inv = 0, o_ind = 0;
  if (host && need_aqs) {
    o_aqsm = a.take((size_t)NB * NC * 16);
    o_aqsn = a.take((size_t)NB * NC * 16);
    o_vc = a.take((size_t)NC * 8);
    o_inv = a.take((size_t)NG * 4);
    o_ind = a.take((size_t)NC * 4);
  }
  if ((rc = ensure(h, h->ff, a.used))) return rc;
  if (host && !slices.empty() &&
      (rc = ensure(h, h->ff_stage, slices.size() * per_slice * 16)))
    return rc;
  char *base = (char *)h->ff.ptr;
  const double2 *aqsm = (const double2 *)p->aqsmtemp;
  const double2 *aqsn = (const double2 *)p->aqsntemp;
  const double *vcoul = p->vcoul;
  const int *inv = p->inv_igp_index, *indinv = p->indinv;
  for (size_t j = 0; j < jobs.size(); ++j) {
    const int kind = slices[job_slice[j]].first, f = slices[job_slice[j]].second;
    if (host)
      jobs[j].slice = (const double2 *)h->ff_stage.ptr + job_slice[j] * per_slice;
    else
      jobs[j].slice = (const double2 *)(kind == kR ? p->I_epsR_array
                                                   : p->I_epsA_array) +
                      (size_t)f * per_slice;
  }
  if (host) {
    gpp::Stager *st = gpp::stager(h, &rc);
    if (!st) return rc;
    auto put = [&](size_t off, const void *src, size_t bytes) -> int {
      if (bytes && src) HIP_OR_RET(h, st->copy(base + off, src, bytes));
      return GPP_OK;
    };
    if (need_aqs) {
      if ((rc = put(o_aqsm, p->aqsmtemp, (size_t)NB * NC * 16)) ||
          (rc = put(o_aqsn, p->aqsntemp, (size_t)NB * NC * 16)) ||
          (rc = put(o_vc, p->vcoul, (size_t)NC * 8)) ||
          (rc = put(o_inv, p->inv_igp_index, (size_t)NG * 4)) ||
          (rc = put(o_ind, p->indinv, (size_t)NC * 4)))
        return rc;
      aqsm = (const double2 *)(base + o_aqsm);
      aqsn = (const double2 *)(base + o_aqsn);
      vcoul = (const double *)(base + o_vc);
      inv = (const int *)(base + o_inv);
      indinv = (const int *)(base + o_ind);
    }
    for (size_t i = 0; i < slices.size(); ++i) {
      const gpp_c128 *src =
          (slices[i].first == kR ? p->I_epsR_array : p->I_epsA_array) +
          (size_t)slices[i].second * per_slice;
      HIP_OR_RET(h, st->copy((double2 *)h->ff_stage.ptr + i * per_slice, src,
                             per_slice * 16));
    }
    HIP_OR_RET(h, st->fence(s));
  }
  if (njobs > 0)
    HIP_OR_RET(h, hipMemcpyAsync(base + o_jobs, jobs.data(),
                                 jobs.size() * sizeof(SliceJob),
                                 hipMemcpyHostToDevice, s));
  if ((rc = ensure_host_out(h, (size_t)njobs * gpp::ff::kRecSlots + 1)))
    return rc;

  // ---- slice sums
  HIP_OR_RET(h, hipEventRecord(h->ev[1], s));
  if (njobs > 0) {
    double2 *cG = (double2 *)(base + o_cg);
    const SliceJob *djobs = (const SliceJob *)(base + o_jobs);
    for (int j0 = 0; j0 < njobs; j0 += jbatch) {
      const int nj = std::min(jbatch, njobs - j0);
      HIP_OR_RET(h, gpp::ff::launch_cg(djobs + j0, nj, aqsm, vcoul, inv,
                                       indinv, NC, NG, cG, s));
      HIP_OR_RET(h, gpp::ff::launch_slices(
                        djobs + j0, nj, cG, aqsn, NC, NG, ngroups,
                        (double *)(base + o_rec) +
                            (size_t)j0 * per_job * gpp::ff::kRecSlots,
                        s));
    }
    HIP_OR_RET(h, gpp::ff::launch_slice_reduce((const double *)(base + o_rec),
                                               njobs, per_job,
                                               (double *)(base + o_T), s));
  }
  HIP_OR_RET(h, hipEventRecord(h->ev[2], s));
  double *ho = h->host_out;
  if (njobs > 0)
    HIP_OR_RET(h, hipMemcpyAsync(ho, base + o_T,
                                 (size_t)njobs * gpp::ff::kRecSlots * 8,
                                 hipMemcpyDeviceToHost, s));
  HIP_OR_RET(h, hipStreamSynchronize(s));

  // ---- assembly
  const auto t_asm = clk::now();
  auto T = [&](int kind, int f, int b) -> Cplx {
    const int slot = need[((size_t)kind * nF + f) * NB + b];
    Cplx c;
    c.re = ho[(size_t)slot * 2];
    c.im = ho[(size_t)slot * 2 + 1];
    return c;
  };
  auto interp = [&](int kind, const Fact &fc, int b) -> Cplx {
    const Cplx x = T(kind, fc.f, b), y = T(kind, fc.f + 1, b);
    Cplx c;
    c.re = fc.f1 * x.re + fc.f2 * y.re;
    c.im = fc.f1 * x.im + fc.f2 * y.im;
    return c;
  };
  Cplx achs;
  if ((parts & FF_PART_ACHS) && sums_possible)
    for (int b = 0; b < NB; ++b) {
      const Cplx t = T(kR, achs_slice, b);
      achs.add(-0.5 * t.re, -0.5 * t.im);
    }
  r->achsDtemp.re = achs.re;
  r->achsDtemp.im = achs.im;
  std::vector<Cplx> asx(NE), cor(NE), a2(NE);
  // Blocks of evaluation frequencies on several host threads; each entry
  // still adds its bands in order, so the sums are those of one thread.
  constexpr int kAsmBlock = 1024;
  auto assemble = [&](int blk) {
    const int i0 = blk * kAsmBlock, i1 = std::min(NE, i0 + kAsmBlock);
    if ((parts & FF_PART_ASX) && sums_possible)
      for (int b = 0; b < NV; ++b)
        for (int iw = i0; iw < i1; ++iw) {
          const Fact &fc = fasx[(size_t)b * NE + iw];
          if (fc.f < 0) continue;
          const Cplx v = interp(kasx[(size_t)b * NE + iw], fc, b);
          asx[iw].add(p->occ * v.re, p->occ * v.im);
        }
    if ((parts & FF_PART_COR) && sums_possible)
      for (int b = 0; b < NB; ++b)
        for (int iw = i0; iw < i1; ++iw) {
          const Fact &fc = fcor[(size_t)b * NE + iw];
          if (fc.f < 0) continue;
          const Cplx x = interp(kR, fc, b), y = interp(kA, fc, b);
          Cplx v;
          v.re = x.re - y.re;
          v.im = x.im - y.im;
          if (pcor[(size_t)b * NE + iw]) {
            a2[iw].add(v.re, v.im);
            if (b >= NV) cor[iw].add(v.re, v.im);
          } else {
            cor[iw].add(-0.5 * v.re, -0.5 * v.im);
          }
        }
  };
  parallel_ranges(&pc->pool, (NE + kAsmBlock - 1) / kAsmBlock,
                  (long long)NB * NE, 20000, assemble);
  // the frequency-grid results (their stream may still be running)
  double asm_ms =
      std::chrono::duration<double, std::milli>(clk::now() - t_asm).count();
  if (do_ach) {
    HIP_OR_RET(h, hipEventSynchronize(h->ev[5]));
    HIP_OR_RET(h, hipStreamWaitEvent(s, h->ev[5], 0));
    join.active = false;
  }
  const auto t_asm2 = clk::now();
  const double *hf = h->host_grid;
  for (int iw = 0; iw < NE; ++iw) {
    const double *q = do_ach ? hf + (size_t)iw * gpp::ff::kFreqSlots : nullptr;
    if (r->asxDtemp) r->asxDtemp[iw] = gpp_c128{asx[iw].re, asx[iw].im};
    if (r->achDtemp) r->achDtemp[iw] = q ? gpp_c128{q[0], q[1]} : gpp_c128{0, 0};
    if (r->achDtemp_corb)
      r->achDtemp_corb[iw] = q ? gpp_c128{q[2], q[3]} : gpp_c128{0, 0};
    if (r->ach2Dtemp)
      r->ach2Dtemp[iw] = gpp_c128{a2[iw].re + (q ? q[4] : 0.0),
                                  a2[iw].im + (q ? q[5] : 0.0)};
    if (r->achDtemp_cor)
      r->achDtemp_cor[iw] = gpp_c128{cor[iw].re + (q ? q[6] : 0.0),
                                     cor[iw].im + (q ? q[7] : 0.0)};
  }
  asm_ms += std::chrono::duration<double, std::milli>(clk::now() - t_asm2).count();
  r->host_plan_ms = plan_ms;
  r->host_assemble_ms = asm_ms;
  float ms = 0.f;
  HIP_OR_RET(h, hipEventElapsedTime(&ms, h->ev[1], h->ev[2]));
  r->slice_kernel_ms = ms;
  HIP_OR_RET(h, hipEventElapsedTime(&ms, h->ev[0], h->ev[2]));
  r->total_ms = ms;
  r->freq_kernel_ms = 0.0;
  if (do_ach) {
    HIP_OR_RET(h, hipEventElapsedTime(&ms, h->ev[3], h->ev[4]));
    r->freq_kernel_ms = ms;
    HIP_OR_RET(h, hipEventElapsedTime(&ms, h->ev[0], h->ev[5]));
    r->total_ms = std::max(r->total_ms, (double)ms);
  }
  long long nslice_bands = 0;
  for (const SliceJob &j : jobs) nslice_bands += j.nb;
  r->slices = (long long)slices.size();
  r->slice_bytes = (double)jobs.size() * per_slice * 16.0;
  r->slice_terms = nslice_bands * (long long)per_slice;
  r->freq_terms = do_ach ? (long long)NB * nF * NE : 0;
  return GPP_OK;
} catch (...) {
  return gpp::exception_code(h);
}

}  // extern "C"

// ---- drop-ins for FFKer_cori/ffKernel.cpp (FF_SEM_CPP, host memory)

namespace {

int run_cpp(const ff_problem &p, int parts, ff_result *r) {
  gpp::DefaultHandle dh;
  if (dh.rc) return dh.rc;
  ff_options o;
  ff_default_options(&o);
  o.parts = parts;
  return dh.ret(ff_sigma_ex(dh.h, &p, &o, r));
}

void add_into(gpp_c128 *dst, const std::vector<gpp_c128> &v) {
  if (!dst) return;
  for (size_t i = 0; i < v.size(); ++i) {
    dst[i].re += v[i].re;
    dst[i].im += v[i].im;
  }
}

}  // namespace

extern "C" {

int ff_achsDtemp_Kernel(int number_bands, int ngpown, int ncouls,
                        const int *inv_igp_index, const int *indinv,
                        const gpp_c128 *aqsntemp, const gpp_c128 *aqsmtemp,
                        const gpp_c128 *I_epsR_array, const double *vcoul,
                        gpp_c128 *achsDtemp) try {
  if (!achsDtemp) return GPP_EINVAL;
  ff_problem p{};
  p.number_bands = number_bands;
  p.ncouls = ncouls;
  p.ngpown = ngpown;
  p.nFreq = 2;  // the kernel reads I_epsR slice 1 only
  p.aqsmtemp = aqsmtemp;
  p.aqsntemp = aqsntemp;
  p.I_epsR_array = I_epsR_array;
  p.vcoul = vcoul;
  p.inv_igp_index = inv_igp_index;
  p.indinv = indinv;
  ff_result r{};
  const int rc = run_cpp(p, FF_PART_ACHS, &r);
  if (rc) return rc;
  *achsDtemp = r.achsDtemp;  // :128 assigns
  return GPP_OK;
} catch (...) {
  return gpp::exception_code(nullptr);
}

int ff_asxDtemp_Kernel(int nvband, int nfreqeval, int ncouls, int ngpown,
                       int numThreads, int nFreq, double freqevalmin,
                       double freqevalstep, double occ, const double *ekq,
                       const double *dFreqGrid, const int *inv_igp_index,
                       const int *indinv, const gpp_c128 *aqsmtemp,
                       const gpp_c128 *aqsntemp, const double *vcoul,
                       const gpp_c128 *I_epsR_array, gpp_c128 *ssxDittt,
                       gpp_c128 *asxDtemp) try {
  (void)numThreads;
  (void)ssxDittt;
  if (!asxDtemp || nfreqeval < 0) return GPP_EINVAL;
  ff_problem p{};
  p.number_bands = nvband;  // only bands n1 < nvband are read
  p.nvband = nvband;
  p.ncouls = ncouls;
  p.ngpown = ngpown;
  p.nFreq = nFreq;
  p.nfreqeval = nfreqeval;
  p.aqsmtemp = aqsmtemp;
  p.aqsntemp = aqsntemp;
  p.I_epsR_array = I_epsR_array;
  p.vcoul = vcoul;
  p.inv_igp_index = inv_igp_index;
  p.indinv = indinv;
  p.ekq = ekq;
  p.dFreqGrid = dFreqGrid;
  p.freqevalmin = freqevalmin;
  p.freqevalstep = freqevalstep;
  p.occ = occ;
  std::vector<gpp_c128> out(nfreqeval);
  ff_result r{};
  r.asxDtemp = out.data();
  const int rc = run_cpp(p, FF_PART_ASX, &r);
  if (rc) return rc;
  add_into(asxDtemp, out);  // :151 adds
  return GPP_OK;
} catch (...) {
  return gpp::exception_code(nullptr);
}

int ff_achDtemp_Kernel(int number_bands, int nvband, int nfreqeval,
                       int ncouls, int ngpown, int nFreq, double freqevalmin,
                       double freqevalstep, const double *ekq, double pref_zb,
                       const double *pref, const double *dFreqGrid,
                       const gpp_c128 *dFreqBrd, const gpp_c128 *schDt_matrix,
                       gpp_c128 *schDi, gpp_c128 *schDi_cor, gpp_c128 *sch2Di,
                       gpp_c128 *achDtemp) try {
  (void)schDi;
  (void)schDi_cor;
  (void)sch2Di;
  if (!achDtemp || nfreqeval < 0) return GPP_EINVAL;
  ff_problem p{};
  p.number_bands = number_bands;
  p.nvband = nvband;
  p.ncouls = ncouls;
  p.ngpown = ngpown;
  p.nFreq = nFreq;
  p.nfreqeval = nfreqeval;
  p.ekq = ekq;
  p.dFreqGrid = dFreqGrid;
  p.pref = pref;
  p.dFreqBrd = dFreqBrd;
  p.schDt_matrix = schDt_matrix;
  p.freqevalmin = freqevalmin;
  p.freqevalstep = freqevalstep;
  p.pref_zb = pref_zb;
  std::vector<gpp_c128> out(nfreqeval);
  ff_result r{};
  r.achDtemp = out.data();
  const int rc = run_cpp(p, FF_PART_ACH, &r);
  if (rc) return rc;
  add_into(achDtemp, out);  // :206 adds
  return GPP_OK;
} catch (...) {
  return gpp::exception_code(nullptr);
}

int ff_achDtemp_cor_Kernel(int number_bands, int nvband, int nfreqeval,
                           int ncouls, int ngpown, int nFreq,
                           double freqevalmin, double freqevalstep,
                           const double *ekq, const double *dFreqGrid,
                           const int *inv_igp_index, const int *indinv,
                           const gpp_c128 *aqsmtemp, const gpp_c128 *aqsntemp,
                           const double *vcoul, const gpp_c128 *I_epsR_array,
                           const gpp_c128 *I_epsA_array, gpp_c128 *schDi_cor,
                           gpp_c128 *schDi_corb, gpp_c128 *sch2Di,
                           gpp_c128 *ach2Dtemp, gpp_c128 *achDtemp_cor,
                           gpp_c128 *achDtemp_corb) try {
  (void)schDi_cor;
  (void)schDi_corb;
  (void)sch2Di;
  if (nfreqeval < 0) return GPP_EINVAL;
  ff_problem p{};
  p.number_bands = number_bands;
  p.nvband = nvband;
  p.ncouls = ncouls;
  p.ngpown = ngpown;
  p.nFreq = nFreq;
  p.nfreqeval = nfreqeval;
  p.aqsmtemp = aqsmtemp;
  p.aqsntemp = aqsntemp;
  p.I_epsR_array = I_epsR_array;
  p.I_epsA_array = I_epsA_array;
  p.vcoul = vcoul;
  p.inv_igp_index = inv_igp_index;
  p.indinv = indinv;
  p.ekq = ekq;
  p.dFreqGrid = dFreqGrid;
  p.freqevalmin = freqevalmin;
  p.freqevalstep = freqevalstep;
  std::vector<gpp_c128> cor(nfreqeval), a2(nfreqeval);
  ff_result r{};
  r.achDtemp_cor = cor.data();
  r.ach2Dtemp = a2.data();
  const int rc = run_cpp(p, FF_PART_COR, &r);
  if (rc) return rc;
  add_into(achDtemp_cor, cor);  // :241-243 add; achDtemp_corb gets zeros
  add_into(ach2Dtemp, a2);
  (void)achDtemp_corb;
  return GPP_OK;
} catch (...) {
  return gpp::exception_code(nullptr);
}

}  // extern "C"
