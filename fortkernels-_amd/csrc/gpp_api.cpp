// gpp_api.cpp — the C ABI of include/gpp_sigma.h: handle lifetime, input
// validation, host->device staging (pipelined with the kernel for host
// inputs), kernel sequencing, result read-back, the batched launch
// (captured in a hipGraph) and the tile autotuner.
//
// The call sequence per evaluation replaces the reference's band loop
// (gppKer.cpp:270-361):
//   prepass (igp map + clamped frequency grid)
//   -> non-occupied kernel over bands [nvband, NB)      (gppKer.cpp:316-345)
//   -> occupied kernel over bands [0, nvband)           (gppKer.cpp:301-312)
//   -> static remainder achstemp                        (gppKer.cpp:274)
//   -> fixed-order reduction of all block records       (gppKer.cpp:365-373)
// Device memory is owned by the handle and grows on demand; nothing is
// allocated or freed between the first and the last kernel of a call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/gpp_sigma.h"
#include "gpp_guard.h"
#include "gpp_handle.h"

using gpp::BatchItem;
using gpp::DevProblem;
using gpp::TileCfg;

using gpp::Arena;
using gpp::DevBuf;
using gpp::ensure;
using gpp::ensure_host_out;
using gpp::set_err;
using gpp::stager;
using gpp::TuneKey;

namespace {

// Mixed precision: |wx - wtilde|^2 below which a term is redone in FP64.
// The FP32 path forms wx - wtilde in FP64 before rounding, so its relative
// error does not grow near resonance; the threshold only has to catch
// FP32 underflow, exact resonances and NaN.
constexpr double kDefaultResonance = 1e-6;

}  // namespace

namespace gpp {

// Error text of the calling thread's last failed handle-less call (the
// drop-in entry points and gpp_sigma_noalloc), for gpp_last_error(NULL).
thread_local std::string g_thread_err;

int set_err(gpp_handle h, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  (h ? h->err : g_thread_err) = buf;
  return code;
}

int exception_code(gpp_handle h) noexcept {
  int code = GPP_EDEVICE;
  const char *what = "internal error";
  try {
    throw;
  } catch (const std::bad_alloc &) {
    code = GPP_ENOMEM;
    what = "host memory allocation failed";
  } catch (const std::exception &e) {
    what = e.what();
  } catch (...) {
  }
  try {
    set_err(h, code, "%s", what);
  } catch (...) {
  }
  return code;
}

// Failure injection for tests: with GPP_FAIL_ALLOC_AFTER=N in the
// environment, the N-th buffer growth of a handle (device buffers and the
// pinned result buffer, counted from 1 per handle) fails as an
// out-of-memory error would, before anything is freed or allocated.
bool simulated_alloc_failure(gpp_handle h) {
  const char *v = std::getenv("GPP_FAIL_ALLOC_AFTER");
  const long long n = ++h->growths;
  return v && *v && n == std::atoll(v);
}

// Failure injection for tests: with GPP_FAIL_HOST_ALLOC set (non-empty),
// the host planning of gpp_sigma_batch and ff_sigma_ex throws
// std::bad_alloc as a failed host allocation would, which the entry
// point's handler turns into GPP_ENOMEM.
void simulated_host_alloc_failure() {
  const char *v = std::getenv("GPP_FAIL_HOST_ALLOC");
  if (v && *v) throw std::bad_alloc();
}

int ensure(gpp_handle h, DevBuf &b, size_t bytes) {
  if (bytes == 0) bytes = 256;
  if (b.cap >= bytes) return GPP_OK;
  if (simulated_alloc_failure(h))
    return set_err(h, GPP_ENOMEM,
                   "device buffer of %zu bytes: simulated allocation failure "
                   "(GPP_FAIL_ALLOC_AFTER)", bytes);
  if (b.ptr) {
    HIP_OR_RET(h, hipStreamSynchronize(h->stream));
    HIP_OR_RET(h, hipDeviceSynchronize());
    HIP_OR_RET(h, hipFree(b.ptr));
    b.ptr = nullptr;
    b.cap = 0;
  }
  // Round up to 2 MiB so repeated calls of similar size reuse the buffer.
  const size_t gran = size_t(2) << 20;
  const size_t want = (bytes + gran - 1) / gran * gran;
  HIP_OR_RET(h, hipMalloc(&b.ptr, want));
  b.cap = want;
  return GPP_OK;
}

int ensure_host_out(gpp_handle h, size_t doubles) {
  if (h->host_out_cap >= doubles) return GPP_OK;
  if (simulated_alloc_failure(h))
    return set_err(h, GPP_ENOMEM,
                   "pinned result buffer: simulated allocation failure "
                   "(GPP_FAIL_ALLOC_AFTER)");
  if (h->host_out) (void)hipHostFree(h->host_out);
  h->host_out = nullptr;
  h->host_out_cap = 0;
  HIP_OR_RET(h, hipHostMalloc((void **)&h->host_out, doubles * sizeof(double),
                              hipHostMallocDefault));
  h->host_out_cap = doubles;
  return GPP_OK;
}

Stager *stager(gpp_handle h, int *rc) {
  *rc = GPP_OK;
  if (!h->stager) h->stager = new (std::nothrow) Stager;
  if (!h->stager) {
    *rc = set_err(h, GPP_ENOMEM, "stager allocation failed");
    return nullptr;
  }
  hipError_t e = h->stager->init();
  if (e != hipSuccess) {
    *rc = set_err(h, e == hipErrorOutOfMemory ? GPP_ENOMEM : GPP_EDEVICE,
                  "pinned staging ring: %s", hipGetErrorString(e));
    return nullptr;
  }
  return h->stager;
}

namespace {
std::mutex g_default_mu;                  // held for a whole drop-in call
constexpr int kMaxDefaultDevices = 64;
gpp_handle g_default[kMaxDefaultDevices];  // one per device, never destroyed
}  // namespace

DefaultHandle::DefaultHandle() {
  g_default_mu.lock();
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  if (dev < 0 || dev >= kMaxDefaultDevices) {
    rc = set_err(nullptr, GPP_EINVAL, "device %d beyond the %d default "
                 "handles", dev, kMaxDefaultDevices);
    return;
  }
  if (!g_default[dev]) {
    rc = gpp_create(&g_default[dev], dev);
    if (rc) {
      set_err(nullptr, rc, "gpp_create on device %d failed (%d)", dev, rc);
      return;
    }
  }
  h = g_default[dev];
}

DefaultHandle::~DefaultHandle() { g_default_mu.unlock(); }

int DefaultHandle::ret(int code) {
  if (code && h) g_thread_err = h->err;
  return code;
}

int validate(gpp_handle h, const gpp_problem *p, const gpp_options *o) {
  if (!p || !o) return set_err(h, GPP_EINVAL, "null problem or options");
  if (p->number_bands < 0 || p->ncouls <= 0 || p->ngpown < 0 ||
      p->nvband < 0)
    return set_err(h, GPP_EINVAL,
                   "bad shape: number_bands=%d nvband=%d ncouls=%d ngpown=%d",
                   p->number_bands, p->nvband, p->ncouls, p->ngpown);
  if ((long long)p->number_bands * 3 > 0x7fffffffLL)
    return set_err(h, GPP_EINVAL, "number_bands too large");
  if (o->precision != GPP_PREC_FP64 && o->precision != GPP_PREC_MIXED &&
      o->precision != GPP_PREC_FP64_FACTORED)
    return set_err(h, GPP_EINVAL, "bad precision %d", o->precision);
  if (o->precision == GPP_PREC_FP64_FACTORED && p->wx_per_band &&
      p->number_bands > 0)
    return set_err(h, GPP_EUNSUPPORTED,
                   "the factored band sum needs wx shared by all bands");
  if (o->semantics < GPP_SEM_GPPKER || o->semantics > GPP_SEM_BGW)
    return set_err(h, GPP_EINVAL, "bad semantics %d", o->semantics);
  if (o->memory != GPP_MEM_HOST && o->memory != GPP_MEM_DEVICE)
    return set_err(h, GPP_EINVAL, "bad memory space %d", o->memory);
  if (o->semantics == GPP_SEM_GPPKER_STRICT && p->ngpown > p->ncouls)
    return set_err(h, GPP_EINVAL,
                   "strict semantics index I_eps[my_igp*ngpown+ig] needs "
                   "ngpown <= ncouls");
  if ((o->flags & GPP_FLAG_ALL_BANDS_UNOCC) && (o->flags & GPP_FLAG_SKIP_UNOCC))
    return set_err(h, GPP_EINVAL, "ALL_BANDS_UNOCC and SKIP_UNOCC conflict");
  const bool need_nb = p->number_bands > 0;
  const bool need_ng = p->ngpown > 0;
  if ((need_nb && (!p->aqsmtemp || !p->aqsntemp)) ||
      (need_ng && (!p->I_eps_array || !p->wtilde_array ||
                   !p->inv_igp_index)) ||
      !p->vcoul || !p->indinv || (need_nb && !p->wx_array))
    return set_err(h, GPP_EINVAL, "null input array");
  if (o->tile_ig != 0 && o->tile_ig != 4 && o->tile_ig != 8 &&
      o->tile_ig != 16 && o->tile_ig != 32)
    return set_err(h, GPP_EINVAL, "tile_ig must be 0, 4, 8, 16 or 32");
  // The kernels are compiled for at most 256 threads per block.
  if (o->waves_per_block < 0 || o->waves_per_block > 4)
    return set_err(h, GPP_EINVAL, "waves_per_block must be in [0,4]");
  return GPP_OK;
}

}  // namespace gpp

namespace {

using gpp::validate;

// The precision argument of the launchers: the public one, or the mixed
// path with compensated run sums.
int kernel_precision(const gpp_options *o) {
  return (o->precision == GPP_PREC_MIXED && (o->flags & GPP_FLAG_MIXED_KAHAN))
             ? gpp::kPrecMixedKahan
             : o->precision;
}

int ilog2_ceil(long long x) {
  int l = 0;
  while ((1LL << l) < x) ++l;
  return l;
}

TuneKey tune_key(const gpp_problem *p, const gpp_options *o) {
  return TuneKey(o->precision, o->semantics == GPP_SEM_BGW ? 1 : 0,
                 ilog2_ceil(p->ncouls), ilog2_ceil(p->ngpown > 0 ? p->ngpown : 1));
}

TileCfg default_tile(int precision) {
  TileCfg t;
  // Fastest untuned tiles at the canonical 512/2/32768/20 shape on MI355X
  // (FP64: K=8 x 4 waves, 33.5 ms; mixed: K=8 x 1 wave, 26.8 ms).
  t.k_ig = 8;
  t.waves = precision == GPP_PREC_MIXED ? 1 : 4;
  return t;
}

TileCfg pick_tile(gpp_handle h, const gpp_problem *p, const gpp_options *o) {
  TileCfg t = default_tile(o->precision);
  auto it = h->tuned.find(tune_key(p, o));
  if (it != h->tuned.end()) t = it->second;
  if (o->tile_ig) t.k_ig = o->tile_ig;
  if (o->waves_per_block) t.waves = o->waves_per_block;
  return t;
}

void load_tune_file(gpp_handle h) {
  const char *f = std::getenv("GPP_TUNE_CACHE");
  if (!f || !*f) return;
  h->tune_file = f;
  FILE *fp = std::fopen(f, "r");
  if (!fp) return;
  int a, b, c, d, k, w;
  while (std::fscanf(fp, "%d %d %d %d %d %d", &a, &b, &c, &d, &k, &w) == 6) {
    TileCfg t;
    t.k_ig = k;
    t.waves = w;
    h->tuned[TuneKey(a, b, c, d)] = t;
  }
  std::fclose(fp);
}

void save_tune_entry(gpp_handle h, const TuneKey &key, const TileCfg &t) {
  if (h->tune_file.empty()) return;
  FILE *fp = std::fopen(h->tune_file.c_str(), "a");
  if (!fp) return;
  std::fprintf(fp, "%d %d %d %d %d %d\n", std::get<0>(key), std::get<1>(key),
               std::get<2>(key), std::get<3>(key), t.k_ig, t.waves);
  std::fclose(fp);
}

// Byte sizes of the eight problem arrays, in gpp_problem order.
void array_bytes(const gpp_problem *p, size_t bytes[8]) {
  const size_t nbnc = (size_t)p->number_bands * p->ncouls;
  const size_t ngnc = (size_t)p->ngpown * p->ncouls;
  bytes[0] = nbnc * 16;
  bytes[1] = nbnc * 16;
  bytes[2] = ngnc * 16;
  bytes[3] = ngnc * 16;
  bytes[4] = (size_t)p->ncouls * 8;
  bytes[5] = (size_t)p->ngpown * 4;
  bytes[6] = (size_t)p->ncouls * 4;
  bytes[7] = (size_t)(p->wx_per_band ? p->number_bands * 3 : 3) * 8;
}

void array_ptrs(const gpp_problem *p, const void *src[8]) {
  src[0] = p->aqsmtemp;
  src[1] = p->aqsntemp;
  src[2] = p->I_eps_array;
  src[3] = p->wtilde_array;
  src[4] = p->vcoul;
  src[5] = p->inv_igp_index;
  src[6] = p->indinv;
  src[7] = p->wx_array;
}

// Everything the kernels need for one problem, in device memory.
struct Prepared {
  DevProblem dp;
  const int *inv = nullptr;
  const int *indinv = nullptr;
  const double *wx_in = nullptr;
  int wx_per_band = 0;
};

// Plans the device layout of `n` problems: staging (host inputs) in
// h->stage, igp/wx/aqsn32 tables in h->work.  Host inputs are copied
// through the pinned ring and `s` is made to wait for them, except I_eps
// and wtilde when `defer_rows` is set (the caller then copies their rows
// itself, chunk by chunk, into q.dp.ieps / q.dp.wtilde).
int prepare(gpp_handle h, const gpp_problem *p, int n, const gpp_options *o,
            bool mixed, hipStream_t s, std::vector<Prepared> &out,
            bool defer_rows = false) {
  int rc;
  out.assign(n, Prepared());
  // Staging of host inputs.
  Arena sa;
  std::vector<size_t> soff((size_t)n * 8);
  gpp::Stager *st = nullptr;
  if (o->memory == GPP_MEM_HOST) {
    for (int i = 0; i < n; ++i) {
      size_t bytes[8];
      array_bytes(p + i, bytes);
      for (int j = 0; j < 8; ++j) soff[(size_t)i * 8 + j] = sa.take(bytes[j]);
    }
    if ((rc = ensure(h, h->stage, sa.used))) return rc;
    if (!(st = stager(h, &rc))) return rc;
  }
  Arena wa;
  std::vector<size_t> woff((size_t)n * 4);
  for (int i = 0; i < n; ++i) {
    const size_t nbnc = (size_t)p[i].number_bands * p[i].ncouls;
    const size_t nbng = (size_t)p[i].number_bands * p[i].ngpown;
    woff[(size_t)i * 4 + 0] = wa.take((size_t)p[i].ngpown * sizeof(int));
    // wx [NB][3] (clamped), the occupied solver's factor [NB][3], the FP32
    // split of wx [NB][3][2]
    woff[(size_t)i * 4 + 1] =
        wa.take((size_t)p[i].number_bands * gpp::kWxDoubles * 8);
    woff[(size_t)i * 4 + 2] =
        mixed ? wa.take((size_t)p[i].number_bands *
                        gpp::padded_nc(p[i].ncouls) * sizeof(float2))
              : 0;
    woff[(size_t)i * 4 + 3] = wa.take(nbng * 16);
  }
  if ((rc = ensure(h, h->work, wa.used))) return rc;
  char *sbase = (char *)h->stage.ptr;
  char *wbase = (char *)h->work.ptr;
  for (int i = 0; i < n; ++i) {
    const void *src[8];
    array_ptrs(p + i, src);
    const void *dev[8];
    if (o->memory == GPP_MEM_HOST) {
      size_t bytes[8];
      array_bytes(p + i, bytes);
      // Smallest arrays first, then aqsm and aqsn, then the G' rows.
      static const int order[8] = {4, 5, 6, 7, 0, 1, 2, 3};
      for (int j : order) {
        void *d = sbase + soff[(size_t)i * 8 + j];
        const bool deferred = defer_rows && (j == 2 || j == 3);
        if (bytes[j] && src[j] && !deferred)
          HIP_OR_RET(h, st->copy(d, src[j], bytes[j]));
        dev[j] = d;
      }
    } else {
      for (int j = 0; j < 8; ++j) dev[j] = src[j];
    }
    Prepared &q = out[i];
    q.dp.aqsm = (const double2 *)dev[0];
    q.dp.aqsn = (const double2 *)dev[1];
    q.dp.ieps = (const double2 *)dev[2];
    q.dp.wtilde = (const double2 *)dev[3];
    q.dp.vcoul = (const double *)dev[4];
    q.inv = (const int *)dev[5];
    q.indinv = (const int *)dev[6];
    q.wx_in = (const double *)dev[7];
    q.wx_per_band = p[i].wx_per_band;
    q.dp.igp_of = (const int *)(wbase + woff[(size_t)i * 4 + 0]);
    q.dp.wx = (const double *)(wbase + woff[(size_t)i * 4 + 1]);
    q.dp.aqsn32 =
        mixed ? (const float2 *)(wbase + woff[(size_t)i * 4 + 2]) : nullptr;
    q.dp.aqsmG = (const double2 *)(wbase + woff[(size_t)i * 4 + 3]);
    q.dp.NB = p[i].number_bands;
    q.dp.NV = std::min(p[i].nvband, p[i].number_bands);
    q.dp.NC = p[i].ncouls;
    q.dp.NG = p[i].ngpown;
  }
  if (st) HIP_OR_RET(h, st->fence(s));
  return GPP_OK;
}

// Host inputs: G' groups (of 64 rows) per pipelined chunk of I_eps/wtilde.
// About a twelfth of the rows per chunk keeps the copy of the first chunk
// (which nothing overlaps) short, while each chunk's launch still has
// thousands of blocks at the canonical 512/2/32768/20 shape.
int stage_groups(int ngrp) { return std::max(1, (ngrp + 11) / 12); }

// The fallback counter lives in the first 64 bytes of h->part.
unsigned long long *counter_ptr(gpp_handle h) {
  return (unsigned long long *)h->part.ptr;
}

// Enqueue prepass (+ FP32 copy) for one prepared problem.
int enqueue_prepass(gpp_handle h, const Prepared &q, bool mixed,
                    unsigned long long *counter, hipStream_t s) {
  HIP_OR_RET(h, gpp::launch_prepass(q.inv, q.indinv, q.wx_in, q.wx_per_band,
                                    q.dp.NB, q.dp.NC, q.dp.NG,
                                    (int *)q.dp.igp_of, (double *)q.dp.wx,
                                    counter, s));
  HIP_OR_RET(h, gpp::launch_gather_aqsm(q.dp.aqsm, q.dp.igp_of, q.dp.NB,
                                        q.dp.NC, q.dp.NG,
                                        (double2 *)q.dp.aqsmG, s));
  if (mixed && q.dp.NB > q.dp.NV)
    HIP_OR_RET(h, gpp::launch_aqsn_to_fp32(q.dp.aqsn, (float2 *)q.dp.aqsn32,
                                           q.dp.NB, q.dp.NC, s));
  return GPP_OK;
}

void fill_result(const double *ho, const double *hach, gpp_result *r) {
  for (int iw = 0; iw < 3; ++iw) {
    r->achtemp[iw].re = ho[2 * iw];
    r->achtemp[iw].im = ho[2 * iw + 1];
    r->asxtemp[iw].re = ho[6 + 2 * iw];
    r->asxtemp[iw].im = ho[6 + 2 * iw + 1];
  }
  r->achstemp.re = hach[0];
  r->achstemp.im = hach[1];
}

// Replays the captured batch graph on s and returns its results.  h->part
// holds [counter 256 B][nrec_total records][n x 2 result slots], as
// gpp_sigma_batch lays it out; ev[0] has been recorded on s.
int finish_batch(gpp_handle h, const gpp_problem *p, int n,
                 const gpp_options *o, bool want_ach, int nrec_total,
                 hipStream_t s, gpp_result *r) {
  unsigned long long *dcount = counter_ptr(h);
  const double *part = (const double *)((char *)h->part.ptr + 256);
  const double *dout = part + (size_t)nrec_total * gpp::kSlots;
  HIP_OR_RET(h, hipGraphLaunch(h->graph_exec, s));
  HIP_OR_RET(h, hipEventRecord(h->ev[1], s));
  double *ho = h->host_out;
  HIP_OR_RET(h, hipMemcpyAsync(ho, dout, (size_t)2 * n * gpp::kSlots * 8,
                               hipMemcpyDeviceToHost, s));
  HIP_OR_RET(h, hipMemcpyAsync(ho + (size_t)2 * n * gpp::kSlots, dcount, 8,
                               hipMemcpyDeviceToHost, s));
  HIP_OR_RET(h, hipStreamSynchronize(s));
  float ms = 0.f;
  HIP_OR_RET(h, hipEventElapsedTime(&ms, h->ev[0], h->ev[1]));
  long long fb = 0;
  std::memcpy(&fb, ho + (size_t)2 * n * gpp::kSlots, 8);
  for (int i = 0; i < n; ++i) {
    fill_result(ho + (size_t)i * gpp::kSlots,
                ho + (size_t)(n + i) * gpp::kSlots, r + i);
    if (!want_ach) r[i].achstemp.re = r[i].achstemp.im = 0.0;
    r[i].unocc_kernel_ms = ms;
    r[i].occ_kernel_ms = 0.0;
    r[i].total_ms = ms;
    r[i].terms = (double)p[i].number_bands * p[i].ngpown *
                 (double)p[i].ncouls * 3.0;
    r[i].fp64_fallback_terms = (i == 0) ? fb : 0;
    if (o->flags & GPP_FLAG_CHECK_FINITE)
      for (int j = 0; j < 12; ++j)
        if (!std::isfinite(ho[(size_t)i * gpp::kSlots + j]))
          return set_err(h, GPP_ENONFINITE,
                         "non-finite result in problem %d slot %d", i, j);
  }
  return GPP_OK;
}

}  // namespace

extern "C" {

const char *gpp_version(void) { return "gpp-mi355x 0.2 (gfx950)"; }

void gpp_default_options(gpp_options *o) {
  if (!o) return;
  std::memset(o, 0, sizeof *o);
  o->precision = GPP_PREC_FP64;
  o->semantics = GPP_SEM_GPPKER;
  o->memory = GPP_MEM_HOST;
}

int gpp_create(gpp_handle *out, int device) try {
  if (!out) return GPP_EINVAL;
  *out = nullptr;
  gpp_handle h = new (std::nothrow) gpp_handle_s;
  if (!h) return GPP_ENOMEM;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0) {
    delete h;
    return GPP_EDEVICE;
  }
  if (device < 0) {
    if (hipGetDevice(&device) != hipSuccess) device = 0;
  }
  if (device >= ndev) {
    delete h;
    return GPP_EINVAL;
  }
  h->device = device;
  if (hipSetDevice(device) != hipSuccess ||
      hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) !=
          hipSuccess) {
    delete h;
    return GPP_EDEVICE;
  }
  for (auto &ev : h->ev) {
    if (hipEventCreate(&ev) != hipSuccess) {
      gpp_destroy(h);
      return GPP_EDEVICE;
    }
  }
  if (ensure_host_out(h, 4 * gpp::kSlots) != GPP_OK) {
    gpp_destroy(h);
    return GPP_ENOMEM;
  }
  load_tune_file(h);
  *out = h;
  return GPP_OK;
} catch (...) {
  return gpp::exception_code(nullptr);
}

int gpp_destroy(gpp_handle h) {
  if (!h) return GPP_OK;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->ff_stream) (void)hipStreamSynchronize(h->ff_stream);
  if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
  DevBuf *bufs[] = {&h->stage, &h->work,     &h->part,   &h->n1log,
                    &h->meta,  &h->ff,       &h->ff_stage, &h->ff_grid};
  for (DevBuf *b : bufs)
    if (b->ptr) (void)hipFree(b->ptr);
  for (auto &ev : h->ev)
    if (ev) (void)hipEventDestroy(ev);
  if (h->host_out) (void)hipHostFree(h->host_out);
  if (h->host_grid) (void)hipHostFree(h->host_grid);
  delete h->stager;
  gpp::ff::free_plan_cache(h->ff_plan);
  if (h->ff_stream) (void)hipStreamDestroy(h->ff_stream);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return GPP_OK;
}

const char *gpp_last_error(gpp_handle h) {
  if (!h) return gpp::g_thread_err.c_str();
  return h->err.c_str();
}

int gpp_sigma_ex(gpp_handle h, const gpp_problem *p, const gpp_options *o,
                 gpp_result *r) try {
  if (!h) return GPP_EINVAL;
  if (!r) return set_err(h, GPP_EINVAL, "null result");
  h->err.clear();
  int rc = validate(h, p, o);
  if (rc) return rc;
  HIP_OR_RET(h, hipSetDevice(h->device));
  hipStream_t s = (o->memory == GPP_MEM_DEVICE && o->stream)
                      ? (hipStream_t)o->stream
                      : h->stream;
  const bool mixed = o->precision == GPP_PREC_MIXED;
  const bool all_unocc = o->flags & GPP_FLAG_ALL_BANDS_UNOCC;
  const bool skip_unocc = o->flags & GPP_FLAG_SKIP_UNOCC;
  const bool want_ach = !(o->flags & GPP_FLAG_NO_ACHSTEMP);
  const bool want_n1 = r->acht_n1_loc != nullptr;

  const bool factored = o->precision == GPP_PREC_FP64_FACTORED;
  if ((o->flags & GPP_FLAG_AUTOTUNE) && !factored && !o->tile_ig &&
      !o->waves_per_block && !h->tuned.count(tune_key(p, o))) {
    int k, w;
    if ((rc = gpp_autotune(h, p, o, &k, &w))) return rc;
  }
  const TileCfg tile = pick_tile(h, p, o);

  const int NB = p->number_bands, NV = std::min(p->nvband, NB);
  const int n1u_begin = all_unocc ? 0 : NV;
  const int n1o_end = all_unocc ? 0 : NV;
  const bool run_unocc = !skip_unocc && n1u_begin < NB;
  const bool run_occ = n1o_end > 0;
  // Host inputs: I_eps and wtilde, most of the bytes, go over in chunks of
  // G' rows; the non-occupied launch of each chunk overlaps the copy of the
  // next.  Its blocks write the records the single launch would write, so the
  // result is bit-identical to that of device-resident inputs.
  const bool pipelined =
      o->memory == GPP_MEM_HOST && run_unocc && p->ngpown > 0;

  std::vector<Prepared> prep;
  if ((rc = prepare(h, p, 1, o, mixed, s, prep, pipelined))) return rc;
  const DevProblem &dp = prep[0].dp;
  const int n_unocc = !run_unocc ? 0
                      : factored ? gpp::unocc_gemm_grid(dp)
                                 : gpp::unocc_grid(dp, tile);
  const int n_occ = run_occ ? gpp::occ_grid(dp) : 0;
  const int n_ach = want_ach ? gpp::achstemp_records(dp, o->semantics) : 0;
  const int nrec = n_unocc + n_occ;
  // h->part: [counter 64 B][records ...][2 result slots]
  const size_t rec_bytes = (size_t)(nrec + n_ach) * gpp::kSlots * 8;
  if ((rc = ensure(h, h->part, 256 + rec_bytes + 2 * gpp::kSlots * 8)))
    return rc;
  unsigned long long *dcount = counter_ptr(h);
  double *part = (double *)((char *)h->part.ptr + 256);
  double *dout = part + (size_t)(nrec + n_ach) * gpp::kSlots;
  double *n1u = nullptr, *n1o = nullptr, *n1out = nullptr;
  // Per-band log records of the non-occupied bands: one per block of the
  // per-term kernels, one per (group, G-vector chunk) in the factored mode.
  const int n_log_u = !(run_unocc && want_n1) ? 0
                      : factored ? gpp::n1log_gemm_records(dp)
                                 : n_unocc;
  if (want_n1) {
    const size_t nu = run_unocc ? (size_t)n_log_u * (NB - n1u_begin) : 0;
    const size_t no = run_occ ? (size_t)n_occ * n1o_end : 0;
    if ((rc = ensure(h, h->n1log, (nu + no + NB + 1) * 16))) return rc;
    n1u = (double *)h->n1log.ptr;
    n1o = n1u + 2 * nu;
    n1out = n1o + 2 * no;
  }
  if ((rc = ensure_host_out(h, 2 * gpp::kSlots + 2 * (size_t)NB + 2)))
    return rc;

  HIP_OR_RET(h, hipEventRecord(h->ev[0], s));
  if ((rc = enqueue_prepass(h, prep[0], mixed, dcount, s))) return rc;
  HIP_OR_RET(h, hipEventRecord(h->ev[1], s));
  if (run_unocc) {
    const double thr =
        o->resonance_threshold > 0 ? o->resonance_threshold : kDefaultResonance;
    const int ngrp = gpp::unocc_groups(dp);
    const int step = pipelined ? stage_groups(ngrp) : ngrp;
    gpp::Stager *st = pipelined ? stager(h, &rc) : nullptr;
    if (pipelined && !st) return rc;
    for (int g = 0; g < ngrp; g += step) {
      const int cnt = std::min(step, ngrp - g);
      if (pipelined) {
        const size_t r0 = (size_t)g * 64;
        const size_t r1 = std::min((size_t)dp.NG, (size_t)(g + cnt) * 64);
        const size_t off = r0 * dp.NC, len = (r1 - r0) * dp.NC * 16;
        HIP_OR_RET(h, st->copy((void *)(dp.ieps + off), p->I_eps_array + off,
                               len));
        HIP_OR_RET(h, st->copy((void *)(dp.wtilde + off),
                               p->wtilde_array + off, len));
        HIP_OR_RET(h, st->fence(s));
      }
      if (factored) {
        HIP_OR_RET(h, gpp::launch_unocc_gemm(dp, n1u_begin, NB, o->semantics,
                                             part, s, g, cnt));
        if (want_n1)
          HIP_OR_RET(h, gpp::launch_n1log_gemm(dp, n1u_begin, NB,
                                               o->semantics, n1u, s, g, cnt));
      }
      else
        HIP_OR_RET(h, gpp::launch_unocc(dp, n1u_begin, NB, kernel_precision(o),
                                        o->semantics, tile, thr, part, n1u,
                                        mixed ? dcount : nullptr, s, g, cnt));
    }
  }
  HIP_OR_RET(h, hipEventRecord(h->ev[2], s));
  if (run_occ)
    HIP_OR_RET(h, gpp::launch_occ(dp, n1o_end, o->semantics,
                                  part + (size_t)n_unocc * gpp::kSlots, n1o,
                                  s, factored));
  HIP_OR_RET(h, hipEventRecord(h->ev[3], s));
  if (n_ach)
    HIP_OR_RET(h, gpp::launch_achstemp(dp, o->semantics,
                                       part + (size_t)nrec * gpp::kSlots, s));
  HIP_OR_RET(h, gpp::launch_reduce(part, nrec, dout, s));
  HIP_OR_RET(h, gpp::launch_reduce(part + (size_t)nrec * gpp::kSlots, n_ach,
                                   dout + gpp::kSlots, s));
  if (want_n1)
    HIP_OR_RET(h, gpp::launch_reduce_n1(n1u, run_unocc ? n_log_u : 0,
                                        n1u_begin,
                                        run_unocc ? NB - n1u_begin : 0, n1o,
                                        run_occ ? n_occ : 0, n1o_end, NB,
                                        n1out, s));
  HIP_OR_RET(h, hipEventRecord(h->ev[4], s));
  double *ho = h->host_out;
  HIP_OR_RET(h, hipMemcpyAsync(ho, dout, 2 * gpp::kSlots * sizeof(double),
                               hipMemcpyDeviceToHost, s));
  if (want_n1 && NB > 0)
    HIP_OR_RET(h, hipMemcpyAsync(ho + 2 * gpp::kSlots, n1out,
                                 (size_t)NB * 16, hipMemcpyDeviceToHost, s));
  if (mixed)
    HIP_OR_RET(h, hipMemcpyAsync(ho + 2 * gpp::kSlots + 2 * (size_t)NB, dcount,
                                 8, hipMemcpyDeviceToHost, s));
  HIP_OR_RET(h, hipStreamSynchronize(s));

  fill_result(ho, ho + gpp::kSlots, r);
  if (want_n1)
    std::memcpy(r->acht_n1_loc, ho + 2 * gpp::kSlots, (size_t)NB * 16);
  float ms = 0.f;
  HIP_OR_RET(h, hipEventElapsedTime(&ms, h->ev[1], h->ev[2]));
  r->unocc_kernel_ms = ms;
  HIP_OR_RET(h, hipEventElapsedTime(&ms, h->ev[2], h->ev[3]));
  r->occ_kernel_ms = ms;
  HIP_OR_RET(h, hipEventElapsedTime(&ms, h->ev[0], h->ev[4]));
  r->total_ms = ms;
  r->terms = (double)NB * dp.NG * (double)dp.NC * 3.0;
  long long fb = 0;
  if (mixed) std::memcpy(&fb, ho + 2 * gpp::kSlots + 2 * (size_t)NB, 8);
  r->fp64_fallback_terms = fb;
  if (o->flags & GPP_FLAG_CHECK_FINITE) {
    for (int i = 0; i < 12; ++i)
      if (!std::isfinite(ho[i]))
        return set_err(h, GPP_ENONFINITE, "non-finite result in slot %d", i);
  }
  return GPP_OK;
} catch (...) {
  return gpp::exception_code(h);
}

int gpp_autotune(gpp_handle h, const gpp_problem *p, const gpp_options *o,
                 int *tile_ig, int *waves_per_block) try {
  if (!h) return GPP_EINVAL;
  if (!tile_ig || !waves_per_block)
    return set_err(h, GPP_EINVAL, "null output");
  int rc = validate(h, p, o);
  if (rc) return rc;
  if (o->precision == GPP_PREC_FP64_FACTORED)
    return set_err(h, GPP_EUNSUPPORTED,
                   "the factored band sum has no tile to tune");
  // Time the non-occupied kernel on a slice of at most 32 unoccupied bands
  // (cost is linear in bands) for every candidate tile.
  gpp_problem q = *p;
  const int nv = std::min(p->nvband, p->number_bands);
  const int nb = std::min(p->number_bands - nv, 32);
  const TuneKey key = tune_key(p, o);
  TileCfg best = default_tile(o->precision);
  if (nb > 0 && p->ngpown > 0) {
    q.number_bands = nb;
    q.nvband = 0;
    q.aqsmtemp = p->aqsmtemp + (size_t)nv * p->ncouls;
    q.aqsntemp = p->aqsntemp + (size_t)nv * p->ncouls;
    if (p->wx_per_band) q.wx_array = p->wx_array + 3 * (size_t)nv;
    gpp_options to = *o;
    to.flags = GPP_FLAG_NO_ACHSTEMP;
    const int ks_fp64[] = {4, 8, 16};
    const int ks_mixed[] = {8, 16, 32};
    const int *ks = o->precision == GPP_PREC_MIXED ? ks_mixed : ks_fp64;
    const int ws[] = {1, 2, 4};
    double best_ms = 1e300;
    for (int i = 0; i < 3; ++i)
      for (int w : ws) {
        to.tile_ig = ks[i];
        to.waves_per_block = w;
        gpp_result r;
        std::memset(&r, 0, sizeof r);
        double t = 1e300;
        for (int rep = 0; rep < 3; ++rep) {
          if ((rc = gpp_sigma_ex(h, &q, &to, &r))) return rc;
          if (rep > 0) t = std::min(t, r.unocc_kernel_ms);
        }
        if (t < best_ms) {
          best_ms = t;
          best.k_ig = ks[i];
          best.waves = w;
        }
      }
  }
  h->tuned[key] = best;
  save_tune_entry(h, key, best);
  *tile_ig = best.k_ig;
  *waves_per_block = best.waves;
  return GPP_OK;
} catch (...) {
  return gpp::exception_code(h);
}

int gpp_sigma_batch(gpp_handle h, const gpp_problem *p, int n,
                    const gpp_options *o, gpp_result *r) try {
  if (!h) return GPP_EINVAL;
  h->err.clear();
  if (n < 0 || (n > 0 && (!p || !r)) || !o)
    return set_err(h, GPP_EINVAL, "bad batch arguments");
  if (n == 0) return GPP_OK;
  if (n > 65535)
    return set_err(h, GPP_EINVAL, "%d problems in one batch (max 65535)", n);
  int rc;
  for (int i = 0; i < n; ++i) {
    if ((rc = validate(h, p + i, o))) return rc;
    if (r[i].acht_n1_loc)
      return set_err(h, GPP_EUNSUPPORTED,
                     "acht_n1_loc is not produced by the batched launch");
  }
  if (o->flags & (GPP_FLAG_ALL_BANDS_UNOCC | GPP_FLAG_SKIP_UNOCC))
    return set_err(h, GPP_EUNSUPPORTED,
                   "band-range flags are not supported in batches");
  if (o->precision == GPP_PREC_FP64_FACTORED) {
    // Band sum first: each problem's matrix-core launch is already a full
    // grid; the problems run one after another on the same stream.
    gpp_options one = *o;
    one.flags &= ~GPP_FLAG_AUTOTUNE;
    for (int i = 0; i < n; ++i) {
      gpp_c128 *keep = r[i].acht_n1_loc;
      std::memset(r + i, 0, sizeof(gpp_result));
      r[i].acht_n1_loc = keep;  // NULL (checked above)
      if ((rc = gpp_sigma_ex(h, p + i, &one, r + i))) return rc;
    }
    return GPP_OK;
  }
  HIP_OR_RET(h, hipSetDevice(h->device));
  hipStream_t s = (o->memory == GPP_MEM_DEVICE && o->stream)
                      ? (hipStream_t)o->stream
                      : h->stream;
  const bool mixed = o->precision == GPP_PREC_MIXED;
  const bool want_ach = !(o->flags & GPP_FLAG_NO_ACHSTEMP);

  // Tile: explicit, else tuned on the largest problem, else default.
  int big = 0;
  for (int i = 1; i < n; ++i)
    if ((double)p[i].number_bands * p[i].ngpown * p[i].ncouls >
        (double)p[big].number_bands * p[big].ngpown * p[big].ncouls)
      big = i;
  if ((o->flags & GPP_FLAG_AUTOTUNE) && !o->tile_ig && !o->waves_per_block &&
      !h->tuned.count(tune_key(p + big, o))) {
    int k, w;
    if ((rc = gpp_autotune(h, p + big, o, &k, &w))) return rc;
  }
  // The occupied items always run 256-thread blocks; the non-occupied ones
  // use the tile's waves per block.
  const TileCfg tile = pick_tile(h, p + big, o);

  // Device inputs and the same call as the one the cached graph was captured
  // for (problems, options, tile, buffers): replay it as it is.  The graph
  // reads the caller's arrays on every replay, so new contents are picked
  // up; host inputs are copied on every call and always take the full path.
  std::vector<unsigned char> key;
  if (o->memory == GPP_MEM_DEVICE) {
    auto put = [&](const void *v, size_t b) {
      const unsigned char *c = (const unsigned char *)v;
      key.insert(key.end(), c, c + b);
    };
    key.reserve(sizeof(gpp_problem) * n + 128);
    put(&n, sizeof n);
    put(o, sizeof *o);
    put(&tile, sizeof tile);
    // Field by field: a caller's structs may carry uninitialised padding.
    for (int i = 0; i < n; ++i) {
      const gpp_problem &q = p[i];
      const int ints[5] = {q.number_bands, q.nvband, q.ncouls, q.ngpown,
                           q.wx_per_band};
      const void *ptrs[8] = {q.aqsmtemp, q.aqsntemp, q.I_eps_array,
                             q.wtilde_array, q.vcoul, q.inv_igp_index,
                             q.indinv, q.wx_array};
      put(ints, sizeof ints);
      put(ptrs, sizeof ptrs);
    }
  }
  void *bases_now[4] = {h->part.ptr, h->meta.ptr, h->work.ptr, (void *)s};
  if (!key.empty() && h->graph_exec && h->batch_key == key &&
      std::memcmp(bases_now, h->batch_bases, sizeof bases_now) == 0 &&
      h->host_out_cap >= (size_t)2 * n * gpp::kSlots + 1) {
    HIP_OR_RET(h, hipEventRecord(h->ev[0], s));
    return finish_batch(h, p, n, o, want_ach, h->batch_nrec, s, r);
  }
  h->batch_key.clear();

  gpp::simulated_host_alloc_failure();
  std::vector<Prepared> prep;
  if ((rc = prepare(h, p, n, o, mixed, s, prep))) return rc;

  // Work list, records and cost model (terms x per-term cost weight).
  std::vector<BatchItem> items;
  std::vector<double> cost;
  std::vector<int> rec_begin(n), rec_count(n), ach_begin(n), ach_count(n);
  int nrec = 0;
  for (int i = 0; i < n; ++i) {
    const DevProblem &dp = prep[i].dp;
    rec_begin[i] = nrec;
    const int nu = dp.NB > dp.NV ? gpp::unocc_grid(dp, tile) : 0;
    const int no = dp.NV > 0 ? gpp::occ_grid(dp) : 0;
    const double cu = (double)(dp.NB - dp.NV) * tile.k_ig * tile.waves * 64;
    const double co = (double)dp.NV * std::min(dp.NC, 2048) * 5.0;
    for (int b = 0; b < nu; ++b) {
      items.push_back(BatchItem{i, 0, b, nrec++});
      cost.push_back(cu);
    }
    for (int b = 0; b < no; ++b) {
      items.push_back(BatchItem{i, 1, b, nrec++});
      cost.push_back(co);
    }
    rec_count[i] = nu + no;
  }
  {
    std::vector<int> order(items.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
    // Non-occupied items first, then occupied ones; largest first in each.
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
      if (items[a].kind != items[b].kind) return items[a].kind < items[b].kind;
      return cost[a] > cost[b];
    });
    std::vector<BatchItem> sorted(items.size());
    for (size_t i = 0; i < order.size(); ++i) sorted[i] = items[order[i]];
    items.swap(sorted);
  }
  int n_u = 0;
  while (n_u < (int)items.size() && items[n_u].kind == 0) ++n_u;
  const int n_o = (int)items.size() - n_u;
  // The bgw static remainder runs per problem on the matrix cores (one
  // launch each, records in that kernel's layout); the gppker one, a term
  // per G' row, in one launch over the batch.
  const bool ach_gemm = o->semantics == GPP_SEM_BGW;
  int nach = 0;
  for (int i = 0; i < n; ++i) {
    ach_begin[i] = nrec + nach;
    ach_count[i] = !want_ach ? 0
                   : ach_gemm ? gpp::achstemp_records(prep[i].dp, o->semantics)
                              : gpp::achstemp_batch_records(prep[i].dp, o->semantics);
    nach += ach_count[i];
  }

  // Device metadata: problem table, prepass inputs, items, record ranges.
  std::vector<DevProblem> table(n);
  std::vector<gpp::PrepItem> preps(n);
  int max_ng_nb3 = 0, max_ach = 0;
  size_t max_nbng = 0, max_nbnc = 0;
  for (int i = 0; i < n; ++i) {
    table[i] = prep[i].dp;
    preps[i] = gpp::PrepItem{prep[i].inv, prep[i].indinv, prep[i].wx_in,
                             prep[i].wx_per_band, 0};
    const DevProblem &d = prep[i].dp;
    max_ng_nb3 = std::max(max_ng_nb3, std::max(d.NG, 3 * d.NB));
    max_nbng = std::max(max_nbng, (size_t)d.NB * d.NG);
    max_nbnc = std::max(max_nbnc, (size_t)d.NB * gpp::padded_nc(d.NC));
    max_ach = std::max(max_ach, ach_count[i]);
  }
  Arena ma;
  const size_t off_tab = ma.take(sizeof(DevProblem) * n);
  const size_t off_prep = ma.take(sizeof(gpp::PrepItem) * n);
  const size_t off_items = ma.take(sizeof(BatchItem) * items.size());
  const size_t off_rb = ma.take(sizeof(int) * n);
  const size_t off_rc = ma.take(sizeof(int) * n);
  const size_t off_ab = ma.take(sizeof(int) * n);
  const size_t off_ac = ma.take(sizeof(int) * n);
  if ((rc = ensure(h, h->meta, ma.used))) return rc;
  char *mb = (char *)h->meta.ptr;
  std::vector<unsigned char> meta_host(ma.used, 0);
  std::memcpy(meta_host.data() + off_tab, table.data(), sizeof(DevProblem) * n);
  std::memcpy(meta_host.data() + off_prep, preps.data(),
              sizeof(gpp::PrepItem) * n);
  std::memcpy(meta_host.data() + off_items, items.data(),
              sizeof(BatchItem) * items.size());
  std::memcpy(meta_host.data() + off_rb, rec_begin.data(), sizeof(int) * n);
  std::memcpy(meta_host.data() + off_rc, rec_count.data(), sizeof(int) * n);
  std::memcpy(meta_host.data() + off_ab, ach_begin.data(), sizeof(int) * n);
  std::memcpy(meta_host.data() + off_ac, ach_count.data(), sizeof(int) * n);

  // h->part: [counter 256 B][records][n x 2 result slots]
  const size_t rec_bytes = (size_t)(nrec + nach) * gpp::kSlots * 8;
  if ((rc = ensure(h, h->part,
                   256 + rec_bytes + (size_t)2 * n * gpp::kSlots * 8)))
    return rc;
  unsigned long long *dcount = counter_ptr(h);
  double *part = (double *)((char *)h->part.ptr + 256);
  double *dout = part + (size_t)(nrec + nach) * gpp::kSlots;
  double *dout_ach = dout + (size_t)n * gpp::kSlots;
  if ((rc = ensure_host_out(h, (size_t)2 * n * gpp::kSlots + 1))) return rc;

  const double thr =
      o->resonance_threshold > 0 ? o->resonance_threshold : kDefaultResonance;

  // Graph signature: everything that determines the launch sequence.
  std::vector<unsigned char> sig(meta_host);
  auto push = [&](const void *v, size_t b) {
    const unsigned char *c = (const unsigned char *)v;
    sig.insert(sig.end(), c, c + b);
  };
  push(o, sizeof *o);
  push(&tile, sizeof tile);
  void *bases[4] = {h->part.ptr, h->meta.ptr, h->work.ptr, (void *)s};
  push(bases, sizeof bases);

  HIP_OR_RET(h, hipMemcpyAsync(mb, meta_host.data(), ma.used,
                               hipMemcpyHostToDevice, s));
  HIP_OR_RET(h, hipEventRecord(h->ev[0], s));
  if (!(h->graph_exec && h->graph_sig == sig)) {
    if (h->graph_exec) {
      (void)hipGraphExecDestroy(h->graph_exec);
      h->graph_exec = nullptr;
    }
    hipGraph_t g = nullptr;
    HIP_OR_RET(h, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int crc = GPP_OK;
    {
      hipError_t e = gpp::launch_prepare_batch(
          (const DevProblem *)(mb + off_tab),
          (const gpp::PrepItem *)(mb + off_prep), n, max_ng_nb3, max_nbng,
          max_nbnc, mixed, dcount, s);
      if (e != hipSuccess)
        crc = set_err(h, GPP_EDEVICE, "batch prepass: %s", hipGetErrorString(e));
    }
    if (crc == GPP_OK) {
      const BatchItem *ditems = (const BatchItem *)(mb + off_items);
      hipError_t e = gpp::launch_batch(
          (const DevProblem *)(mb + off_tab), ditems, n_u, ditems + n_u, n_o,
          tile.k_ig, kernel_precision(o), o->semantics, thr, part, dcount,
          64 * tile.waves, (const int *)(mb + off_rb),
          (const int *)(mb + off_rc), n, dout, s);
      if (e != hipSuccess)
        crc = set_err(h, GPP_EDEVICE, "batch launch: %s", hipGetErrorString(e));
    }
    if (crc == GPP_OK && want_ach) {
      hipError_t e = hipSuccess;
      if (ach_gemm) {
        for (int i = 0; i < n && e == hipSuccess; ++i)
          if (ach_count[i] > 0)
            e = gpp::launch_achstemp(prep[i].dp, o->semantics,
                                     part + (size_t)ach_begin[i] * gpp::kSlots, s);
      } else {
        e = gpp::launch_achstemp_batch(
            (const DevProblem *)(mb + off_tab), (const int *)(mb + off_ab), n,
            max_ach, o->semantics, part, s);
      }
      if (e != hipSuccess)
        crc = set_err(h, GPP_EDEVICE, "achstemp: %s", hipGetErrorString(e));
    }
    if (crc == GPP_OK && want_ach) {
      hipError_t e = gpp::launch_reduce_batch(
          part, (const int *)(mb + off_ab), (const int *)(mb + off_ac), n,
          dout_ach, s);
      if (e != hipSuccess)
        crc = set_err(h, GPP_EDEVICE, "achstemp reduce: %s",
                      hipGetErrorString(e));
    }
    hipError_t ee = hipStreamEndCapture(s, &g);
    if (crc != GPP_OK) {
      if (g) (void)hipGraphDestroy(g);
      return crc;
    }
    HIP_OR_RET(h, ee);
    hipError_t ie = hipGraphInstantiate(&h->graph_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIP_OR_RET(h, ie);
    h->graph_sig.swap(sig);
  }
  if (!key.empty()) {
    h->batch_key.swap(key);
    std::memcpy(h->batch_bases, bases, sizeof bases);
    h->batch_nrec = nrec + nach;
  }
  return finish_batch(h, p, n, o, want_ach, nrec + nach, s, r);
} catch (...) {
  return gpp::exception_code(h);
}

int gpp_sigma(int number_bands, int nvband, int ncouls, int ngpown,
              const gpp_c128 *aqsmtemp, const gpp_c128 *aqsntemp,
              const gpp_c128 *I_eps_array, const gpp_c128 *wtilde_array,
              const double *vcoul, const int *inv_igp_index,
              const int *indinv, const double *wx_array, gpp_c128 *achtemp,
              gpp_c128 *asxtemp) try {
  if (!achtemp || !asxtemp) return GPP_EINVAL;
  gpp::DefaultHandle dh;
  if (dh.rc) return dh.rc;
  gpp_problem p = {number_bands, nvband,   ncouls,        ngpown,
                   aqsmtemp,     aqsntemp, I_eps_array,   wtilde_array,
                   vcoul,        inv_igp_index, indinv,   wx_array,
                   0};
  gpp_options o;
  gpp_default_options(&o);
  // One wx[3] for all bands: the band-sum-first FP64 evaluation applies.
  o.precision = GPP_PREC_FP64_FACTORED;
  gpp_result res;
  std::memset(&res, 0, sizeof res);
  const int rc = dh.ret(gpp_sigma_ex(dh.h, &p, &o, &res));
  if (rc) return rc;
  for (int i = 0; i < 3; ++i) {
    achtemp[i] = res.achtemp[i];
    asxtemp[i] = res.asxtemp[i];
  }
  return GPP_OK;
} catch (...) {
  return gpp::exception_code(nullptr);
}

int gpp_achtemp_kernel(int ncouls, int ngpown, const double *wx_array,
                       const gpp_c128 *wtilde_array, const gpp_c128 *aqsmtemp,
                       const gpp_c128 *aqsntemp, const gpp_c128 *I_eps_array,
                       const double *vcoul, const int *inv_igp_index,
                       const int *indinv, int number_bands,
                       double *achtemp_re0, double *achtemp_re1,
                       double *achtemp_re2, double *achtemp_im0,
                       double *achtemp_im1, double *achtemp_im2) try {
  if (!achtemp_re0 || !achtemp_re1 || !achtemp_re2 || !achtemp_im0 ||
      !achtemp_im1 || !achtemp_im2)
    return GPP_EINVAL;
  gpp::DefaultHandle dh;
  if (dh.rc) return dh.rc;
  gpp_problem p = {number_bands, 0,        ncouls,        ngpown,
                   aqsmtemp,     aqsntemp, I_eps_array,   wtilde_array,
                   vcoul,        inv_igp_index, indinv,   wx_array,
                   0};
  gpp_options o;
  gpp_default_options(&o);
  o.precision = GPP_PREC_FP64_FACTORED;  // one wx[3] for all bands
  o.flags = GPP_FLAG_ALL_BANDS_UNOCC | GPP_FLAG_NO_ACHSTEMP;
  gpp_result res;
  std::memset(&res, 0, sizeof res);
  const int rc = dh.ret(gpp_sigma_ex(dh.h, &p, &o, &res));
  if (rc) return rc;
  // The reference adds into its reduction variables (:156-157,194-199).
  *achtemp_re0 += res.achtemp[0].re;
  *achtemp_re1 += res.achtemp[1].re;
  *achtemp_re2 += res.achtemp[2].re;
  *achtemp_im0 += res.achtemp[0].im;
  *achtemp_im1 += res.achtemp[1].im;
  *achtemp_im2 += res.achtemp[2].im;
  return GPP_OK;
} catch (...) {
  return gpp::exception_code(nullptr);
}

int gpp_tillNvband_kernel(const int *inv_igp_index, const int *indinv,
                          gpp_c128 *asxtemp, const double *wx_array,
                          const gpp_c128 *wtilde_array,
                          const gpp_c128 *aqsmtemp, const gpp_c128 *aqsntemp,
                          const gpp_c128 *I_eps_array, int ncouls, int nvband,
                          int ngpown, const double *vcoul, int nstart,
                          int nend, double occ, int number_bands) try {
  if (!asxtemp || nstart < 0 || nend > 3 || nstart > nend) return GPP_EINVAL;
  gpp::DefaultHandle dh;
  if (dh.rc) return dh.rc;
  gpp_problem p = {number_bands, nvband,   ncouls,        ngpown,
                   aqsmtemp,     aqsntemp, I_eps_array,   wtilde_array,
                   vcoul,        inv_igp_index, indinv,   wx_array,
                   0};
  gpp_options o;
  gpp_default_options(&o);
  // gppKer_gpuComplex.cpp:99,119 read I_eps[my_igp*ngpown+ig].
  o.semantics = GPP_SEM_GPPKER_STRICT;
  o.flags = GPP_FLAG_SKIP_UNOCC | GPP_FLAG_NO_ACHSTEMP;
  gpp_result res;
  std::memset(&res, 0, sizeof res);
  const int rc = dh.ret(gpp_sigma_ex(dh.h, &p, &o, &res));
  if (rc) return rc;
  for (int iw = nstart; iw < nend; ++iw) {
    asxtemp[iw].re += res.asxtemp[iw].re * occ;
    asxtemp[iw].im += res.asxtemp[iw].im * occ;
  }
  return GPP_OK;
} catch (...) {
  return gpp::exception_code(nullptr);
}

}  // extern "C"
