// gpp_api.cpp — the C ABI of include/gpp_sigma.h: handle lifetime, input
// validation, host->device staging, kernel sequencing, result read-back.
//
// The call sequence per evaluation replaces the reference's band loop
// (gppKer.cpp:270-361):
//   prepass (igp map + clamped frequency grid)
//   -> non-occupied kernel over bands [nvband, NB)      (gppKer.cpp:316-345)
//   -> occupied kernel over bands [0, nvband)           (gppKer.cpp:301-312)
//   -> static remainder achstemp                        (gppKer.cpp:274)
//   -> fixed-order reduction of all block records       (gppKer.cpp:365-373)
// Device memory is owned by the handle and grows on demand; it is never
// allocated or freed inside the timed kernel sequence.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/gpp_sigma.h"
#include "gpp_internal.h"

using gpp::DevProblem;
using gpp::TileCfg;

namespace {

struct DevBuf {
  void *ptr = nullptr;
  size_t cap = 0;
};

}  // namespace

struct gpp_handle_s {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[5] = {};
  DevBuf in[8];       // staging of host inputs
  DevBuf igp_of, wx, part, out, aqsn32, count;
  double *host_out = nullptr;        // pinned, 2*kSlots doubles
  unsigned long long *host_count = nullptr;
  std::string err;
};

namespace {

int set_err(gpp_handle h, int code, const char *fmt, ...) {
  if (h) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    h->err = buf;
  }
  return code;
}

#define HIP_OR_RET(h, expr)                                                  \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess)                                                    \
      return set_err((h), e_ == hipErrorOutOfMemory ? GPP_ENOMEM             \
                                                    : GPP_EDEVICE,           \
                     "%s: %s (%s:%d)", #expr, hipGetErrorString(e_),         \
                     __FILE__, __LINE__);                                    \
  } while (0)

int ensure(gpp_handle h, DevBuf &b, size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (b.cap >= bytes) return GPP_OK;
  if (b.ptr) {
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

int validate(gpp_handle h, const gpp_problem *p, const gpp_options *o) {
  if (!p || !o) return set_err(h, GPP_EINVAL, "null problem or options");
  if (p->number_bands < 0 || p->ncouls <= 0 || p->ngpown < 0 ||
      p->nvband < 0)
    return set_err(h, GPP_EINVAL,
                   "bad shape: number_bands=%d nvband=%d ncouls=%d ngpown=%d",
                   p->number_bands, p->nvband, p->ncouls, p->ngpown);
  if ((long long)p->number_bands * 3 > 0x7fffffffLL)
    return set_err(h, GPP_EINVAL, "number_bands too large");
  if (o->precision != GPP_PREC_FP64 && o->precision != GPP_PREC_MIXED)
    return set_err(h, GPP_EINVAL, "bad precision %d", o->precision);
  if (o->semantics < GPP_SEM_GPPKER || o->semantics > GPP_SEM_BGW)
    return set_err(h, GPP_EINVAL, "bad semantics %d", o->semantics);
  if (o->memory != GPP_MEM_HOST && o->memory != GPP_MEM_DEVICE)
    return set_err(h, GPP_EINVAL, "bad memory space %d", o->memory);
  if (o->precision == GPP_PREC_MIXED && o->semantics == GPP_SEM_BGW)
    return set_err(h, GPP_EUNSUPPORTED,
                   "mixed precision supports the gppker semantics only");
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
  if (o->waves_per_block < 0 || o->waves_per_block > 16)
    return set_err(h, GPP_EINVAL, "waves_per_block must be in [0,16]");
  return GPP_OK;
}

TileCfg pick_tile(const gpp_problem *p, const gpp_options *o) {
  TileCfg t;
  t.k_ig = o->precision == GPP_PREC_MIXED ? 16 : 8;
  t.waves = 4;
  if (o->tile_ig) t.k_ig = o->tile_ig;
  if (o->waves_per_block) t.waves = o->waves_per_block;
  (void)p;
  return t;
}

}  // namespace

extern "C" {

const char *gpp_version(void) { return "gpp-mi355x 0.1 (gfx950)"; }

void gpp_default_options(gpp_options *o) {
  if (!o) return;
  std::memset(o, 0, sizeof *o);
  o->precision = GPP_PREC_FP64;
  o->semantics = GPP_SEM_GPPKER;
  o->memory = GPP_MEM_HOST;
}

int gpp_create(gpp_handle *out, int device) {
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
  if (hipHostMalloc((void **)&h->host_out, 2 * gpp::kSlots * sizeof(double),
                    hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void **)&h->host_count, sizeof(unsigned long long),
                    hipHostMallocDefault) != hipSuccess) {
    gpp_destroy(h);
    return GPP_ENOMEM;
  }
  *out = h;
  return GPP_OK;
}

int gpp_destroy(gpp_handle h) {
  if (!h) return GPP_OK;
  hipSetDevice(h->device);
  if (h->stream) hipStreamSynchronize(h->stream);
  for (auto &b : h->in)
    if (b.ptr) hipFree(b.ptr);
  DevBuf *bufs[] = {&h->igp_of, &h->wx, &h->part, &h->out, &h->aqsn32,
                    &h->count};
  for (DevBuf *b : bufs)
    if (b->ptr) hipFree(b->ptr);
  for (auto &ev : h->ev)
    if (ev) hipEventDestroy(ev);
  if (h->host_out) hipHostFree(h->host_out);
  if (h->host_count) hipHostFree(h->host_count);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
  return GPP_OK;
}

const char *gpp_last_error(gpp_handle h) {
  if (!h) return "null handle";
  return h->err.c_str();
}

int gpp_sigma_ex(gpp_handle h, const gpp_problem *p, const gpp_options *o,
                 gpp_result *r) {
  if (!h) return GPP_EINVAL;
  if (!r) return set_err(h, GPP_EINVAL, "null result");
  h->err.clear();
  int rc = validate(h, p, o);
  if (rc) return rc;
  HIP_OR_RET(h, hipSetDevice(h->device));

  const int NB = p->number_bands, NC = p->ncouls, NG = p->ngpown;
  const int NV = p->nvband < NB ? p->nvband : NB;
  const size_t nbnc = (size_t)NB * NC, ngnc = (size_t)NG * NC;
  hipStream_t s = (o->memory == GPP_MEM_DEVICE && o->stream)
                      ? (hipStream_t)o->stream
                      : h->stream;

  // Device views of the inputs.
  const void *src[8] = {p->aqsmtemp, p->aqsntemp,      p->I_eps_array,
                        p->wtilde_array, p->vcoul,     p->inv_igp_index,
                        p->indinv,     p->wx_array};
  const size_t bytes[8] = {nbnc * 16,
                           nbnc * 16,
                           ngnc * 16,
                           ngnc * 16,
                           (size_t)NC * 8,
                           (size_t)NG * 4,
                           (size_t)NC * 4,
                           (size_t)(p->wx_per_band ? NB * 3 : 3) * 8};
  const void *dev[8];
  if (o->memory == GPP_MEM_HOST) {
    for (int i = 0; i < 8; ++i) {
      if ((rc = ensure(h, h->in[i], bytes[i]))) return rc;
      if (bytes[i] && src[i])
        HIP_OR_RET(h, hipMemcpyAsync(h->in[i].ptr, src[i], bytes[i],
                                     hipMemcpyHostToDevice, s));
      dev[i] = h->in[i].ptr;
    }
  } else {
    for (int i = 0; i < 8; ++i) dev[i] = src[i];
  }

  if ((rc = ensure(h, h->igp_of, (size_t)NG * sizeof(int)))) return rc;
  if ((rc = ensure(h, h->wx, (size_t)NB * 3 * sizeof(double)))) return rc;
  if ((rc = ensure(h, h->out, 2 * gpp::kSlots * sizeof(double)))) return rc;
  if ((rc = ensure(h, h->count, sizeof(unsigned long long)))) return rc;

  DevProblem dp;
  dp.aqsm = (const double2 *)dev[0];
  dp.aqsn = (const double2 *)dev[1];
  dp.ieps = (const double2 *)dev[2];
  dp.wtilde = (const double2 *)dev[3];
  dp.vcoul = (const double *)dev[4];
  dp.igp_of = (const int *)h->igp_of.ptr;
  dp.wx = (const double *)h->wx.ptr;
  dp.aqsn32 = nullptr;
  dp.NB = NB;
  dp.NV = NV;
  dp.NC = NC;
  dp.NG = NG;

  const TileCfg tile = pick_tile(p, o);
  const bool all_unocc = o->flags & GPP_FLAG_ALL_BANDS_UNOCC;
  const bool skip_unocc = o->flags & GPP_FLAG_SKIP_UNOCC;
  const bool want_ach = !(o->flags & GPP_FLAG_NO_ACHSTEMP);
  const int n1_unocc_begin = all_unocc ? 0 : NV;
  const int n1_occ_end = all_unocc ? 0 : NV;
  const bool run_unocc = !skip_unocc && n1_unocc_begin < NB;
  const bool run_occ = n1_occ_end > 0;
  const int n_unocc = run_unocc ? gpp::unocc_grid(dp, tile) : 0;
  const int n_occ = run_occ ? gpp::occ_grid(dp) : 0;
  const int n_ach = want_ach ? gpp::achstemp_records(dp, o->semantics) : 0;
  const int nrec = n_unocc + n_occ;
  if ((rc = ensure(h, h->part,
                   (size_t)(nrec + n_ach + 1) * gpp::kSlots * sizeof(double))))
    return rc;
  double *part = (double *)h->part.ptr;
  double *dout = (double *)h->out.ptr;

  if (o->precision == GPP_PREC_MIXED && run_unocc) {
    if ((rc = ensure(h, h->aqsn32, nbnc * sizeof(float2)))) return rc;
    dp.aqsn32 = (const float2 *)h->aqsn32.ptr;
  }

  HIP_OR_RET(h, hipEventRecord(h->ev[0], s));
  HIP_OR_RET(h, gpp::launch_prepass((const int *)dev[5], (const int *)dev[6],
                                    (const double *)dev[7], p->wx_per_band,
                                    NB, NC, NG, (int *)h->igp_of.ptr,
                                    (double *)h->wx.ptr,
                                    (unsigned long long *)h->count.ptr, s));
  if (o->precision == GPP_PREC_MIXED && run_unocc)
    HIP_OR_RET(h, gpp::launch_aqsn_to_fp32(dp.aqsn,
                                           (float2 *)h->aqsn32.ptr, nbnc, s));
  unsigned long long *dcount = (unsigned long long *)h->count.ptr;
  HIP_OR_RET(h, hipEventRecord(h->ev[1], s));
  if (run_unocc) {
    const double thr =
        o->resonance_threshold > 0 ? o->resonance_threshold : 1e-2;
    HIP_OR_RET(h, gpp::launch_unocc(dp, n1_unocc_begin, NB, o->precision,
                                    o->semantics, tile, thr, part,
                                    o->precision == GPP_PREC_MIXED ? dcount
                                                                   : nullptr,
                                    s));
  }
  HIP_OR_RET(h, hipEventRecord(h->ev[2], s));
  if (run_occ)
    HIP_OR_RET(h, gpp::launch_occ(dp, n1_occ_end, o->semantics,
                                  part + (size_t)n_unocc * gpp::kSlots, s));
  HIP_OR_RET(h, hipEventRecord(h->ev[3], s));
  if (n_ach)
    HIP_OR_RET(h, gpp::launch_achstemp(dp, o->semantics,
                                       part + (size_t)nrec * gpp::kSlots, s));
  HIP_OR_RET(h, gpp::launch_reduce(part, nrec, dout, s));
  HIP_OR_RET(h, gpp::launch_reduce(part + (size_t)nrec * gpp::kSlots, n_ach,
                                   dout + gpp::kSlots, s));
  HIP_OR_RET(h, hipEventRecord(h->ev[4], s));
  HIP_OR_RET(h, hipMemcpyAsync(h->host_out, dout,
                               2 * gpp::kSlots * sizeof(double),
                               hipMemcpyDeviceToHost, s));
  if (o->precision == GPP_PREC_MIXED)
    HIP_OR_RET(h, hipMemcpyAsync(h->host_count, dcount, sizeof *dcount,
                                 hipMemcpyDeviceToHost, s));
  HIP_OR_RET(h, hipStreamSynchronize(s));

  const double *ho = h->host_out;
  for (int iw = 0; iw < 3; ++iw) {
    r->achtemp[iw].re = ho[2 * iw];
    r->achtemp[iw].im = ho[2 * iw + 1];
    r->asxtemp[iw].re = ho[6 + 2 * iw];
    r->asxtemp[iw].im = ho[6 + 2 * iw + 1];
  }
  r->achstemp.re = ho[gpp::kSlots + 0];
  r->achstemp.im = ho[gpp::kSlots + 1];
  float ms = 0.f;
  HIP_OR_RET(h, hipEventElapsedTime(&ms, h->ev[1], h->ev[2]));
  r->unocc_kernel_ms = ms;
  HIP_OR_RET(h, hipEventElapsedTime(&ms, h->ev[2], h->ev[3]));
  r->occ_kernel_ms = ms;
  HIP_OR_RET(h, hipEventElapsedTime(&ms, h->ev[0], h->ev[4]));
  r->total_ms = ms;
  r->terms = (double)NB * NG * (double)NC * 3.0;
  r->fp64_fallback_terms =
      o->precision == GPP_PREC_MIXED ? (long long)*h->host_count : 0;
  if (r->acht_n1_loc) {
    // Per-band log is not produced by the fused kernels yet.
    return set_err(h, GPP_EUNSUPPORTED, "acht_n1_loc output not supported");
  }
  if (o->flags & GPP_FLAG_CHECK_FINITE) {
    for (int i = 0; i < 12; ++i)
      if (!std::isfinite(ho[i]))
        return set_err(h, GPP_ENONFINITE, "non-finite result in slot %d", i);
  }
  return GPP_OK;
}

int gpp_sigma_batch(gpp_handle h, const gpp_problem *p, int n,
                    const gpp_options *o, gpp_result *r) {
  if (!h) return GPP_EINVAL;
  if (n < 0 || (n > 0 && (!p || !r)))
    return set_err(h, GPP_EINVAL, "bad batch arguments");
  for (int i = 0; i < n; ++i) {
    int rc = gpp_sigma_ex(h, p + i, o, r + i);
    if (rc) return rc;
  }
  return GPP_OK;
}

// Library-owned handle for the reference-shaped entry points.
static gpp_handle default_handle(int *rc) {
  static std::mutex mu;
  static gpp_handle h = nullptr;
  std::lock_guard<std::mutex> lock(mu);
  *rc = GPP_OK;
  if (!h) *rc = gpp_create(&h, -1);
  return h;
}

int gpp_sigma(int number_bands, int nvband, int ncouls, int ngpown,
              const gpp_c128 *aqsmtemp, const gpp_c128 *aqsntemp,
              const gpp_c128 *I_eps_array, const gpp_c128 *wtilde_array,
              const double *vcoul, const int *inv_igp_index,
              const int *indinv, const double *wx_array, gpp_c128 *achtemp,
              gpp_c128 *asxtemp) {
  if (!achtemp || !asxtemp) return GPP_EINVAL;
  int rc;
  gpp_handle h = default_handle(&rc);
  if (rc) return rc;
  gpp_problem p = {number_bands, nvband,   ncouls,        ngpown,
                   aqsmtemp,     aqsntemp, I_eps_array,   wtilde_array,
                   vcoul,        inv_igp_index, indinv,   wx_array,
                   0};
  gpp_options o;
  gpp_default_options(&o);
  gpp_result res;
  std::memset(&res, 0, sizeof res);
  rc = gpp_sigma_ex(h, &p, &o, &res);
  if (rc) return rc;
  for (int i = 0; i < 3; ++i) {
    achtemp[i] = res.achtemp[i];
    asxtemp[i] = res.asxtemp[i];
  }
  return GPP_OK;
}

int gpp_achtemp_kernel(int ncouls, int ngpown, const double *wx_array,
                       const gpp_c128 *wtilde_array, const gpp_c128 *aqsmtemp,
                       const gpp_c128 *aqsntemp, const gpp_c128 *I_eps_array,
                       const double *vcoul, const int *inv_igp_index,
                       const int *indinv, int number_bands,
                       double *achtemp_re0, double *achtemp_re1,
                       double *achtemp_re2, double *achtemp_im0,
                       double *achtemp_im1, double *achtemp_im2) {
  if (!achtemp_re0 || !achtemp_re1 || !achtemp_re2 || !achtemp_im0 ||
      !achtemp_im1 || !achtemp_im2)
    return GPP_EINVAL;
  int rc;
  gpp_handle h = default_handle(&rc);
  if (rc) return rc;
  gpp_problem p = {number_bands, 0,        ncouls,        ngpown,
                   aqsmtemp,     aqsntemp, I_eps_array,   wtilde_array,
                   vcoul,        inv_igp_index, indinv,   wx_array,
                   0};
  gpp_options o;
  gpp_default_options(&o);
  o.flags = GPP_FLAG_ALL_BANDS_UNOCC | GPP_FLAG_NO_ACHSTEMP;
  gpp_result res;
  std::memset(&res, 0, sizeof res);
  rc = gpp_sigma_ex(h, &p, &o, &res);
  if (rc) return rc;
  // The reference adds into its reduction variables (:156-157,194-199).
  *achtemp_re0 += res.achtemp[0].re;
  *achtemp_re1 += res.achtemp[1].re;
  *achtemp_re2 += res.achtemp[2].re;
  *achtemp_im0 += res.achtemp[0].im;
  *achtemp_im1 += res.achtemp[1].im;
  *achtemp_im2 += res.achtemp[2].im;
  return GPP_OK;
}

int gpp_tillNvband_kernel(const int *inv_igp_index, const int *indinv,
                          gpp_c128 *asxtemp, const double *wx_array,
                          const gpp_c128 *wtilde_array,
                          const gpp_c128 *aqsmtemp, const gpp_c128 *aqsntemp,
                          const gpp_c128 *I_eps_array, int ncouls, int nvband,
                          int ngpown, const double *vcoul, int nstart,
                          int nend, double occ, int number_bands) {
  if (!asxtemp || nstart < 0 || nend > 3 || nstart > nend) return GPP_EINVAL;
  int rc;
  gpp_handle h = default_handle(&rc);
  if (rc) return rc;
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
  rc = gpp_sigma_ex(h, &p, &o, &res);
  if (rc) return rc;
  for (int iw = nstart; iw < nend; ++iw) {
    asxtemp[iw].re += res.asxtemp[iw].re * occ;
    asxtemp[iw].im += res.asxtemp[iw].im * occ;
  }
  return GPP_OK;
}

}  // extern "C"
