// gpp_noalloc.hip — gpp_sigma_noalloc: one GPP sigma evaluation of
// device-resident inputs that allocates nothing.
//
// The evaluation runs the same kernels, in the same order and with the same
// record layout, as gpp_sigma_ex with GPP_MEM_DEVICE (FP64 per term or band
// sum first), so its results are bit-identical to that entry point's for the
// same tile.  Its scratch memory (igp table, clamped frequencies, gathered
// aqsm rows, block records, result slots) is carved out of one file-scope
// device array instead of handle buffers, and the results go to a device
// buffer the caller passes in: the call only enqueues kernels and one
// device-to-device copy on the caller's stream and returns, so it can sit
// inside a captured HIP graph or a loop that must not allocate.
//
// The array exists once per GPU (the code object's own copy) and is shared
// by every call on that GPU: calls that overlap on one GPU (two streams, two
// host threads) race on it and must be serialised by the caller.  Nothing in
// it needs to start at zero: every kernel of the sequence writes each record
// it owns in full, and the prepass clears its counter with a kernel store.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <mutex>

#include "gpp_guard.h"
#include "../../include/gpp_sigma.h"
#include "gpp_handle.h"

namespace {

constexpr size_t kScratchBytes = GPP_NOALLOC_SCRATCH_BYTES;
static_assert(kScratchBytes % 256 == 0, "scratch size is 256-byte granular");

// The scratch memory.  In host code this name is only a handle for
// hipGetSymbolAddress; kernels get the device address resolved below.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::once_flag g_once[kMaxDevices];
void *g_ptr[kMaxDevices];               // device address of scratch per GPU
hipError_t g_err[kMaxDevices];

// Resolves the address of this GPU's copy of `scratch` once per device
// (with that device current, as the caller has it).
int scratch_of(int dev, unsigned char **out) {
  if (dev < 0 || dev >= kMaxDevices)
    return gpp::set_err(nullptr, GPP_EINVAL, "device %d beyond %d devices",
                        dev, kMaxDevices);
  std::call_once(g_once[dev], [dev] {
    void *p = nullptr;
    g_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_ptr[dev] = (g_err[dev] == hipSuccess) ? p : nullptr;
  });
  if (g_err[dev] != hipSuccess || !g_ptr[dev])
    return gpp::set_err(nullptr, GPP_EDEVICE,
                        "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                        hipGetErrorString(g_err[dev]));
  *out = (unsigned char *)g_ptr[dev];
  return GPP_OK;
}

#define NOALLOC_HIP(expr)                                                    \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess)                                                    \
      return gpp::set_err(nullptr, GPP_EDEVICE, "%s: %s", #expr,             \
                          hipGetErrorString(e_));                            \
  } while (0)

}  // namespace

extern "C" int gpp_sigma_noalloc(const gpp_problem *p, const gpp_options *o,
                                 gpp_c128 *out) try {
  int rc = gpp::validate(nullptr, p, o);
  if (rc) return rc;
  if (!out) return gpp::set_err(nullptr, GPP_EINVAL, "null output buffer");
  if (o->memory != GPP_MEM_DEVICE)
    return gpp::set_err(nullptr, GPP_EINVAL,
                        "gpp_sigma_noalloc takes device-resident inputs only "
                        "(memory = GPP_MEM_DEVICE)");
  if (o->precision == GPP_PREC_MIXED)
    return gpp::set_err(nullptr, GPP_EUNSUPPORTED,
                        "gpp_sigma_noalloc runs FP64 (per term or band sum "
                        "first); the mixed path's FP32 copy of aqsn does not "
                        "fit its scratch");
  if (o->flags & (GPP_FLAG_AUTOTUNE | GPP_FLAG_CHECK_FINITE))
    return gpp::set_err(nullptr, GPP_EUNSUPPORTED,
                        "gpp_sigma_noalloc does not tune or read results back "
                        "(AUTOTUNE, CHECK_FINITE)");
  const bool factored = o->precision == GPP_PREC_FP64_FACTORED;
  const bool all_unocc = o->flags & GPP_FLAG_ALL_BANDS_UNOCC;
  const bool skip_unocc = o->flags & GPP_FLAG_SKIP_UNOCC;
  const bool want_ach = !(o->flags & GPP_FLAG_NO_ACHSTEMP);
  gpp::TileCfg tile;
  tile.k_ig = o->tile_ig ? o->tile_ig : 8;
  tile.waves = o->waves_per_block ? o->waves_per_block : 4;

  const int NB = p->number_bands, NV = std::min(p->nvband, NB);
  const int NC = p->ncouls, NG = p->ngpown;
  const int n1u_begin = all_unocc ? 0 : NV;
  const int n1o_end = all_unocc ? 0 : NV;
  const bool run_unocc = !skip_unocc && n1u_begin < NB;
  const bool run_occ = n1o_end > 0;

  // Layout of the scratch array, as gpp_sigma_ex lays out its handle
  // buffers: [counter][igp_of NG][wx NB*9][aqsmG NB*NG][records][2 slots].
  gpp::Arena a;
  const size_t off_cnt = a.take(256);
  const size_t off_igp = a.take((size_t)NG * sizeof(int));
  const size_t off_wx = a.take((size_t)NB * gpp::kWxDoubles * sizeof(double));
  const size_t off_aqsmG = a.take((size_t)NB * NG * 16);
  gpp::DevProblem dp{};
  dp.aqsm = (const double2 *)p->aqsmtemp;
  dp.aqsn = (const double2 *)p->aqsntemp;
  dp.ieps = (const double2 *)p->I_eps_array;
  dp.wtilde = (const double2 *)p->wtilde_array;
  dp.vcoul = p->vcoul;
  dp.aqsn32 = nullptr;
  dp.NB = NB;
  dp.NV = NV;
  dp.NC = NC;
  dp.NG = NG;
  const int n_unocc = !run_unocc ? 0
                      : factored ? gpp::unocc_gemm_grid(dp)
                                 : gpp::unocc_grid(dp, tile);
  const int n_occ = run_occ ? gpp::occ_grid(dp) : 0;
  const int n_ach = want_ach ? gpp::achstemp_records(dp, o->semantics) : 0;
  const int nrec = n_unocc + n_occ;
  const size_t off_rec = a.take((size_t)(nrec + n_ach) * gpp::kSlots * 8);
  const size_t off_out = a.take(2 * gpp::kSlots * 8);
  if (a.used > kScratchBytes)
    return gpp::set_err(nullptr, GPP_ESCRATCH,
                        "gpp_sigma_noalloc: this problem (number_bands=%d "
                        "ncouls=%d ngpown=%d) needs %zu bytes of scratch; the "
                        "allocation-free entry point holds %zu",
                        NB, NC, NG, a.used, kScratchBytes);
  // Nothing above touched the device: bad arguments and oversized problems
  // are refused before the first HIP call.
  int dev = 0;
  NOALLOC_HIP(hipGetDevice(&dev));
  unsigned char *base = nullptr;
  if ((rc = scratch_of(dev, &base))) return rc;
  hipStream_t s = (hipStream_t)o->stream;

  unsigned long long *dcount = (unsigned long long *)(base + off_cnt);
  dp.igp_of = (const int *)(base + off_igp);
  dp.wx = (const double *)(base + off_wx);
  dp.aqsmG = (const double2 *)(base + off_aqsmG);
  double *part = (double *)(base + off_rec);
  double *dres = (double *)(base + off_out);

  NOALLOC_HIP(gpp::launch_prepass(p->inv_igp_index, p->indinv, p->wx_array,
                                  p->wx_per_band, NB, NC, NG,
                                  (int *)dp.igp_of, (double *)dp.wx, dcount,
                                  s));
  NOALLOC_HIP(gpp::launch_gather_aqsm(dp.aqsm, dp.igp_of, NB, NC, NG,
                                      (double2 *)dp.aqsmG, s));
  if (run_unocc) {
    if (factored)
      NOALLOC_HIP(gpp::launch_unocc_gemm(dp, n1u_begin, NB, o->semantics,
                                         part, s));
    else
      NOALLOC_HIP(gpp::launch_unocc(dp, n1u_begin, NB, GPP_PREC_FP64,
                                    o->semantics, tile, 0.0, part, nullptr,
                                    nullptr, s));
  }
  if (run_occ)
    NOALLOC_HIP(gpp::launch_occ(dp, n1o_end, o->semantics,
                                part + (size_t)n_unocc * gpp::kSlots, nullptr,
                                s, factored));
  if (n_ach)
    NOALLOC_HIP(gpp::launch_achstemp(dp, o->semantics,
                                     part + (size_t)nrec * gpp::kSlots, s));
  // achtemp and asxtemp straight into out[0..5]; achstemp via the slots.
  NOALLOC_HIP(gpp::launch_reduce(part, nrec, (double *)out, s));
  NOALLOC_HIP(gpp::launch_reduce(part + (size_t)nrec * gpp::kSlots, n_ach,
                                 dres, s));
  NOALLOC_HIP(hipMemcpyAsync(out + 6, dres, sizeof(gpp_c128),
                             hipMemcpyDeviceToDevice, s));
  return GPP_OK;
} catch (...) {
  return gpp::exception_code(nullptr);
}
