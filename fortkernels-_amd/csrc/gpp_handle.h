// gpp_handle.h — the library handle and the host-side helpers shared by the
// GPP (gpp_api.cpp) and FF (ff_api.cpp) entry points.  Not public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/gpp_sigma.h"
#include "gpp_internal.h"
#include "gpp_stage.h"

namespace gpp {
namespace ff {
// FF host planning buffers of a handle (defined in ff_api.cpp).
struct PlanCache;
void free_plan_cache(PlanCache *c);
}  // namespace ff

// Device buffer owned by a handle, grown on demand.
struct DevBuf {
  void *ptr = nullptr;
  size_t cap = 0;
};

// Tuned tile per shape class: (precision, masked, log2 NC, log2 NG).
using TuneKey = std::tuple<int, int, int, int>;

// Bump allocator over one DevBuf, 256-byte aligned slices.
struct Arena {
  size_t used = 0;
  size_t take(size_t bytes) {
    const size_t at = used;
    used += (bytes + 255) / 256 * 256;
    return at;
  }
};

#define GPP_INTERNAL __attribute__((visibility("hidden")))

// Records the message in the handle and returns `code`.
GPP_INTERNAL int set_err(gpp_handle h, int code, const char *fmt, ...);
// Grows `b` to at least `bytes` (2 MiB granularity); synchronises first if
// it has to reallocate.
GPP_INTERNAL int ensure(gpp_handle h, DevBuf &b, size_t bytes);
// Grows the handle's pinned result buffer to `doubles` doubles.
GPP_INTERNAL int ensure_host_out(gpp_handle h, size_t doubles);
// The handle's pinned-ring stager, created and initialised on first use.
GPP_INTERNAL Stager *stager(gpp_handle h, int *rc);
// Library-owned handle on the current device for the reference-shaped
// entry points: one per device, created on first use.  The object holds a
// process-wide lock for its whole lifetime, so that concurrent drop-in
// calls from several host threads run one after another on the shared
// handle (its buffers, stream and error text).  On failure the handle's
// error text is also kept as the calling thread's handle-less error
// (gpp_last_error(NULL)).
struct DefaultHandle {
  DefaultHandle();
  ~DefaultHandle();
  DefaultHandle(const DefaultHandle &) = delete;
  DefaultHandle &operator=(const DefaultHandle &) = delete;
  int rc = 0;            // GPP_OK once h is usable
  gpp_handle h = nullptr;
  // Passes a drop-in call's return code through, recording the handle's
  // error text for gpp_last_error(NULL) when it is nonzero.
  int ret(int code);
};

// Throws std::bad_alloc when GPP_FAIL_HOST_ALLOC is set (tests of the
// entry points' exception handlers).
GPP_INTERNAL void simulated_host_alloc_failure();

// Validation shared by every entry point that takes a gpp_problem.
GPP_INTERNAL int validate(gpp_handle h, const gpp_problem *p,
                          const gpp_options *o);

}  // namespace gpp

#define HIP_OR_RET(h, expr)                                                  \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess)                                                    \
      return gpp::set_err((h), e_ == hipErrorOutOfMemory ? GPP_ENOMEM        \
                                                         : GPP_EDEVICE,      \
                          "%s: %s (%s:%d)", #expr, hipGetErrorString(e_),    \
                          __FILE__, __LINE__);                               \
  } while (0)

struct gpp_handle_s {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[6] = {};
  gpp::DevBuf stage;            // host-input staging arena
  gpp::DevBuf work;             // per-call workspace arena
  gpp::DevBuf part;             // block records
  gpp::DevBuf n1log;            // per-band records (acht_n1_loc)
  gpp::DevBuf meta;             // batch: problem table, items, record ranges
  gpp::DevBuf ff;               // FF: tables, job list, records, results
  gpp::DevBuf ff_stage;         // FF: staged I_eps slices of host inputs
  gpp::DevBuf ff_grid;          // FF: frequency-grid constants and sums
  hipStream_t ff_stream = nullptr;  // FF: frequency-grid terms (created on
                                    // first use, default priority)
  double *host_out = nullptr;   // pinned result slots
  size_t host_out_cap = 0;      // doubles
  double *host_grid = nullptr;  // FF: pinned frequency-grid results
  size_t host_grid_cap = 0;     // doubles
  std::map<gpp::TuneKey, gpp::TileCfg> tuned;
  std::string tune_file;
  // Cached graph of the last batch (replayed when the call is identical).
  hipGraphExec_t graph_exec = nullptr;
  std::vector<unsigned char> graph_sig;
  // Device-memory batches: the call (problems, options, tile) and buffers
  // the graph was captured for, and its record count, so that a repeated
  // call replays it without planning again.
  std::vector<unsigned char> batch_key;
  void *batch_bases[4] = {nullptr, nullptr, nullptr, nullptr};
  int batch_nrec = 0;
  std::string err;
  gpp::Stager *stager = nullptr;  // host-input copies (created on first use)
  gpp::ff::PlanCache *ff_plan = nullptr;  // FF host planning buffers
  long long growths = 0;          // buffer growths (failure injection count)
};
