// gpp_guard.h — what the C entry points return when an exception (a failed
// host allocation, a thread that cannot start) reaches them: each exported
// function body is a function-try-block whose handler returns
// exception_code(h), so no exception crosses the C ABI.  Not public ABI.
#pragma once

#include "../../include/gpp_sigma.h"

namespace gpp {

// From inside a catch block: GPP_ENOMEM for std::bad_alloc, GPP_EDEVICE for
// anything else; the message goes to the handle's error text, or to the
// calling thread's (gpp_last_error(NULL)) when h is null, if there is
// memory left for it.
__attribute__((visibility("hidden"))) int exception_code(gpp_handle h) noexcept;

}  // namespace gpp
