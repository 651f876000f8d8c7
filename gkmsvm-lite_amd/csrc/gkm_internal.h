// gkm_internal.h -- what the engine's HIP translation units share: the
// thread-local error message behind gkm_last_error(), the launch check and
// the export macro of the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#define GKM_API extern "C" __attribute__((visibility("default")))

// Record an error message (printf format) for gkm_last_error(); returns code.
__attribute__((visibility("hidden"))) int fail(int code, const char *fmt, ...);

static inline hipStream_t S_(void *stream) { return (hipStream_t)stream; }

#define GKM_CHECK_LAUNCH(what)                                                  \
    do {                                                                        \
        hipError_t e_ = hipGetLastError();                                      \
        if (e_ != hipSuccess)                                                   \
            return fail(-3, "%s: launch failed: %s", what, hipGetErrorString(e_)); \
    } while (0)
