// Error reporting shared by the library's host code: a failing call records a message for
// pxfdtd_last_error() and returns its status.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/pxfdtd.h"

int px_fail(int code, const char *fmt, ...);

#define HIPCHK(x)                                                                        \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess)                                                            \
            return px_fail(e_ == hipErrorOutOfMemory ? PXFDTD_ERR_MEMORY : PXFDTD_ERR_DEVICE, \
                           "%s: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

#define TRY(x)                       \
    do {                             \
        int rc_ = (x);               \
        if (rc_ != PXFDTD_OK) return rc_; \
    } while (0)
