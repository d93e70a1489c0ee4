// Definitions shared by the translation units of libbolt_hip.so (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/bolt.h"

// Records a per-thread error message and returns code.
int bolt_fail(int code, const char* fmt, ...);
const char* bolt_error_text();
#define fail bolt_fail

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return bolt_fail(BOLT_EDEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                             __FILE__, __LINE__);                                          \
    } while (0)

struct bolt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string info;
};
