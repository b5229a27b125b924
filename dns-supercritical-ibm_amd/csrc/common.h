// Shared host-side helpers of the dnsproj library: status codes, the
// thread-local error message and HIP call checking.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "dnsproj.h"

namespace dnsp {

void set_error(const char* fmt, ...);
void clear_error();

#define DNSP_HIP_TRY(expr)                                                          \
    do {                                                                            \
        hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess) {                                                     \
            ::dnsp::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                              __FILE__, __LINE__);                                  \
            return (e_ == hipErrorOutOfMemory) ? DNSP_ERR_NOMEM : DNSP_ERR_HIP;     \
        }                                                                           \
    } while (0)

#define DNSP_ARG_CHECK(cond, ...)                    \
    do {                                             \
        if (!(cond)) {                               \
            ::dnsp::set_error(__VA_ARGS__);          \
            return DNSP_ERR_ARG;                     \
        }                                            \
    } while (0)

// Checks the launch error of the kernels enqueued just before.
#define DNSP_LAUNCH_CHECK() DNSP_HIP_TRY(hipGetLastError())

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace dnsp
