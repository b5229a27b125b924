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

// Device bytes a handle holds: every allocation it owns at the moment
// (poisson.hip, ibm.hip).
size_t poisson_device_bytes(const dnsp_poisson* h);
size_t ibm_device_bytes(const dnsp_ibm* h);

// A rough-wall column-slab pair (bottom, top; n ints each) on the device:
// copies the host slabs in (allocating on first use, *bytes grows by the
// allocation when given), or frees the pair when both are NULL.
inline int set_slab_pair(int** d_bot, int** d_top, const int* bot, const int* top, size_t n, size_t* bytes) {
    DNSP_ARG_CHECK((bot == nullptr) == (top == nullptr), "give both column slabs or neither");
    if (!bot) {
        if (*d_bot && bytes) *bytes -= 2 * sizeof(int) * n;
        (void)hipFree(*d_bot);
        (void)hipFree(*d_top);
        *d_bot = *d_top = nullptr;
        return DNSP_OK;
    }
    if (!*d_bot) {
        DNSP_HIP_TRY(hipMalloc(d_bot, sizeof(int) * n));
        DNSP_HIP_TRY(hipMalloc(d_top, sizeof(int) * n));
        if (bytes) *bytes += 2 * sizeof(int) * n;
    }
    DNSP_HIP_TRY(hipMemcpy(*d_bot, bot, sizeof(int) * n, hipMemcpyHostToDevice));
    DNSP_HIP_TRY(hipMemcpy(*d_top, top, sizeof(int) * n, hipMemcpyHostToDevice));
    return DNSP_OK;
}

}  // namespace dnsp
