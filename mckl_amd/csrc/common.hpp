// common.hpp -- status codes, error reporting and small host helpers shared
// by the HIP translation units of libmckl_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/mckl_hip.h"

namespace mckl_hip {

std::string &last_error_ref();

inline int set_error(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    last_error_ref() = buf;
    return code;
}

inline int hip_status(hipError_t err, const char *what)
{
    if (err == hipSuccess)
        return MCKL_HIP_OK;
    return set_error(MCKL_HIP_EHIP - (int)err, "%s: %s", what,
        hipGetErrorString(err));
}

#define MCKL_HIP_TRY(expr)                                                    \
    do {                                                                      \
        int mckl_hip_rc_ = hip_status((expr), #expr);                         \
        if (mckl_hip_rc_ != MCKL_HIP_OK)                                      \
            return mckl_hip_rc_;                                              \
    } while (0)

#define MCKL_HIP_CHECK_LAUNCH(what)                                           \
    do {                                                                      \
        int mckl_hip_rc_ = hip_status(hipGetLastError(), what);               \
        if (mckl_hip_rc_ != MCKL_HIP_OK)                                      \
            return mckl_hip_rc_;                                              \
    } while (0)

inline hipStream_t as_stream(void *s) { return static_cast<hipStream_t>(s); }

constexpr uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// Makes sure per-device constant tables are uploaded (AES T0).  Returns a
// status code.
int ensure_device_tables();

} // namespace mckl_hip
