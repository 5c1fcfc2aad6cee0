// phc_common.hpp -- shared host-side helpers of libphc.so: thread-local error
// text, HIP call checking, device buffers and small math helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "phc.h"

namespace phc {

// A roctx range over a host-side API call (the run, the step, the tuning):
// rocprofv3 --marker-trace shows it beside the kernels it enqueued.
struct RoctxRange {
    explicit RoctxRange(const char *name) { roctxRangePushA(name); }
    ~RoctxRange() { roctxRangePop(); }
    RoctxRange(const RoctxRange &) = delete;
    RoctxRange &operator=(const RoctxRange &) = delete;
};

void set_error(const char *fmt, ...);
phc_status fail(phc_status code, const char *fmt, ...);

// Evaluate a HIP call; on failure record "<what>: <hip error>" and return
// PHC_ERR_HIP from the enclosing function.
#define PHC_HIP(call)                                                          \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess)                                                  \
            return ::phc::fail(PHC_ERR_HIP, "%s failed: %s (%s:%d)", #call,    \
                               hipGetErrorString(e_), __FILE__, __LINE__);     \
    } while (0)

#define PHC_TRY(expr)                                                          \
    do {                                                                       \
        phc_status s_ = (expr);                                                \
        if (s_ != PHC_OK) return s_;                                           \
    } while (0)

// Owning device allocation (hipMalloc): move-only, freed by release() or
// when it goes out of scope.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : ptr(o.ptr), bytes(o.bytes) { o.ptr = nullptr; o.bytes = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept
    {
        if (this != &o) {
            release();
            ptr = o.ptr; bytes = o.bytes;
            o.ptr = nullptr; o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    phc_status alloc(size_t n);
    void release();
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

inline int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

} // namespace phc
