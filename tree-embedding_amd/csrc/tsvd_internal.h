// tsvd_internal.h -- shared host/device helpers of the Tree-SVD HIP library.
#pragma once

#include <hip/hip_runtime.h>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdarg>
#include <cstring>
#include <string>
#include <vector>
#include <new>

#define TREESVD_BUILDING 1
#include "../../include/treesvd.h"
#include "tsvd_profile.h"

namespace tsvd {

// ----------------------------------------------------------------- errors
int set_error(int code, const char *fmt, ...);

struct Error {
    int code;
};

#define TSVD_HIP(call)                                                                          \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            throw ::tsvd::Error{::tsvd::set_error(TSVD_EHIP, "%s failed: %s (%s:%d)", #call,     \
                                                  hipGetErrorString(e_), __FILE__, __LINE__)};  \
    } while (0)

#define TSVD_REQUIRE(cond, ...)                                                       \
    do {                                                                              \
        if (!(cond)) throw ::tsvd::Error{::tsvd::set_error(TSVD_EINVAL, __VA_ARGS__)}; \
    } while (0)

// Run `body` and turn exceptions into status codes (no exception crosses the ABI).
template <class F>
int guard(F &&body)
{
    try {
        body();
        return TSVD_OK;
    } catch (const Error &e) {
        return e.code;
    } catch (const std::bad_alloc &) {
        return set_error(TSVD_ENOMEM, "host allocation failed");
    } catch (...) {
        return set_error(TSVD_EHIP, "unexpected exception");
    }
}

// ------------------------------------------------------------ device memory
inline double wall_ms_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

template <class T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf &operator=(DBuf &&o) noexcept
    {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DBuf() { release(); }
    // hipMalloc and hipFree wall times are profiled as "hip_malloc" and
    // "hip_free" (a large allocation can wait for the driver to clear freed
    // memory)
    void alloc(size_t count)
    {
        release();
        if (count == 0) return;
        const auto t0 = std::chrono::steady_clock::now();
        hipError_t e = hipMalloc((void **)&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            throw Error{set_error(TSVD_ENOMEM, "hipMalloc of %zu bytes failed: %s", count * sizeof(T),
                                  hipGetErrorString(e))};
        }
        n = count;
        if (g_profile) prof_add_time("hip_malloc", wall_ms_since(t0));
    }
    // grow (contents discarded) when the request exceeds the current size
    void ensure(size_t count)
    {
        if (count > n) alloc(count);
    }
    // grow (contents discarded): the first allocation is exact, a regrowth
    // takes 1/8 headroom -- for workspaces whose need creeps up from step to
    // step (with the graph's edge count, the batch size, the dirty set), so
    // that steps rarely reallocate them
    void grow(size_t count)
    {
        if (count > n) alloc(n == 0 ? count : count + count / 8);
    }
    void release()
    {
        if (p) {
            const auto t0 = std::chrono::steady_clock::now();
            (void)hipFree(p);
            if (g_profile) prof_add_time("hip_free", wall_ms_since(t0));
        }
        p = nullptr;
        n = 0;
    }
    size_t bytes() const { return n * sizeof(T); }
};

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// A pointer read from a descriptor in memory is generic to the compiler, so
// its accesses become flat instructions, which also count against the LDS
// wait counter: a flat load in flight then stalls every LDS wait after it.
// Device buffers are global memory; these casts make the accesses global.
template <class T>
using gptr = T __attribute__((address_space(1))) *;
template <class T>
__device__ inline gptr<const T> gbl(const T *p)
{
    return (gptr<const T>)p;
}
template <class T>
__device__ inline gptr<T> gbl(T *p)
{
    return (gptr<T>)p;
}
inline int64_t round_up(int64_t a, int64_t b) { return ceil_div(a, b) * b; }

// Event-bracketed timing on one stream.
struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    Timer()
    {
        TSVD_HIP(hipEventCreate(&a));
        TSVD_HIP(hipEventCreate(&b));
    }
    ~Timer()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
    void start(hipStream_t s) { TSVD_HIP(hipEventRecord(a, s)); }
    void stop(hipStream_t s) { TSVD_HIP(hipEventRecord(b, s)); }
    float ms()
    {
        TSVD_HIP(hipEventSynchronize(b));
        float t = 0.f;
        TSVD_HIP(hipEventElapsedTime(&t, a, b));
        return t;
    }
};

}  // namespace tsvd
