// tsvd_internal.h -- shared host/device helpers of the Tree-SVD HIP library.
#pragma once

#include <hip/hip_runtime.h>
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
    void alloc(size_t count)
    {
        release();
        if (count == 0) return;
        hipError_t e = hipMalloc((void **)&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            throw Error{set_error(TSVD_ENOMEM, "hipMalloc of %zu bytes failed: %s", count * sizeof(T),
                                  hipGetErrorString(e))};
        }
        n = count;
    }
    // grow (contents discarded) when the request exceeds the current size
    void ensure(size_t count)
    {
        if (count > n) alloc(count);
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    size_t bytes() const { return n * sizeof(T); }
};

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
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
