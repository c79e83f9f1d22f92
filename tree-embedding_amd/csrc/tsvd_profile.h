// tsvd_profile.h -- optional per-kernel timing with HIP events.
//
// When enabled (tsvd_profile_enable), every library launch site wrapped in
// PROF(name, stream) records a HIP event pair on the stream the kernel runs
// on; tsvd_profile_read sums the elapsed times per name.  Disabled, a scope
// costs one branch.
#pragma once
#include <hip/hip_runtime.h>

namespace tsvd {

extern bool g_profile;

struct ProfScope {
    const char *name;
    hipStream_t st;
    hipEvent_t a = nullptr;
    ProfScope(const char *n, hipStream_t s);
    ~ProfScope();
};

// Algorithmic work (bytes, flops) credited to a kernel name; only counted
// while profiling is enabled.
void prof_add_work(const char *name, double bytes, double flops);
// One launch of `name` timed elsewhere (kernels replayed from a captured
// graph, timed by event nodes of that graph).
void prof_add_time(const char *name, double ms);

}  // namespace tsvd

#define TSVD_PROF_CAT2(a, b) a##b
#define TSVD_PROF_CAT(a, b) TSVD_PROF_CAT2(a, b)
#define PROF(name, st) ::tsvd::ProfScope TSVD_PROF_CAT(prof_scope_, __LINE__)(name, st)
