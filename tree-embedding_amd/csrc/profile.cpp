// profile.cpp -- per-kernel HIP-event timing registry (see tsvd_profile.h).
#include "tsvd_gemm.h"

#include <cstring>
#include "tsvd_profile.h"
#include "tsvd_internal.h"

#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

namespace tsvd {

bool g_profile = false;

namespace {

struct Rec {
    std::string name;
    hipEvent_t a, b;
};

struct Sum {
    int64_t launches = 0;
    double ms = 0.0, bytes = 0.0, flops = 0.0;
};

std::mutex g_mu;
std::vector<Rec> g_pending;
std::vector<hipEvent_t> g_pool;
std::map<std::string, Sum> g_sum;

hipEvent_t take_event()
{
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

void drain_locked()
{
    for (auto &r : g_pending) {
        (void)hipEventSynchronize(r.b);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, r.a, r.b);
        Sum &s = g_sum[r.name];
        s.launches += 1;
        s.ms += ms;
        g_pool.push_back(r.a);
        g_pool.push_back(r.b);
    }
    g_pending.clear();
}

}  // namespace

ProfScope::ProfScope(const char *n, hipStream_t s) : name(n), st(s)
{
    if (!g_profile) return;
    std::lock_guard<std::mutex> lk(g_mu);
    a = take_event();
    (void)hipEventRecord(a, st);
}

ProfScope::~ProfScope()
{
    if (!a) return;
    std::lock_guard<std::mutex> lk(g_mu);
    hipEvent_t b = take_event();
    (void)hipEventRecord(b, st);
    g_pending.push_back(Rec{name, a, b});
    if (g_pending.size() > 4096) drain_locked();
}

void prof_add_work(const char *name, double bytes, double flops)
{
    if (!g_profile) return;
    std::lock_guard<std::mutex> lk(g_mu);
    Sum &s = g_sum[name];
    s.bytes += bytes;
    s.flops += flops;
}

void prof_add_time(const char *name, double ms)
{
    if (!g_profile) return;
    std::lock_guard<std::mutex> lk(g_mu);
    Sum &s = g_sum[name];
    s.launches += 1;
    s.ms += ms;
}

}  // namespace tsvd

extern "C" int tsvd_profile_enable(int on)
{
    return tsvd::guard([&] { tsvd::g_profile = on != 0; });
}

extern "C" int tsvd_gemm_tuning(int on)
{
    return tsvd::guard([&] { tsvd::gemm_tuning_enable(on != 0); });
}

extern "C" int tsvd_gemm_tuning_read(char *json, int len)
{
    return tsvd::guard([&] {
        TSVD_REQUIRE(json && len > 2, "null or short buffer");
        const std::string s = tsvd::gemm_tuning_json();
        TSVD_REQUIRE((int)s.size() < len, "buffer too short for the tuning record (%d bytes)", (int)s.size() + 1);
        memcpy(json, s.c_str(), s.size() + 1);
    });
}

extern "C" int tsvd_profile_reset(void)
{
    return tsvd::guard([&] {
        std::lock_guard<std::mutex> lk(tsvd::g_mu);
        tsvd::drain_locked();
        tsvd::g_sum.clear();
    });
}

// JSON object {"name": [launches, total_ms, algorithmic_bytes, flops], ...}
extern "C" int tsvd_profile_read(char *json, int len)
{
    return tsvd::guard([&] {
        TSVD_REQUIRE(json && len > 2, "null or short buffer");
        std::lock_guard<std::mutex> lk(tsvd::g_mu);
        tsvd::drain_locked();
        std::string s = "{";
        bool first = true;
        for (auto &kv : tsvd::g_sum) {
            char buf[320];
            snprintf(buf, sizeof(buf), "%s\"%s\": [%lld, %.6f, %.6e, %.6e]", first ? "" : ", ", kv.first.c_str(),
                     (long long)kv.second.launches, kv.second.ms, kv.second.bytes, kv.second.flops);
            s += buf;
            first = false;
        }
        s += "}";
        TSVD_REQUIRE((int)s.size() < len, "buffer too small for the profile (%zu bytes)", s.size());
        memcpy(json, s.c_str(), s.size() + 1);
    });
}
