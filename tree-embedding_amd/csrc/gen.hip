// gen.hip -- seeded synthetic R-MAT and Chung-Lu edges generated in HBM.
//
// No reference counterpart.  The measured configurations need seeded
// synthetic graphs; at 1.8e8-1.47e9 edges drawing them on the host takes
// longer than the run itself, so they are drawn here.  Each edge e is drawn
// from a counter-based hash of (seed, round, e):
//  * R-MAT: at each of `scale` levels a quadrant with probabilities a, b, c,
//    1-a-b-c (16 hash bits per level), most significant bit first; ids are
//    taken modulo n;
//  * Chung-Lu: u with probability proportional to (u+1)^(-1/(out_exp-1)),
//    v with probability proportional to (pi(v)+1)^(-1/(in_exp-1)) for a
//    seeded random permutation pi, each by a binary search of a 53-bit
//    uniform in the cumulative weights;
// self-loops are dropped.  The distinct keys (u * n + v) are kept by a radix
// sort and a unique pass, topped up by further rounds of draws until there
// are at least m; then every distinct key gets a seeded random tag, and
// sorting by tag gives both a uniform choice of m of them and their order.
#include "tsvd_internal.h"

#include <hipcub/hipcub.hpp>

namespace tsvd {

namespace {

constexpr int TPB = 256;
constexpr uint64_t NONE = ~0ull;

__device__ __host__ inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void k_rmat_draw(uint64_t *__restrict__ keys, int64_t count, int32_t scale, int32_t n, uint32_t ta,
                            uint32_t tab, uint32_t tabc, uint64_t seed, uint64_t round)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    const uint64_t base = mix64(seed ^ mix64(round * 0xD1B54A32D192ED03ull + (uint64_t)e));
    uint64_t u = 0, v = 0, h = 0;
    for (int32_t l = 0; l < scale; l++) {
        if ((l & 3) == 0) h = mix64(base + (uint64_t)(l >> 2) * 0x9E3779B97F4A7C15ull);
        const uint32_t r = (uint32_t)(h & 0xFFFF);
        h >>= 16;
        const uint64_t bu = r >= tab ? 1 : 0;                  // quadrants c, d: lower half
        const uint64_t bv = (r >= ta && r < tab) || r >= tabc ? 1 : 0;   // quadrants b, d: right half
        u = (u << 1) | bu;
        v = (v << 1) | bv;
    }
    u %= (uint64_t)n;
    v %= (uint64_t)n;
    keys[e] = u == v ? NONE : u * (uint64_t)n + v;
}

__global__ void k_weight_out(double *__restrict__ w, int32_t n, double expo)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) w[i] = pow((double)(i + 1), expo);
}

__global__ void k_iota_tags(uint64_t *__restrict__ tags, int32_t *__restrict__ idx, int32_t n, uint64_t seed)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    tags[i] = mix64(seed ^ mix64((uint64_t)i));
    idx[i] = (int32_t)i;
}

// order[pos] = the vertex at position pos of the seeded permutation
__global__ void k_weight_in(double *__restrict__ w, const int32_t *__restrict__ order, int32_t n, double expo)
{
    const int64_t pos = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pos < n) w[order[pos]] = pow((double)(pos + 1), expo);
}

// first index i with cdf[i] > x (cdf ascending, n entries), clipped to n-1
__device__ inline int32_t upper_index(const double *__restrict__ cdf, int32_t n, double x)
{
    int32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const int32_t mid = lo + ((hi - lo) >> 1);
        if (cdf[mid] > x) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

__global__ void k_chung_lu_draw(uint64_t *__restrict__ keys, int64_t count, int32_t n,
                                const double *__restrict__ cdf_out, const double *__restrict__ cdf_in,
                                uint64_t seed, uint64_t round)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    const uint64_t base = mix64(seed ^ mix64(round * 0xD1B54A32D192ED03ull + (uint64_t)e));
    const double x1 = (double)(mix64(base + 0x9E3779B97F4A7C15ull) >> 11) * 0x1.0p-53;
    const double x2 = (double)(mix64(base + 0x3C6EF372FE94F82Aull) >> 11) * 0x1.0p-53;
    const uint64_t u = (uint64_t)upper_index(cdf_out, n, x1 * cdf_out[n - 1]);
    const uint64_t v = (uint64_t)upper_index(cdf_in, n, x2 * cdf_in[n - 1]);
    keys[e] = u == v ? NONE : u * (uint64_t)n + v;
}

__global__ void k_tag(const uint64_t *__restrict__ keys, uint64_t *__restrict__ tags, int64_t count, uint64_t seed)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) tags[i] = mix64(seed ^ mix64(keys[i]));
}

__global__ void k_split(const uint64_t *__restrict__ keys, int64_t count, int32_t n, int32_t *__restrict__ src,
                        int32_t *__restrict__ dst)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint64_t k = keys[i];
    src[i] = (int32_t)(k / (uint64_t)n);
    dst[i] = (int32_t)(k % (uint64_t)n);
}

// sorted unique of keys[0, count) into out; returns the count without NONE
int64_t sort_unique(uint64_t *keys, uint64_t *alt, int64_t count, DBuf<uint8_t> &tmp, DBuf<int64_t> &num,
                    hipStream_t st)
{
    size_t need = 0, need2 = 0;
    TSVD_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, need, keys, alt, count, 0, 64, st));
    TSVD_HIP(hipcub::DeviceSelect::Unique(nullptr, need2, alt, keys, num.p, count, st));
    tmp.ensure(std::max(need, need2) + 256);
    need = tmp.n;
    TSVD_HIP(hipcub::DeviceRadixSort::SortKeys(tmp.p, need, keys, alt, count, 0, 64, st));
    need = tmp.n;
    TSVD_HIP(hipcub::DeviceSelect::Unique(tmp.p, need, alt, keys, num.p, count, st));
    int64_t u = 0;
    TSVD_HIP(hipMemcpyAsync(&u, num.p, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    uint64_t last = 0;
    TSVD_HIP(hipStreamSynchronize(st));
    if (u > 0) {
        TSVD_HIP(hipMemcpyAsync(&last, keys + u - 1, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
        TSVD_HIP(hipStreamSynchronize(st));
        if (last == NONE) u--;
    }
    return u;
}

// Draws rounds of candidate keys with `draw(keys + have, count, round)` until
// at least m distinct ones exist, then picks m of them uniformly at random in
// a random order and splits them into device arrays *src, *dst (hipMalloc'd,
// released by the caller with tsvd_device_free).
template <class Draw>
void distinct_edges(int32_t n, int64_t m, uint64_t seed, hipStream_t st, Draw draw, int32_t **src, int32_t **dst)
{
    // keys[0, have) hold the distinct keys so far; a round appends draws behind them
    const int64_t cap = m + m / 2 + 4096;
    DBuf<uint64_t> keys((size_t)cap), alt((size_t)cap);
    DBuf<uint8_t> tmp;
    DBuf<int64_t> num(1);
    int64_t have = 0;
    for (uint64_t round = 0; have < m; round++) {
        TSVD_REQUIRE(round < 16, "only %lld distinct edges after 16 rounds of draws (need %lld)",
                     (long long)have, (long long)m);
        const int64_t cnt = std::min(cap - have, (m - have) + (m - have) / 4 + 4096);
        draw(keys.p + have, cnt, round);
        TSVD_HIP(hipGetLastError());
        have = sort_unique(keys.p, alt.p, have + cnt, tmp, num, st);
    }
    // seeded random tags; sorting by tag picks m keys uniformly and orders them
    DBuf<uint64_t> tags((size_t)have), tags2((size_t)have);
    hipLaunchKernelGGL(k_tag, dim3((unsigned)ceil_div(have, TPB)), dim3(TPB), 0, st, keys.p, tags.p, have,
                       mix64(seed + 0x5eed));
    size_t need = 0;
    TSVD_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, need, tags.p, tags2.p, keys.p, alt.p, have, 0, 64, st));
    tmp.ensure(need + 256);
    need = tmp.n;
    TSVD_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, need, tags.p, tags2.p, keys.p, alt.p, have, 0, 64, st));
    int32_t *s = nullptr, *d = nullptr;
    TSVD_HIP(hipMalloc((void **)&s, sizeof(int32_t) * (size_t)m));
    if (hipMalloc((void **)&d, sizeof(int32_t) * (size_t)m) != hipSuccess) {
        (void)hipFree(s);
        throw Error{set_error(TSVD_ENOMEM, "hipMalloc of the edge arrays failed")};
    }
    hipLaunchKernelGGL(k_split, dim3((unsigned)ceil_div(m, TPB)), dim3(TPB), 0, st, alt.p, m, n, s, d);
    TSVD_HIP(hipGetLastError());
    TSVD_HIP(hipStreamSynchronize(st));
    *src = s;
    *dst = d;
}

struct StreamGuard {
    hipStream_t s = nullptr;
    ~StreamGuard() { if (s) (void)hipStreamDestroy(s); }
};

}  // namespace

}  // namespace tsvd

using namespace tsvd;

extern "C" int tsvd_gen_rmat(int32_t scale, int32_t n, int64_t m, double a, double b, double c, uint64_t seed,
                             int32_t **src, int32_t **dst)
{
    return guard([&] {
        TSVD_REQUIRE(src && dst, "null output");
        TSVD_REQUIRE(scale >= 1 && scale <= 31 && n >= 2 && n <= (int64_t)1 << scale, "bad scale or n");
        TSVD_REQUIRE(m >= 1 && m < (int64_t)INT32_MAX, "m out of range");
        TSVD_REQUIRE(a > 0 && b >= 0 && c >= 0 && a + b + c < 1.0, "bad quadrant probabilities");
        StreamGuard sg;
        TSVD_HIP(hipStreamCreateWithFlags(&sg.s, hipStreamNonBlocking));
        hipStream_t st = sg.s;
        const uint32_t ta = (uint32_t)std::lround(a * 65536.0), tab = (uint32_t)std::lround((a + b) * 65536.0),
                       tabc = (uint32_t)std::lround((a + b + c) * 65536.0);
        distinct_edges(n, m, seed, st, [&](uint64_t *keys, int64_t cnt, uint64_t round) {
            hipLaunchKernelGGL(k_rmat_draw, dim3((unsigned)ceil_div(cnt, TPB)), dim3(TPB), 0, st, keys, cnt, scale,
                               n, ta, tab, tabc, seed, round);
        }, src, dst);
    });
}

extern "C" int tsvd_gen_chung_lu(int32_t n, int64_t m, double out_exp, double in_exp, uint64_t seed, int32_t **src,
                                 int32_t **dst)
{
    return guard([&] {
        TSVD_REQUIRE(src && dst, "null output");
        TSVD_REQUIRE(n >= 2, "n must be >= 2");
        TSVD_REQUIRE(m >= 1 && m < (int64_t)INT32_MAX && (double)m < 0.5 * (double)n * (double)(n - 1),
                     "m out of range");
        TSVD_REQUIRE(out_exp > 1.0 && in_exp > 1.0, "power-law exponents must exceed 1");
        StreamGuard sg;
        TSVD_HIP(hipStreamCreateWithFlags(&sg.s, hipStreamNonBlocking));
        hipStream_t st = sg.s;
        const unsigned gn = (unsigned)ceil_div(n, TPB);
        DBuf<double> w((size_t)n), cdf_out((size_t)n), cdf_in((size_t)n);
        DBuf<uint8_t> tmp;
        size_t need = 0;
        // out-weights (i+1)^(-1/(out_exp-1)), cumulated
        hipLaunchKernelGGL(k_weight_out, dim3(gn), dim3(TPB), 0, st, w.p, n, -1.0 / (out_exp - 1.0));
        TSVD_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, need, w.p, cdf_out.p, n, st));
        tmp.ensure(need + 256);
        need = tmp.n;
        TSVD_HIP(hipcub::DeviceScan::InclusiveSum(tmp.p, need, w.p, cdf_out.p, n, st));
        // in-weights through a seeded random permutation (sort of hashed tags)
        {
            DBuf<uint64_t> tags((size_t)n), tags2((size_t)n);
            DBuf<int32_t> idx((size_t)n), order((size_t)n);
            hipLaunchKernelGGL(k_iota_tags, dim3(gn), dim3(TPB), 0, st, tags.p, idx.p, n, mix64(seed + 0x9e37));
            need = 0;
            TSVD_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, need, tags.p, tags2.p, idx.p, order.p, n, 0, 64,
                                                        st));
            tmp.ensure(need + 256);
            need = tmp.n;
            TSVD_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, need, tags.p, tags2.p, idx.p, order.p, n, 0, 64, st));
            hipLaunchKernelGGL(k_weight_in, dim3(gn), dim3(TPB), 0, st, w.p, order.p, n, -1.0 / (in_exp - 1.0));
            TSVD_HIP(hipStreamSynchronize(st));
        }
        need = 0;
        TSVD_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, need, w.p, cdf_in.p, n, st));
        tmp.ensure(need + 256);
        need = tmp.n;
        TSVD_HIP(hipcub::DeviceScan::InclusiveSum(tmp.p, need, w.p, cdf_in.p, n, st));
        TSVD_HIP(hipStreamSynchronize(st));
        distinct_edges(n, m, seed, st, [&](uint64_t *keys, int64_t cnt, uint64_t round) {
            hipLaunchKernelGGL(k_chung_lu_draw, dim3((unsigned)ceil_div(cnt, TPB)), dim3(TPB), 0, st, keys, cnt, n,
                               cdf_out.p, cdf_in.p, seed, round);
        }, src, dst);
    });
}
