// gemm.hip -- grouped dense products on CDNA4 matrix cores.
//
// Used by the blockwise frPCA (the products A*X and A^T*Y of every dirty
// column block in one launch; the reference calls mkl_dcsrmm once per block,
// matrix_vector_functions_intel_mkl_ext.c:364-398), by the FP64 Gram matrices
// of the orthonormalisations and of the tree merges (dgemm / dgesvd work of
// frpca.c:99-136 and Tree_u_NC.cpp:238-262), and by the basis rotations.
//
// A workgroup of 4 wavefronts owns a 64 x (16*NT) tile of C; wave w owns rows
// [16w, 16w+16) and all NT 16-column MFMA tiles, so one A fragment feeds NT
// matrix instructions (NT = 9 covers rank + oversampling = 133 in one tile).
// K advances in 32-deep slabs staged through LDS (double-buffered through
// registers: the next slab is fetched from HBM while the current one is fed
// to the MFMAs).  MFMA 16x16x4 operand layout: lane l holds A(l%16, l/16) and
// B(l/16, l%16); the FP32 accumulator holds C(4*(l/16)+r, l%16) and the FP64
// accumulator C(l/16 + 4r, l%16), r = 0..3 (both checked on gfx950).
#include "tsvd_gemm.h"

#include <map>
#include <mutex>

namespace tsvd {

namespace {

constexpr int BM = 64;
constexpr int BK = 16;    // K slab of the bf16x3 products
constexpr int KB = 32;    // K slab of the FP32 / FP64 products
constexpr int TPB = 256;

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <class T>
struct Acc;
template <>
struct Acc<float> {
    typedef f32x4 V;
    static __device__ inline V mfma(float a, float b, V c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    static __device__ inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};
template <>
struct Acc<double> {
    typedef f64x4 V;
    static __device__ inline V mfma(double a, double b, V c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    static __device__ inline int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};

__device__ inline void atomic_add_out(float *p, double v) { atomicAdd(p, (float)v); }
__device__ inline void atomic_add_out(double *p, double v) { atomicAdd(p, v); }

template <class TA, class TB, class TC, class TACC, int NT>
__global__ void __launch_bounds__(TPB) k_gemm(const GemmProb *__restrict__ probs, int nprob)
{
    constexpr int BN = 16 * NT;
    constexpr int A_PER = BM * KB / TPB;             // 8
    constexpr int B_PER = (BN * KB + TPB - 1) / TPB; // 18 or 10
    // operands stay in their own type in registers and LDS (converted to
    // TACC as they feed the MFMAs); one padding element per row: the stash
    // writes 32 consecutive k of a column into 32 rows, which land in
    // different banks
    __shared__ TA As[KB][BM + 1];
    __shared__ TB Bs[KB][BN + 1];

    // locate the problem (binary search over tile_begin)
    const int64_t tile = blockIdx.x;
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (probs[mid].tile_begin <= tile) lo = mid;
        else hi = mid - 1;
    }
    const GemmProb &P = probs[lo];
    int64_t t = tile - P.tile_begin;
    const int tm = (int)(t % P.tiles_m);
    t /= P.tiles_m;
    const int tn = (int)(t % P.tiles_n);
    const int tk = (int)(t / P.tiles_n);
    const int M = P.M, N = P.N;
    const int i0 = tm * BM, j0 = tn * BN;
    int kbeg = 0, kend = P.K;
    if (P.tiles_k > 1) {
        kbeg = tk * P.kchunk;
        kend = min(P.K, kbeg + P.kchunk);
    }
    const gptr<const TA> __restrict__ A = gbl((const TA *)P.A);
    const gptr<const TB> __restrict__ B = gbl((const TB *)P.B);
    const int64_t lda = P.lda, ldb = P.ldb;
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;

    typename Acc<TACC>::V acc[NT];
#pragma unroll
    for (int n = 0; n < NT; n++) acc[n] = typename Acc<TACC>::V{0, 0, 0, 0};

    TA ra[A_PER];
    TB rb[B_PER];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int q = 0; q < A_PER; q++) {
            const int e = tid + q * TPB;
            int ii, kk;
            if (P.transA) { ii = e / KB; kk = e % KB; }
            else { kk = e / BM; ii = e % BM; }
            const int gi = i0 + ii, gk = k0 + kk;
            TA v = TA(0);
            if (gi < M && gk < kend) v = P.transA ? A[(int64_t)gi * lda + gk] : A[(int64_t)gk * lda + gi];
            ra[q] = v;
        }
#pragma unroll
        for (int q = 0; q < B_PER; q++) {
            const int e = tid + q * TPB;
            int jj, kk;
            if (P.transB) { kk = e / BN; jj = e % BN; }
            else { jj = e / KB; kk = e % KB; }
            const int gj = j0 + jj, gk = k0 + kk;
            TB v = TB(0);
            if (e < BN * KB && gj < N && gk < kend)
                v = P.transB ? B[(int64_t)gk * ldb + gj] : B[(int64_t)gj * ldb + gk];
            rb[q] = v;
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int q = 0; q < A_PER; q++) {
            const int e = tid + q * TPB;
            int ii, kk;
            if (P.transA) { ii = e / KB; kk = e % KB; }
            else { kk = e / BM; ii = e % BM; }
            As[kk][ii] = ra[q];
        }
#pragma unroll
        for (int q = 0; q < B_PER; q++) {
            const int e = tid + q * TPB;
            if (e < BN * KB) {
                if (P.transB) Bs[e / BN][e % BN] = rb[q];
                else Bs[e % KB][e / KB] = rb[q];
            }
        }
    };

    if (kbeg < kend) fetch(kbeg);
    for (int k0 = kbeg; k0 < kend; k0 += KB) {
        __syncthreads();
        stash();
        __syncthreads();
        if (k0 + KB < kend) fetch(k0 + KB);
        // two k-steps of operands at a time (a full unroll hoists all eight
        // steps' LDS reads and halves the occupancy)
#pragma unroll 2
        for (int ks = 0; ks < KB; ks += 4) {
            const TACC a = (TACC)As[ks + (lane >> 4)][wave * 16 + (lane & 15)];
#pragma unroll
            for (int n = 0; n < NT; n++) {
                const TACC b = (TACC)Bs[ks + (lane >> 4)][n * 16 + (lane & 15)];
                acc[n] = Acc<TACC>::mfma(a, b, acc[n]);
            }
        }
    }

    TC *__restrict__ C = (TC *)P.C;
    const int64_t ldc = P.ldc;
    const double alpha = P.alpha, beta = P.beta;
#pragma unroll
    for (int n = 0; n < NT; n++) {
        const int gj = j0 + n * 16 + (lane & 15);
        if (gj >= N) continue;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int gi = i0 + wave * 16 + Acc<TACC>::row(lane, r);
            if (gi >= M) continue;
            TC *c = C + (int64_t)gj * ldc + gi;
            const double v = alpha * (double)acc[n][r];
            if (P.tiles_k > 1) atomic_add_out(c, v);
            else *c = (TC)(beta == 0.0 ? v : v + beta * (double)*c);
        }
    }
}

// bf16x3 products (the merge products of the FP32 pipelines): every FP32
// operand x is split into bf16 hi = rne(x) and lo = rne(x - hi) when staged to
// LDS, and each 16 x 16 x 16 slab is hi*hi + hi*lo + lo*hi on
// v_mfma_f32_16x16x16bf16_1k with FP32 accumulation (about 16 significant
// bits; lo*lo is dropped).  Lane l holds A(l % 16, 4 (l / 16) + e), B(4 (l /
// 16) + e, l % 16), e = 0..3, and the accumulator C(4 (l / 16) + r, l % 16)
// (checked on gfx950).  A is M x K column-major, B K x N column-major (FP32
// or FP64, rounded to FP32 first), no transposes, no K split.
typedef short bf16x4 __attribute__((ext_vector_type(4)));

__device__ inline unsigned short bf16_rne(float x)
{
    const unsigned u = __float_as_uint(x);
    return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

__device__ inline float bf16_val(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

template <class TB, class TC, int NT>
__global__ void __launch_bounds__(TPB) k_gemm_bf3(const GemmProb *__restrict__ probs, int nprob)
{
    constexpr int BN = 16 * NT;
    constexpr int A_PER = BM * BK / TPB;
    constexpr int B_PER = (BN * BK + TPB - 1) / TPB;
    __shared__ unsigned short Ah[BK][BM], Al[BK][BM];
    __shared__ unsigned short Bh[BK][BN], Bl[BK][BN];
    const int64_t tile = blockIdx.x;
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (probs[mid].tile_begin <= tile) lo = mid;
        else hi = mid - 1;
    }
    const GemmProb &P = probs[lo];
    int64_t t = tile - P.tile_begin;
    const int tm = (int)(t % P.tiles_m);
    const int tn = (int)(t / P.tiles_m);
    const int M = P.M, N = P.N, K = P.K;
    const int i0 = tm * BM, j0 = tn * BN;
    const gptr<const float> __restrict__ A = gbl((const float *)P.A);
    const gptr<const TB> __restrict__ B = gbl((const TB *)P.B);
    const int64_t lda = P.lda, ldb = P.ldb;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    f32x4 acc[NT];
#pragma unroll
    for (int n = 0; n < NT; n++) acc[n] = f32x4{0.f, 0.f, 0.f, 0.f};
    float ra[A_PER], rb[B_PER];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int q = 0; q < A_PER; q++) {
            const int e = tid + q * TPB, kk = e / BM, ii = e % BM;
            const int gi = i0 + ii, gk = k0 + kk;
            // clamped (not conditional) loads, so that they are all in flight together
            const float v = A[(int64_t)min(gk, K - 1) * lda + min(gi, M - 1)];
            ra[q] = (gi < M && gk < K) ? v : 0.f;
        }
#pragma unroll
        for (int q = 0; q < B_PER; q++) {
            const int e = tid + q * TPB, jj = e / BK, kk = e % BK;
            const int gj = j0 + jj, gk = k0 + kk;
            const float v = (float)B[(int64_t)min(gj, N - 1) * ldb + min(gk, K - 1)];
            rb[q] = (e < BN * BK && gj < N && gk < K) ? v : 0.f;
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int q = 0; q < A_PER; q++) {
            const int e = tid + q * TPB, kk = e / BM, ii = e % BM;
            const unsigned short h = bf16_rne(ra[q]);
            Ah[kk][ii] = h;
            Al[kk][ii] = bf16_rne(ra[q] - bf16_val(h));
        }
#pragma unroll
        for (int q = 0; q < B_PER; q++) {
            const int e = tid + q * TPB;
            if (e < BN * BK) {
                const unsigned short h = bf16_rne(rb[q]);
                Bh[e % BK][e / BK] = h;
                Bl[e % BK][e / BK] = bf16_rne(rb[q] - bf16_val(h));
            }
        }
    };
    if (K > 0) fetch(0);
    for (int k0 = 0; k0 < K; k0 += BK) {
        __syncthreads();
        stash();
        __syncthreads();
        if (k0 + BK < K) fetch(k0 + BK);
        bf16x4 ah, al;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int kk = 4 * (lane >> 4) + e, ii = wave * 16 + (lane & 15);
            ah[e] = (short)Ah[kk][ii];
            al[e] = (short)Al[kk][ii];
        }
#pragma unroll
        for (int n = 0; n < NT; n++) {
            bf16x4 bh, bl;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int kk = 4 * (lane >> 4) + e, jj = n * 16 + (lane & 15);
                bh[e] = (short)Bh[kk][jj];
                bl[e] = (short)Bl[kk][jj];
            }
            acc[n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh, acc[n], 0, 0, 0);
            acc[n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl, acc[n], 0, 0, 0);
            acc[n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh, acc[n], 0, 0, 0);
        }
    }
    TC *__restrict__ C = (TC *)P.C;
    const int64_t ldc = P.ldc;
    const double alpha = P.alpha, beta = P.beta;
#pragma unroll
    for (int n = 0; n < NT; n++) {
        const int gj = j0 + n * 16 + (lane & 15);
        if (gj >= N) continue;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int gi = i0 + wave * 16 + 4 * (lane >> 4) + r;
            if (gi >= M) continue;
            TC *c = C + (int64_t)gj * ldc + gi;
            const double v = alpha * (double)acc[n][r];
            *c = (TC)(beta == 0.0 ? v : v + beta * (double)*c);
        }
    }
}

template <class TB, class TC>
void launch_bf3(hipStream_t st, const GemmProb *d, int nprob, int64_t tiles, bool wide)
{
    if (wide) hipLaunchKernelGGL((k_gemm_bf3<TB, TC, 9>), dim3((unsigned)tiles), dim3(TPB), 0, st, d, nprob);
    else hipLaunchKernelGGL((k_gemm_bf3<TB, TC, 5>), dim3((unsigned)tiles), dim3(TPB), 0, st, d, nprob);
}

template <class TA, class TB, class TC, class TACC>
void launch_t(hipStream_t st, const GemmProb *d, int nprob, int64_t tiles, bool wide)
{
    if (wide) hipLaunchKernelGGL((k_gemm<TA, TB, TC, TACC, 9>), dim3((unsigned)tiles), dim3(TPB), 0, st, d, nprob);
    else hipLaunchKernelGGL((k_gemm<TA, TB, TC, TACC, 5>), dim3((unsigned)tiles), dim3(TPB), 0, st, d, nprob);
}

}  // namespace

void GemmRunner::run(hipStream_t st, std::vector<GemmProb> &probs, GemmTypes t, bool wide, const char *label)
{
    int nlive = 0;
    for (auto &p : probs) nlive += (p.M > 0 && p.N > 0);
    dprobs_.ensure((size_t)std::max(nlive, 1));
    run_into(st, probs.data(), (int)probs.size(), t, wide, dprobs_.p, nlive, label);
}

namespace {

struct TuneRec {
    int wide = -1;
    double ms[2] = {0.0, 0.0};
    int64_t calls = 0;
};

std::mutex g_tune_mu;
bool g_tune_on = false;
std::map<std::string, TuneRec> g_tuned;

std::string tune_key(const char *label, GemmTypes t, const GemmProb *probs, int nprobs)
{
    int64_t nmax = 0, kmax = 0;
    for (int i = 0; i < nprobs; i++) {
        nmax = std::max<int64_t>(nmax, probs[i].N);
        kmax = std::max<int64_t>(kmax, probs[i].K);
    }
    int kcls = 0;
    while ((1ll << kcls) < kmax) kcls++;
    char buf[160];
    snprintf(buf, sizeof buf, "%s|%d%d%d%d|N%lld|K2^%d", label ? label : "gemm", (int)t.a, (int)t.b, (int)t.c,
             (int)t.acc, (long long)nmax, kcls);
    return buf;
}

bool idempotent(const GemmProb *probs, int nprobs)
{
    for (int i = 0; i < nprobs; i++) {
        const GemmProb &p = probs[i];
        if (p.M <= 0 || p.N <= 0) continue;
        if ((p.kchunk > 0 && p.K > p.kchunk) || p.beta != 0.0 || p.C == p.A || p.C == p.B) return false;
    }
    return true;
}

}  // namespace

void gemm_tuning_enable(bool on)
{
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune_on = on;
}

std::string gemm_tuning_json()
{
    std::lock_guard<std::mutex> g(g_tune_mu);
    std::string out = "{";
    bool first = true;
    for (auto &kv : g_tuned) {
        char buf[320];
        snprintf(buf, sizeof buf, "%s\"%s\": {\"tile_n\": %d, \"ms_tile80\": %.4f, \"ms_tile144\": %.4f, \"calls\": %lld}",
                 first ? "" : ", ", kv.first.c_str(), kv.second.wide ? 144 : 80, kv.second.ms[0], kv.second.ms[1],
                 (long long)kv.second.calls);
        out += buf;
        first = false;
    }
    return out + "}";
}

void GemmRunner::run_into(hipStream_t st, GemmProb *probs, int nprobs, GemmTypes t, bool wide, GemmProb *dev,
                          int cap, const char *label)
{
    const std::string key = tune_key(label, t, probs, nprobs);
    bool tune_now = false;
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end() && it->second.wide >= 0) {
            wide = it->second.wide != 0;
            it->second.calls++;
        } else if (g_tune_on && idempotent(probs, nprobs)) {
            tune_now = true;
        }
    }
    if (!tune_now) {
        launch_once(st, probs, nprobs, t, wide, dev, cap, label);
        return;
    }
    // both widths on the same inputs (each overwrites C), the faster one kept
    hipEvent_t ev[3];
    for (auto &e : ev) TSVD_HIP(hipEventCreate(&e));
    TSVD_HIP(hipEventRecord(ev[0], st));
    launch_once(st, probs, nprobs, t, false, dev, cap, label);
    TSVD_HIP(hipEventRecord(ev[1], st));
    launch_once(st, probs, nprobs, t, true, dev, cap, label);
    TSVD_HIP(hipEventRecord(ev[2], st));
    TSVD_HIP(hipEventSynchronize(ev[2]));
    float a = 0.f, b = 0.f;
    TSVD_HIP(hipEventElapsedTime(&a, ev[0], ev[1]));
    TSVD_HIP(hipEventElapsedTime(&b, ev[1], ev[2]));
    for (auto &e : ev) TSVD_HIP(hipEventDestroy(e));
    std::lock_guard<std::mutex> g(g_tune_mu);
    TuneRec &r = g_tuned[key];
    r.ms[0] = a;
    r.ms[1] = b;
    r.wide = b < a ? 1 : 0;
    r.calls++;
}

void GemmRunner::launch_once(hipStream_t st, GemmProb *probs, int nprobs, GemmTypes t, bool wide, GemmProb *dev,
                             int cap, const char *label)
{
    const int BN = wide ? 144 : 80;
    int64_t tiles = 0;
    GemmProb live[64];
    int np = 0;
    TSVD_REQUIRE(cap >= 0, "bad descriptor capacity");
    for (int i = 0; i < nprobs; i++) {
        GemmProb &p = probs[i];
        if (p.M <= 0 || p.N <= 0) continue;
        TSVD_REQUIRE(p.K >= 0, "negative K");
        p.tiles_m = (int32_t)ceil_div(p.M, BM);
        p.tiles_n = (int32_t)ceil_div(p.N, BN);
        p.tiles_k = 1;
        if (p.kchunk > 0 && p.K > p.kchunk) {
            TSVD_REQUIRE(p.kchunk % KB == 0, "kchunk must be a multiple of 32");
            p.tiles_k = (int32_t)ceil_div(p.K, p.kchunk);
        }
        p.tile_begin = tiles;
        tiles += (int64_t)p.tiles_m * p.tiles_n * p.tiles_k;
        TSVD_REQUIRE(np < cap, "more GEMM problems (%d) than descriptor slots (%d)", np + 1, cap);
        // copied in chunks of 64 problems
        live[np % 64] = p;
        np++;
        if (np % 64 == 0) {
            TSVD_HIP(hipMemcpyAsync(dev + np - 64, live, sizeof(GemmProb) * 64, hipMemcpyHostToDevice, st));
            // the staging array is reused for the next chunk
            TSVD_HIP(hipStreamSynchronize(st));
        }
    }
    if (np == 0 || tiles == 0) return;
    TSVD_REQUIRE(tiles < (int64_t)INT32_MAX, "too many GEMM tiles");
    if (np % 64 != 0)
        TSVD_HIP(hipMemcpyAsync(dev + (np / 64) * 64, live, sizeof(GemmProb) * (np % 64), hipMemcpyHostToDevice, st));
    // the staging array must outlive the copy
    TSVD_HIP(hipStreamSynchronize(st));
    if (g_profile) {
        auto sz = [](DT d) { return d == DT::F64 ? 8.0 : 4.0; };   // BF3 is never an operand type
        double flops = 0.0, bytes = 0.0;
        for (int i = 0; i < nprobs; i++) {
            const GemmProb &p = probs[i];
            if (p.M <= 0 || p.N <= 0) continue;
            flops += 2.0 * p.M * p.N * p.K;
            bytes += (double)p.M * p.K * sz(t.a) + (double)p.K * p.N * sz(t.b) + (double)p.M * p.N * sz(t.c);
        }
        prof_add_work(label, bytes, flops);
    }
    PROF(label, st);
    if (t.acc == DT::BF3) {
        for (int i = 0; i < nprobs; i++)
            TSVD_REQUIRE(probs[i].transA == 0 && probs[i].transB == 0 && probs[i].tiles_k <= 1,
                         "bf16x3 products take untransposed operands and no K split");
        TSVD_REQUIRE(t.a == DT::F32, "bf16x3 products take an FP32 A");
        if (t.b == DT::F64 && t.c == DT::F32) launch_bf3<double, float>(st, dev, np, tiles, wide);
        else if (t.b == DT::F64 && t.c == DT::F64) launch_bf3<double, double>(st, dev, np, tiles, wide);
        else if (t.b == DT::F32 && t.c == DT::F32) launch_bf3<float, float>(st, dev, np, tiles, wide);
        else TSVD_REQUIRE(false, "unsupported bf16x3 type combination");
    } else if (t.acc == DT::F32) {
        TSVD_REQUIRE(t.a == DT::F32 && t.b == DT::F32 && t.c == DT::F32, "FP32 accumulate takes FP32 operands");
        launch_t<float, float, float, float>(st, dev, np, tiles, wide);
    } else {
        if (t.a == DT::F32 && t.b == DT::F32 && t.c == DT::F64) launch_t<float, float, double, double>(st, dev, np, tiles, wide);
        else if (t.a == DT::F32 && t.b == DT::F64 && t.c == DT::F32) launch_t<float, double, float, double>(st, dev, np, tiles, wide);
        else if (t.a == DT::F64 && t.b == DT::F64 && t.c == DT::F64) launch_t<double, double, double, double>(st, dev, np, tiles, wide);
        else if (t.a == DT::F32 && t.b == DT::F64 && t.c == DT::F64) launch_t<float, double, double, double>(st, dev, np, tiles, wide);
        else if (t.a == DT::F64 && t.b == DT::F64 && t.c == DT::F32) launch_t<double, double, float, double>(st, dev, np, tiles, wide);
        else if (t.a == DT::F64 && t.b == DT::F32 && t.c == DT::F32) launch_t<double, float, float, double>(st, dev, np, tiles, wide);
        else TSVD_REQUIRE(false, "unsupported GEMM type combination");
    }
    TSVD_HIP(hipGetLastError());
    launches++;
}

}  // namespace tsvd
