// frpca_noalloc.hip -- frPCA of a device-resident CSR matrix that allocates nothing.
//
// The computation of tsvd_frpca_csr (frPCA, frpca.c:239-326: Gaussian sketch,
// q passes of A X / A^T Y with eigSVD orthonormalisation, eigSVD of the last
// A^T Q) for callers that cannot allocate device memory: code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc.  The matrix
// and the outputs are the caller's device arrays; every temporary lives in
// `scratch`, one file-scope device array per GPU (held until the process
// ends), sized for m <= 4096 rows, n <= 32768 columns and k + 5 <= 136:
//   Qt   n x l   (row-major: one row of the sketch side per column of A)
//   Q,Q2 m x l   (row-major)
//   G, W l x l   (FP64 Gram / rotation), lam l, GEMM descriptors, flags,
//   and the eigensolver's workspace (eig_workspace_bytes(136, 1)).
// The sparse products are hand-written CSR kernels on those row-major
// operands (one wavefront per row of A, its column ids and values broadcast
// from one coalesced load; A^T Y scatters with FP64 atomic adds), the Grams
// and rotations run on the grouped MFMA GEMM, the eigensolves on the
// library's own batched solver (eig.hip) in that workspace.  The sketch draws the same
// numbers as tsvd_frpca_csr with the same seed, so the two agree to rounding.
// Calls are serialised by a mutex; concurrent calls from other processes use
// their own copy of the array.
#include "tsvd_pipeline.h"

#include <mutex>

namespace tsvd {
namespace {

constexpr int64_t NA_MMAX = 4096;
constexpr int64_t NA_NMAX = 32768;
constexpr int64_t NA_LMAX = 136;
constexpr int NA_DESC = 8;
constexpr size_t na_round(size_t x) { return (x + 255) / 256 * 256; }
constexpr size_t OFF_FLAG = 0;
constexpr size_t OFF_DESC = 256;
constexpr size_t OFF_QT = na_round(OFF_DESC + NA_DESC * sizeof(GemmProb));
constexpr size_t OFF_Q = na_round(OFF_QT + NA_NMAX * NA_LMAX * sizeof(double));
constexpr size_t OFF_Q2 = na_round(OFF_Q + NA_MMAX * NA_LMAX * sizeof(double));
constexpr size_t OFF_G = na_round(OFF_Q2 + NA_MMAX * NA_LMAX * sizeof(double));
constexpr size_t OFF_W = na_round(OFF_G + NA_LMAX * NA_LMAX * sizeof(double));
constexpr size_t OFF_LAM = na_round(OFF_W + NA_LMAX * NA_LMAX * sizeof(double));
constexpr size_t OFF_EIG = na_round(OFF_LAM + NA_LMAX * sizeof(double));
constexpr size_t EIG_BYTES = eig_workspace_bytes((int)NA_LMAX, 1);
constexpr size_t SCRATCH_BYTES = na_round(OFF_EIG + EIG_BYTES);

}  // namespace
}  // namespace tsvd

// The workspace of tsvd_frpca_csr_noalloc (about 45 MB per GPU).  In host code
// this name is only a handle for hipGetSymbolAddress; kernels get the device
// address that call returns.
static __device__ __align__(256) uint8_t scratch[tsvd::SCRATCH_BYTES];

namespace tsvd {
namespace {

constexpr int TPB = 256;

__device__ __host__ inline uint64_t na_mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void k_na_fill(double *__restrict__ x, int64_t count, double v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) x[i] = v;
}

__global__ void k_na_zero_flag(int32_t *flag)
{
    if (threadIdx.x == 0) flag[0] = 0;
}

// 1-based CSR rows in range and column ids in [1, n]; bits of *flag on failure
__global__ void k_na_validate(int32_t m, int32_t n, int32_t nnz, const int32_t *__restrict__ cols,
                              const int32_t *__restrict__ pb, const int32_t *__restrict__ pe, int32_t *flag)
{
    const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    const int32_t b = pb[r], e = pe[r];
    if (b < 1 || e < b || e - 1 > nnz) {
        atomicOr(flag, 1);
        return;
    }
    for (int32_t t = b - 1; t < e - 1; t++) {
        const int32_t c = cols[t];
        if (c < 1 || c > n) {
            atomicOr(flag, 2);
            return;
        }
    }
}

// X (rows x l, row-major): entry (i, c) is the Gaussian that the column-major
// sketch of tsvd_frpca_csr puts at index c * rows + i (same hash, same
// Box-Muller transform)
__global__ void k_na_sketch(double *__restrict__ X, int64_t rows, int32_t l, uint64_t seed)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= rows * l) return;
    const int64_t c = t / rows, i = t - c * rows;
    const uint64_t h = na_mix64(seed ^ na_mix64((uint64_t)t));
    const double u1 = ((double)(h >> 40) + 1.0) / 16777217.0;
    const double u2 = (double)(h & 0xFFFFFFull) / 16777216.0;
    X[i * l + c] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// Y (m x l) = A X, X (n x l); one wavefront per row of A, lanes over columns
__global__ void __launch_bounds__(TPB) k_na_spmm(int32_t m, int32_t l, const double *__restrict__ vals,
                                                 const int32_t *__restrict__ cols, const int32_t *__restrict__ pb,
                                                 const int32_t *__restrict__ pe, const double *__restrict__ X,
                                                 double *__restrict__ Y)
{
    const int lane = threadIdx.x & 63;
    const int32_t r = blockIdx.x * (TPB / 64) + (threadIdx.x >> 6);
    if (r >= m) return;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    const int32_t b = pb[r] - 1, e = pe[r] - 1;
    for (int32_t t0 = b; t0 < e; t0 += 64) {
        const int32_t t = t0 + lane;
        int32_t cj = 0;
        double av = 0.0;
        if (t < e) {
            cj = cols[t] - 1;
            av = vals[t];
        }
        const int cnt = min(64, e - t0);
        for (int u = 0; u < cnt; u++) {
            const int32_t j = __shfl(cj, u);
            const double a = __shfl(av, u);
            const double *__restrict__ x = X + (int64_t)j * l;
            if (lane < l) a0 += a * x[lane];
            if (lane + 64 < l) a1 += a * x[lane + 64];
            if (lane + 128 < l) a2 += a * x[lane + 128];
        }
    }
    double *__restrict__ y = Y + (int64_t)r * l;
    if (lane < l) y[lane] = a0;
    if (lane + 64 < l) y[lane + 64] = a1;
    if (lane + 128 < l) y[lane + 128] = a2;
}

// Y (n x l) += A^T X, X (m x l): one wavefront per row r of A scatters
// a_rj X[r, :] into Y[j, :] with FP64 atomic adds (Y zeroed beforehand)
__global__ void __launch_bounds__(TPB) k_na_spmm_t(int32_t m, int32_t l, const double *__restrict__ vals,
                                                   const int32_t *__restrict__ cols, const int32_t *__restrict__ pb,
                                                   const int32_t *__restrict__ pe, const double *__restrict__ X,
                                                   double *__restrict__ Y)
{
    const int lane = threadIdx.x & 63;
    const int32_t r = blockIdx.x * (TPB / 64) + (threadIdx.x >> 6);
    if (r >= m) return;
    const double *__restrict__ x = X + (int64_t)r * l;
    const double x0 = lane < l ? x[lane] : 0.0;
    const double x1 = lane + 64 < l ? x[lane + 64] : 0.0;
    const double x2 = lane + 128 < l ? x[lane + 128] : 0.0;
    const int32_t b = pb[r] - 1, e = pe[r] - 1;
    for (int32_t t0 = b; t0 < e; t0 += 64) {
        const int32_t t = t0 + lane;
        int32_t cj = 0;
        double av = 0.0;
        if (t < e) {
            cj = cols[t] - 1;
            av = vals[t];
        }
        const int cnt = min(64, e - t0);
        for (int u = 0; u < cnt; u++) {
            const int32_t j = __shfl(cj, u);
            const double a = __shfl(av, u);
            double *__restrict__ y = Y + (int64_t)j * l;
            if (lane < l) atomicAdd(y + lane, a * x0);
            if (lane + 64 < l) atomicAdd(y + lane + 64, a * x1);
            if (lane + 128 < l) atomicAdd(y + lane + 128, a * x2);
        }
    }
}

// S[c] = sqrt(lam[l - k + c]) (ascending, the reference's order), zero at or
// below the eigSVD guard 1e-4 (frpca.c:106-122)
__global__ void k_na_sigma(double *__restrict__ S, const double *__restrict__ lam, int32_t l, int32_t k)
{
    const int c = threadIdx.x + blockIdx.x * blockDim.x;
    if (c >= k) return;
    const double s = sqrt(fmax(lam[l - k + c], 0.0));
    S[c] = s > 1e-4 ? s : 0.0;
}

struct NoAlloc {
    std::mutex mu;
    uint8_t *ptr[64] = {};
    GemmRunner gemm;
};

NoAlloc &na()
{
    static NoAlloc s;
    return s;
}

// the calling device's copy of `scratch` (its address is looked up once per device)
uint8_t *na_scratch(NoAlloc &s)
{
    int dev = -1;
    TSVD_HIP(hipGetDevice(&dev));
    TSVD_REQUIRE(dev >= 0 && dev < 64, "device index %d out of range", dev);
    if (!s.ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        TSVD_REQUIRE(e == hipSuccess && p, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        s.ptr[dev] = (uint8_t *)p;
    }
    return s.ptr[dev];
}

void na_fill(double *x, int64_t count, double v, hipStream_t st)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(k_na_fill, dim3((unsigned)ceil_div(count, TPB)), dim3(TPB), 0, st, x, count, v);
    TSVD_HIP(hipGetLastError());
}

struct NaCsr {
    int32_t m, n, nnz;
    const double *vals;
    const int32_t *cols, *pb, *pe;
};

void na_product(const NaCsr &a, int32_t l, bool transposed, const double *X, double *Y, hipStream_t st)
{
    const unsigned grid = (unsigned)ceil_div(a.m, TPB / 64);
    if (!transposed) {
        PROF("spmm_noalloc", st);
        hipLaunchKernelGGL(k_na_spmm, dim3(grid), dim3(TPB), 0, st, a.m, l, a.vals, a.cols, a.pb, a.pe, X, Y);
    } else {
        na_fill(Y, (int64_t)a.n * l, 0.0, st);
        PROF("spmm_t_noalloc", st);
        hipLaunchKernelGGL(k_na_spmm_t, dim3(grid), dim3(TPB), 0, st, a.m, l, a.vals, a.cols, a.pb, a.pe, X, Y);
    }
    TSVD_HIP(hipGetLastError());
}

// G = X^T X (FP64) of a row-major rows x l matrix, split over K with atomic adds
void na_gram(const double *X, int64_t rows, int32_t l, double *G, GemmProb *desc, hipStream_t st)
{
    na_fill(G, (int64_t)l * l, 0.0, st);
    GemmProb p{};
    p.A = X; p.lda = l; p.transA = 0;        // A(i, r) = X[r * l + i]
    p.B = X; p.ldb = l; p.transB = 1;        // B(r, j) = X[r * l + j]
    p.C = G; p.ldc = l;
    p.M = l; p.N = l; p.K = (int32_t)rows; p.alpha = 1.0; p.beta = 0.0; p.kchunk = 512;
    na().gemm.run_into(st, &p, 1, GemmTypes{DT::F64, DT::F64, DT::F64, DT::F64}, l > 80, desc, NA_DESC,
                       "gemm_gram_noalloc");
}

// Q <- Q W with W = V diag(lam^-1/2) (eigSVD's orthonormalisation); returns the new Q
double *na_orth(double *Q, double *Q2, int32_t m, int32_t l, uint8_t *sc, hipStream_t st)
{
    double *G = (double *)(sc + OFF_G), *W = (double *)(sc + OFF_W), *lam = (double *)(sc + OFF_LAM);
    GemmProb *desc = (GemmProb *)(sc + OFF_DESC);
    na_gram(Q, m, l, G, desc, st);
    eig_batched(st, G, l, 1, lam, 0, sc + OFF_EIG, EIG_BYTES, na().gemm, desc, NA_DESC);
    eig_rotation(W, G, lam, l, 1, 1e-4, 1e-13, st);
    GemmProb p{};
    p.A = W; p.lda = l; p.transA = 1;        // op(A)(i, k) = W[i * l + k] = W(k, i)
    p.B = Q; p.ldb = l;                      // B(k, r) = Q[r * l + k]
    p.C = Q2; p.ldc = l;                     // C(i, r) = Q2[r * l + i]
    p.M = l; p.N = m; p.K = l; p.alpha = 1.0; p.beta = 0.0;
    na().gemm.run_into(st, &p, 1, GemmTypes{DT::F64, DT::F64, DT::F64, DT::F64}, l > 80, desc, NA_DESC,
                       "gemm_rotate_noalloc");
    return Q2;
}

// Out (rows x k, column-major, ld rows) = X (rows x l, row-major) Wk (l x k)
void na_project(const double *X, int64_t rows, int32_t l, const double *Wk, int32_t k, double *Out, GemmProb *desc,
                hipStream_t st)
{
    GemmProb p{};
    p.A = X; p.lda = l; p.transA = 1;        // op(A)(r, c) = X[r * l + c]
    p.B = Wk; p.ldb = l;
    p.C = Out; p.ldc = rows;
    p.M = (int32_t)rows; p.N = k; p.K = l; p.alpha = 1.0; p.beta = 0.0;
    na().gemm.run_into(st, &p, 1, GemmTypes{DT::F64, DT::F64, DT::F64, DT::F64}, k > 80, desc, NA_DESC,
                       "gemm_project_noalloc");
}

}  // namespace
}  // namespace tsvd

using namespace tsvd;

extern "C" int tsvd_frpca_csr_noalloc(int32_t m, int32_t n, int32_t nnz, const double *values, const int32_t *cols,
                                      const int32_t *pointerB, const int32_t *pointerE, double *U, double *S,
                                      double *V, int32_t k, int32_t q, uint64_t seed)
{
    return guard([&] {
        TSVD_REQUIRE(m >= 1 && m <= NA_MMAX, "rows m = %d outside [1, %lld] (the no-allocation workspace limit)", m,
                     (long long)NA_MMAX);
        TSVD_REQUIRE(n >= 1 && n <= NA_NMAX, "columns n = %d outside [1, %lld] (the no-allocation workspace limit)",
                     n, (long long)NA_NMAX);
        TSVD_REQUIRE(k >= 1 && k + 5 <= NA_LMAX, "rank k = %d outside [1, %lld] (the no-allocation workspace limit)",
                     k, (long long)NA_LMAX - 5);
        TSVD_REQUIRE(q >= 1, "q must be >= 1");
        TSVD_REQUIRE(nnz >= 0 && (nnz == 0 || (values && cols)) && pointerB && pointerE && U && S && V,
                     "null argument");
        NoAlloc &s = na();
        std::lock_guard<std::mutex> lk(s.mu);
        uint8_t *sc = na_scratch(s);
        hipStream_t st = nullptr;   // the null stream: nothing to create
        const int32_t l = k + 5;
        int32_t *flag = (int32_t *)(sc + OFF_FLAG);
        GemmProb *desc = (GemmProb *)(sc + OFF_DESC);
        double *Qt = (double *)(sc + OFF_QT), *Q = (double *)(sc + OFF_Q), *Q2 = (double *)(sc + OFF_Q2);
        double *G = (double *)(sc + OFF_G), *W = (double *)(sc + OFF_W), *lam = (double *)(sc + OFF_LAM);
        // the matrix is checked on the device before any product reads it
        hipLaunchKernelGGL(k_na_zero_flag, dim3(1), dim3(64), 0, st, flag);
        hipLaunchKernelGGL(k_na_validate, dim3((unsigned)ceil_div(m, TPB)), dim3(TPB), 0, st, m, n, nnz, cols,
                           pointerB, pointerE, flag);
        TSVD_HIP(hipGetLastError());
        int32_t hflag = 0;
        TSVD_HIP(hipMemcpyAsync(&hflag, flag, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        TSVD_HIP(hipStreamSynchronize(st));
        TSVD_REQUIRE(!(hflag & 1), "bad row pointers (1-based pointerB/pointerE within nnz expected)");
        TSVD_REQUIRE(!(hflag & 2), "column index outside [1, %d]", n);
        const NaCsr a{m, n, nnz, values, cols, pointerB, pointerE};
        // frpca.c:239-326 with eigSVD orthonormalisation between passes (as tsvd_frpca_csr)
        if (q % 2 == 0) {
            hipLaunchKernelGGL(k_na_sketch, dim3((unsigned)ceil_div((int64_t)n * l, TPB)), dim3(TPB), 0, st, Qt,
                               (int64_t)n, l, seed);
            na_product(a, l, false, Qt, Q, st);
            double *nq = na_orth(Q, Q2, m, l, sc, st);
            Q2 = Q;
            Q = nq;
        } else {
            hipLaunchKernelGGL(k_na_sketch, dim3((unsigned)ceil_div((int64_t)m * l, TPB)), dim3(TPB), 0, st, Q,
                               (int64_t)m, l, seed);
        }
        TSVD_HIP(hipGetLastError());
        const int niter = (q - 1) / 2;
        for (int i = 1; i <= niter; i++) {
            na_product(a, l, true, Q, Qt, st);
            na_product(a, l, false, Qt, Q, st);
            double *nq = na_orth(Q, Q2, m, l, sc, st);
            Q2 = Q;
            Q = nq;
        }
        na_product(a, l, true, Q, Qt, st);
        // eigSVD(Qt): U = Q V[:, sel], V = Qt V[:, sel] / S, S ascending (frpca.c:296-316)
        na_gram(Qt, n, l, G, desc, st);
        eig_batched(st, G, l, 1, lam, 0, sc + OFF_EIG, EIG_BYTES, na().gemm, desc, NA_DESC);
        eig_select(W, G, lam, l, k, 1, 3, 1, 1e-4, st);
        na_project(Q, m, l, W, k, U, desc, st);
        eig_select(W, G, lam, l, k, 1, 2, 1, 1e-4, st);
        na_project(Qt, n, l, W, k, V, desc, st);
        hipLaunchKernelGGL(k_na_sigma, dim3(1), dim3(256), 0, st, S, lam, l, k);
        TSVD_HIP(hipGetLastError());
        TSVD_HIP(hipStreamSynchronize(st));
    });
}
