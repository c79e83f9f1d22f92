// frpca.hip -- blockwise randomized SVD (frPCA) on the device.
//
// Replaces frPCA (frpca.c:239-326) with its helpers eigSVD (frpca.c:99-136),
// LUfraction (frpca.c:4-41), initialize_random_matrix (matrix_vector_functions_
// intel_mkl.c:479-512) and the MKL sparse products (..._ext.c:364-398), called
// by the reference once per dirty block (Tree_u_NC.cpp:679-710, serially).
//
// All dirty blocks are factored together: every product below is one grouped
// MFMA launch over the blocks (gemm.hip) and every small eigenproblem one
// batched FP64 solve.  Passes follow frpca.c (q even, s = 5, l = k + s):
//     Qt <- Gaussian (nc x l)                 (counter-based, seeded)
//     Q  <- A Qt ;  Q <- orth(Q)
//     (q-2)/2 times:  Qt <- A^T Q ;  Q <- A Qt ;  Q <- orth(Q)
//     Qt <- A^T Q ;  [V, lam] = eig(Qt^T Qt)   (eigSVD of frpca.c:99-136)
//     U = Q V[:, top k],  S = sqrt(lam[top k]),  V = Qt V[:, top k] / S.
// orth(Q) is eigSVD's own orthonormalisation, Q V diag(lam^-1/2) with the FP64
// Gram of Q, used at every step (the reference uses LU normalisation between
// power iterations and eigSVD on the last one; both only re-basis span(Q)).
// Products run in the matrix precision (FP32 MFMA for FP32 blocks, FP64 MFMA
// for FP64 blocks); Grams and eigensolves are FP64.
#include "tsvd_pipeline.h"

namespace tsvd {

static GemmTypes types_for(int precision, DT a, DT b, DT c, DT acc)
{
    (void)precision;
    return GemmTypes{a, b, c, acc};
}

// G = X^T X (l x l, FP64, split over K with atomic adds) of a rows x l
// operand: column-major with leading dimension ldx, or row-major (ld l).
static GemmProb gram_prob(const void *X, int64_t ldx, int32_t rows, int32_t l, bool rowmajor, double *G)
{
    GemmProb p{};
    if (!rowmajor) {
        p.A = X; p.lda = ldx; p.transA = 1;       // op(A)(i, k) = X(k, i)
        p.B = X; p.ldb = ldx;
    } else {
        p.A = X; p.lda = l; p.transA = 0;         // A(i, k) = X[k * l + i] = X(k, i)
        p.B = X; p.ldb = l; p.transB = 1;         // B(k, j) = X[k * l + j]
    }
    p.C = G; p.ldc = l;
    p.M = l; p.N = l; p.K = rows; p.alpha = 1.0; p.beta = 0.0; p.kchunk = 512;
    return p;
}

// Out (rows x n, column-major, leading dimension ldo) = X W with W l x n
// column-major; X column-major (leading dimension ldx) or row-major (ld l).
static GemmProb right_prob(const void *X, int64_t ldx, int32_t rows, int32_t l, bool rowmajor, const double *W,
                           int32_t n, void *Out, int64_t ldo)
{
    GemmProb p{};
    p.A = X;
    if (!rowmajor) { p.lda = ldx; p.transA = 0; }
    else { p.lda = l; p.transA = 1; }             // op(A)(r, c) = X[r * l + c]
    p.B = W; p.ldb = l;
    p.C = Out; p.ldc = ldo;
    p.M = rows; p.N = n; p.K = l; p.alpha = 1.0; p.beta = 0.0;
    return p;
}

// Q <- orth(Q) for every item: G = Q^T Q (FP64), eig, Q2 = Q V lam^-1/2, swap.
static void orth(std::vector<FrpcaItem> &items, int precision, int32_t l, double *G, double *W, double *lam,
                 GemmRunner &gemm, Eig &eig, hipStream_t st)
{
    const DT T = precision == TSVD_FP64 ? DT::F64 : DT::F32;
    const int cnt = (int)items.size();
    TSVD_HIP(hipMemsetAsync(G, 0, sizeof(double) * (size_t)l * l * cnt, st));
    std::vector<GemmProb> gp;
    for (int i = 0; i < cnt; i++) {
        const FrpcaItem &it = items[i];
        gp.push_back(gram_prob(it.Q, it.ldq, it.m, l, it.sp >= 0, G + (size_t)i * l * l));
    }
    gemm.run(st, gp, types_for(precision, T, T, DT::F64, DT::F64), l > 80, "gemm_gram");
    eig.solve(st, G, l, cnt, lam);
    const double eps_rel = precision == TSVD_FP64 ? 1e-13 : 1e-6;
    eig_rotation(W, G, lam, l, cnt, 1e-4, eps_rel, st);
    std::vector<GemmProb> rp, rp_rm;
    for (int i = 0; i < cnt; i++) {
        FrpcaItem &it = items[i];
        GemmProb p{};
        if (it.sp < 0) {
            p.A = it.Q; p.lda = it.ldq; p.transA = 0;
            p.B = W + (size_t)i * l * l; p.ldb = l;
            p.C = it.Q2; p.ldc = it.ldq;
            p.M = it.m; p.N = l; p.K = l;
        } else {
            // row-major Q2 is the column-major Q2^T = W^T Q^T (l x m, ld l)
            p.A = W + (size_t)i * l * l; p.lda = l; p.transA = 1;
            p.B = it.Q; p.ldb = l;
            p.C = it.Q2; p.ldc = l;
            p.M = l; p.N = it.m; p.K = l;
        }
        p.alpha = 1.0; p.beta = 0.0;
        (it.sp < 0 ? rp : rp_rm).push_back(p);
    }
    gemm.run(st, rp, types_for(precision, T, DT::F64, T, DT::F64), l > 80, "gemm_rotate");
    // row-major items: the FP64 rotation is the first operand
    gemm.run(st, rp_rm, types_for(precision, DT::F64, T, T, DT::F64), l > 80, "gemm_rotate");
    for (auto &it : items) std::swap(it.Q, it.Q2);
}

static void product(std::vector<FrpcaItem> &items, int precision, int32_t l, bool transposed, GemmRunner &gemm,
                     hipStream_t st, SparseLeaves *sp)
{
    const DT T = precision == TSVD_FP64 ? DT::F64 : DT::F32;
    std::vector<GemmProb> ps;
    bool any_sparse = false;
    for (auto &it : items) {
        if (it.sp >= 0) {
            sp->items[(size_t)it.sp].Q = it.Q;
            sp->items[(size_t)it.sp].Qt = it.Qt;
            any_sparse = true;
            continue;
        }
        GemmProb p{};
        p.A = it.A; p.lda = it.lda; p.alpha = 1.0; p.beta = 0.0; p.N = l;
        if (!transposed) {   // Q = A Qt
            p.transA = 0; p.B = it.Qt; p.ldb = it.ldqt; p.C = it.Q; p.ldc = it.ldq;
            p.M = it.m; p.K = it.nc;
        } else {             // Qt = A^T Q
            p.transA = 1; p.B = it.Q; p.ldb = it.ldq; p.C = it.Qt; p.ldc = it.ldqt;
            p.M = it.nc; p.K = it.m;
        }
        ps.push_back(p);
    }
    gemm.run(st, ps, GemmTypes{T, T, T, T}, l > 80, transposed ? "gemm_sketch_t" : "gemm_sketch_n");
    if (any_sparse) leaf_spmm(*sp, precision, l, transposed, st);
}

void frpca_core(std::vector<FrpcaItem> &items, int precision, int32_t l, int32_t q, double *G, double *W,
                double *lam, GemmRunner &gemm, Eig &eig, hipStream_t st, SparseLeaves *sp)
{
    TSVD_REQUIRE(q >= 1, "frPCA needs q >= 1");
    const DT T = precision == TSVD_FP64 ? DT::F64 : DT::F32;
    std::vector<GaussItem> sk;
    if (q % 2 == 0) {
        for (auto &it : items) sk.push_back(GaussItem{it.Qt, it.nc, l, it.ldqt, it.seed, it.sp >= 0});
        gaussian_fill(sk, precision, st);
        product(items, precision, l, false, gemm, st, sp);
        orth(items, precision, l, G, W, lam, gemm, eig, st);
    } else {
        for (auto &it : items) sk.push_back(GaussItem{it.Q, it.m, l, it.ldq, it.seed, it.sp >= 0});
        gaussian_fill(sk, precision, st);
    }
    const int niter = (q - 1) / 2;
    for (int i = 1; i <= niter; i++) {
        product(items, precision, l, true, gemm, st, sp);
        product(items, precision, l, false, gemm, st, sp);
        orth(items, precision, l, G, W, lam, gemm, eig, st);
    }
    product(items, precision, l, true, gemm, st, sp);
    // eigSVD(Qt): Gram of Qt and its eigenvectors
    const int cnt = (int)items.size();
    TSVD_HIP(hipMemsetAsync(G, 0, sizeof(double) * (size_t)l * l * cnt, st));
    std::vector<GemmProb> gp;
    for (int i = 0; i < cnt; i++) {
        const FrpcaItem &it = items[i];
        gp.push_back(gram_prob(it.Qt, it.ldqt, it.nc, l, it.sp >= 0, G + (size_t)i * l * l));
    }
    gemm.run(st, gp, GemmTypes{T, T, DT::F64, DT::F64}, l > 80, "gemm_gram");
    eig.solve(st, G, l, cnt, lam);
}

void frpca_setup(FrpcaWork &w, const Blocks &b, int32_t k, int32_t q)
{
    TSVD_REQUIRE(k >= 1, "rank must be positive");
    w.k = k;
    w.l = k + 5;
    w.q = q;
    const size_t es = b.esize;
    w.Qt.alloc((size_t)b.n * w.l * es);
    w.Q.alloc((size_t)b.nparts * b.s_pad * w.l * es);
    w.Q2.alloc((size_t)b.nparts * b.s_pad * w.l * es);
    w.G.alloc((size_t)b.nparts * w.l * w.l);
    w.W.alloc((size_t)b.nparts * w.l * w.l);
    w.lam.alloc((size_t)b.nparts * w.l);
    // TSVD_SPARSE_DENSITY overrides the density below which a dirty block is
    // factored from its sparse copies (0: always dense)
    if (const char *e = getenv("TSVD_SPARSE_DENSITY")) w.sparse_density = atof(e);
    TSVD_HIP(hipMemset(w.Q.p, 0, w.Q.bytes()));
    TSVD_HIP(hipMemset(w.Q2.p, 0, w.Q2.bytes()));
}

void frpca_blocks(FrpcaWork &w, Blocks &b, const std::vector<int32_t> &dirty, void *leaves, GemmRunner &gemm,
                  Eig &eig, uint64_t seed, hipStream_t st, std::vector<double> &sig2)
{
    const int32_t l = w.l, k = w.k;
    const size_t es = b.esize;
    std::vector<FrpcaItem> items;
    for (size_t i = 0; i < dirty.size(); i++) {
        const int32_t j = dirty[i];
        FrpcaItem it{};
        it.A = b.A.p + (size_t)b.start[j] * b.s_pad * es;
        it.lda = b.s_pad;
        it.m = b.nsrc;
        it.nc = b.ncols[j];
        it.Qt = w.Qt.p + (size_t)b.start[j] * l * es;
        it.ldqt = std::max<int64_t>(b.width[j], 1);
        it.Q = w.Q.p + i * (size_t)b.s_pad * l * es;
        it.Q2 = w.Q2.p + i * (size_t)b.s_pad * l * es;
        it.ldq = b.s_pad;
        it.seed = seed * 0x9E3779B97F4A7C15ull + (uint64_t)j * 0xD1B54A32D192ED03ull + 1;
        items.push_back(it);
    }
    // blocks with few enough non-zeros are factored from CSR / CSC copies
    std::vector<SpItem> sparse;
    w.last_sparse = 0;
    if (w.sparse_density > 0.0 && l <= 192) {
        std::vector<SpItem> cand;
        for (auto &it : items) cand.push_back(SpItem{it.A, it.lda, it.m, it.nc, 0, it.Q, it.Qt});
        const std::vector<double> dens = leaf_densities(w.sp, cand, b.precision, st);
        for (size_t i = 0; i < items.size(); i++) {
            if (items[i].nc <= 0 || dens[i] >= w.sparse_density) continue;
            items[i].sp = (int32_t)sparse.size();
            sparse.push_back(cand[i]);
        }
    }
    w.last_sparse = (int64_t)sparse.size();
    leaf_sparse_build(w.sp, std::move(sparse), b.precision, st);
    frpca_core(items, b.precision, l, w.q, w.G.p, w.W.p, w.lam.p, gemm, eig, st, &w.sp);
    const int cnt = (int)items.size();
    // leaf = U diag(S) = Q V[:, top k] diag(sqrt(lam))   (Tree_u_NC.cpp:147-154)
    eig_select(w.W.p, w.G.p, w.lam.p, l, k, cnt, 0, 0, 1e-4, st);
    const DT T = b.precision == TSVD_FP64 ? DT::F64 : DT::F32;
    std::vector<GemmProb> lp;
    for (int i = 0; i < cnt; i++)
        lp.push_back(right_prob(items[i].Q, items[i].ldq, b.nsrc, l, items[i].sp >= 0, w.W.p + (size_t)i * l * k, k,
                                (uint8_t *)leaves + (size_t)dirty[i] * b.s_pad * k * es, b.s_pad));
    gemm.run(st, lp, GemmTypes{T, DT::F64, T, DT::F64}, k > 80, "gemm_leaf");
    std::vector<double> lam((size_t)cnt * l);
    TSVD_HIP(hipMemcpyAsync(lam.data(), w.lam.p, sizeof(double) * lam.size(), hipMemcpyDeviceToHost, st));
    TSVD_HIP(hipStreamSynchronize(st));
    sig2.assign(cnt, 0.0);
    for (int i = 0; i < cnt; i++)
        for (int c = l - k; c < l; c++) {
            const double lv = std::max(lam[(size_t)i * l + c], 0.0);
            if (std::sqrt(lv) > 1e-4) sig2[i] += lv;
        }
}

}  // namespace tsvd

using namespace tsvd;

// ------------------------------------------------------------ drop-in frPCA
// Host CSR in (1-based cols, pointerB/pointerE, ext.h:22-33), host U, S, V out
// (col-major FP64, caller-allocated, S ascending as frpca.c:296-312 leaves it).
namespace {

__global__ void k_csr_to_dense(const double *__restrict__ vals, const int *__restrict__ cols,
                               const int *__restrict__ pb, const int *__restrict__ pe, int nrows, double *__restrict__ D,
                               int64_t ld)
{
    const int r = blockIdx.x;
    if (r >= nrows) return;
    const int b = pb[r] - 1, e = pe[r] - 1;
    for (int t = b + threadIdx.x; t < e; t += blockDim.x) D[(int64_t)(cols[t] - 1) * ld + r] = vals[t];
}

struct DropIn {
    GemmRunner gemm;
    Eig eig;
};

// frPCA of `count` independent CSR matrices in one grouped run (item i uses
// sketch seed seed + i): every product is one grouped launch over the items
// and every eigensolve one batched call.
int frpca_csr_batch_impl(const mat_csr *const *A, mat *const *U, mat *const *S, mat *const *V, int count, int k,
                         int q, uint64_t seed)
{
    return guard([&] {
        TSVD_REQUIRE(count >= 1 && A && U && S && V, "null argument or empty batch");
        TSVD_REQUIRE(k >= 1 && q >= 1, "k and q must be positive");
        const int l = k + 5;
        for (int i = 0; i < count; i++) {
            TSVD_REQUIRE(A[i] && U[i] && S[i] && V[i], "null matrix in item %d", i);
            const mat_csr *a = A[i];
            const int m = a->nrows, n = a->ncols;
            TSVD_REQUIRE(m >= 1 && n >= 1 && a->nnz >= 0, "bad matrix shape in item %d", i);
            TSVD_REQUIRE(U[i]->nrows == m && U[i]->ncols == k && S[i]->nrows == k && V[i]->nrows == n &&
                             V[i]->ncols == k,
                         "item %d: output shapes must be U m x k, S k x 1, V n x k", i);
            // the CSR is validated on the host (1-based, in range) before any copy
            for (int r = 0; r < m; r++) {
                TSVD_REQUIRE(a->pointerB[r] >= 1 && a->pointerE[r] >= a->pointerB[r] && a->pointerE[r] - 1 <= a->nnz,
                             "item %d: bad row pointers at row %d", i, r);
                for (int t = a->pointerB[r] - 1; t < a->pointerE[r] - 1; t++)
                    TSVD_REQUIRE(a->cols[t] >= 1 && a->cols[t] <= n, "item %d: column index out of range at row %d",
                                 i, r);
            }
        }
        hipStream_t st;
        TSVD_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() { (void)hipStreamDestroy(s); }
        } sg{st};
        static DropIn *ctx = nullptr;   // handles reused across calls (one per process)
        if (!ctx) ctx = new DropIn();
        struct Dev {
            DBuf<double> A, Qt, Q, Q2, U, V, vals;
            DBuf<int> cols, pb, pe;
            int64_t ld;
        };
        std::vector<Dev> dv((size_t)count);
        std::vector<FrpcaItem> items((size_t)count);
        for (int i = 0; i < count; i++) {
            const mat_csr *a = A[i];
            const int m = a->nrows, n = a->ncols;
            Dev &d = dv[(size_t)i];
            d.ld = round_up(m, 64);
            d.A.alloc((size_t)d.ld * n);
            d.Qt.alloc((size_t)n * l);
            d.Q.alloc((size_t)d.ld * l);
            d.Q2.alloc((size_t)d.ld * l);
            d.U.alloc((size_t)m * k);
            d.V.alloc((size_t)n * k);
            d.vals.alloc((size_t)std::max(a->nnz, 1));
            d.cols.alloc((size_t)std::max(a->nnz, 1));
            d.pb.alloc((size_t)m);
            d.pe.alloc((size_t)m);
            TSVD_HIP(hipMemsetAsync(d.A.p, 0, d.A.bytes(), st));
            TSVD_HIP(hipMemsetAsync(d.Q.p, 0, d.Q.bytes(), st));
            TSVD_HIP(hipMemsetAsync(d.Q2.p, 0, d.Q2.bytes(), st));
            if (a->nnz > 0) {
                TSVD_HIP(hipMemcpyAsync(d.vals.p, a->values, sizeof(double) * a->nnz, hipMemcpyHostToDevice, st));
                TSVD_HIP(hipMemcpyAsync(d.cols.p, a->cols, sizeof(int) * a->nnz, hipMemcpyHostToDevice, st));
            }
            TSVD_HIP(hipMemcpyAsync(d.pb.p, a->pointerB, sizeof(int) * m, hipMemcpyHostToDevice, st));
            TSVD_HIP(hipMemcpyAsync(d.pe.p, a->pointerE, sizeof(int) * m, hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_csr_to_dense, dim3(m), dim3(256), 0, st, d.vals.p, d.cols.p, d.pb.p, d.pe.p, m, d.A.p,
                               d.ld);
            TSVD_HIP(hipGetLastError());
            items[(size_t)i] = FrpcaItem{d.A.p, d.ld, m, n, d.Qt.p, n, d.Q.p, d.Q2.p, d.ld, seed + (uint64_t)i};
        }
        DBuf<double> G((size_t)l * l * count), W((size_t)l * l * count), lam((size_t)l * count),
            Wk((size_t)l * k * count);
        frpca_core(items, TSVD_FP64, l, q, G.p, W.p, lam.p, ctx->gemm, ctx->eig, st);
        std::vector<double> hl((size_t)l * count);
        TSVD_HIP(hipMemcpyAsync(hl.data(), lam.p, sizeof(double) * hl.size(), hipMemcpyDeviceToHost, st));
        // U = Q V[:, s..s+k-1] (frpca.c:314-316), ascending
        eig_select(Wk.p, G.p, lam.p, l, k, count, 3, 1, 1e-4, st);
        std::vector<GemmProb> pu((size_t)count), pv((size_t)count);
        for (int i = 0; i < count; i++) {
            const Dev &d = dv[(size_t)i];
            const FrpcaItem &it = items[(size_t)i];
            GemmProb g{};
            g.A = it.Q; g.lda = d.ld; g.B = Wk.p + (size_t)i * l * k; g.ldb = l; g.C = d.U.p; g.ldc = it.m;
            g.M = it.m; g.N = k; g.K = l; g.alpha = 1.0;
            pu[(size_t)i] = g;
        }
        ctx->gemm.run(st, pu, GemmTypes{DT::F64, DT::F64, DT::F64, DT::F64}, k > 80);
        // V = Qt V[:, sel] / S (eigSVD's U of Qt, frpca.c:118-131)
        eig_select(Wk.p, G.p, lam.p, l, k, count, 2, 1, 1e-4, st);
        for (int i = 0; i < count; i++) {
            const Dev &d = dv[(size_t)i];
            const FrpcaItem &it = items[(size_t)i];
            GemmProb g{};
            g.A = d.Qt.p; g.lda = it.nc; g.B = Wk.p + (size_t)i * l * k; g.ldb = l; g.C = d.V.p; g.ldc = it.nc;
            g.M = it.nc; g.N = k; g.K = l; g.alpha = 1.0;
            pv[(size_t)i] = g;
        }
        ctx->gemm.run(st, pv, GemmTypes{DT::F64, DT::F64, DT::F64, DT::F64}, k > 80);
        for (int i = 0; i < count; i++) {
            const Dev &d = dv[(size_t)i];
            TSVD_HIP(hipMemcpyAsync(U[i]->d, d.U.p, d.U.bytes(), hipMemcpyDeviceToHost, st));
            TSVD_HIP(hipMemcpyAsync(V[i]->d, d.V.p, d.V.bytes(), hipMemcpyDeviceToHost, st));
        }
        TSVD_HIP(hipStreamSynchronize(st));
        for (int i = 0; i < count; i++)
            for (int c = 0; c < k; c++) {
                const double sv = std::sqrt(std::max(hl[(size_t)i * l + (size_t)(l - k + c)], 0.0));
                S[i]->d[c] = sv > 1e-4 ? sv : 0.0;
            }
    });
}

int frpca_csr_impl(const mat_csr *A, mat *U, mat *S, mat *V, int k, int q, uint64_t seed)
{
    return frpca_csr_batch_impl(&A, &U, &S, &V, 1, k, q, seed);
}

}  // namespace

extern "C" int tsvd_frpca_csr(const mat_csr *A, mat *U, mat *S, mat *V, int k, int q, uint64_t seed)
{
    return frpca_csr_impl(A, U, S, V, k, q, seed);
}

extern "C" int tsvd_frpca_csr_batch(const mat_csr *const *A, mat *const *U, mat *const *S, mat *const *V,
                                    int count, int k, int q, uint64_t seed)
{
    return frpca_csr_batch_impl(A, U, S, V, count, k, q, seed);
}

extern "C" void frPCA(mat_csr *A, mat **U, mat **S, mat **V, int k, int q)
{
    static uint64_t calls = 0;
    if (!A || !U || !S || !V || !*U || !*S || !*V) {
        set_error(TSVD_EINVAL, "frPCA: null argument");
        return;
    }
    if (frpca_csr_impl(A, *U, *S, *V, k, q, 0x5eed0000ull + (calls++)) != TSVD_OK) {
        // void like the reference: the outputs are left zero, the reason is in tsvd_last_error()
        mat *o[3] = {*U, *S, *V};
        for (mat *x : o)
            if (x->d && x->nrows > 0 && x->ncols > 0) memset(x->d, 0, sizeof(double) * (size_t)x->nrows * x->ncols);
    }
}
