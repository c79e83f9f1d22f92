// tree.hip -- dirty-path merges up the fan-out tree and the root embedding.
//
// Replaces the merge loop of the drivers (Tree_u_NC.cpp:723-850, Tree_d_NC.cpp:
// 995-1114, Tree_u_LP.cpp:873-989) with append_matrices_horizontally
// (matrix_vector_functions_intel_mkl.c:1024-1054), dense_sub_svd_function
// (Tree_u_NC.cpp:238-262) and truncated_singular_value_decomposition
// (..._intel_mkl.c:1259-1322, LAPACK dgesvd), run there one parent at a time.
//
// A parent's input C = [child_1 | ... | child_f] is never copied: the children
// of a parent are consecutive s_pad x k blocks of their level's array, i.e. one
// s_pad x (f*k) column-major matrix.  All dirty parents of a level are merged
// together: one grouped FP64 Gram C^T C, one batched FP64 eigensolve, one
// grouped product parent = C V_k (= U_k S_k, the reference's U_d * Sigma_d).
// The root gives E = C V_k S_k^-1/2 = U_k sqrt(S_k) (NC, Tree_u_NC.cpp:765-772)
// or C V_k S_k^-1 = U_k (LP, Tree_u_LP.cpp:907-912).  Levels whose size is not
// a power of f (ragged trees) are padded with zero nodes, whose columns only
// add zero eigenvalues.
#include "tsvd_pipeline.h"

#include <algorithm>
#include <set>

namespace tsvd {

void tree_setup(Tree &t, int32_t nparts, int32_t fanout, int32_t k, int32_t nsrc, int32_t s_pad, size_t esize,
                hipStream_t st)
{
    TSVD_REQUIRE(fanout >= 2, "fan-out must be at least 2");
    t.fanout = fanout;
    t.k = k;
    t.nsrc = nsrc;
    t.s_pad = s_pad;
    t.esize = esize;
    t.sizes.clear();
    t.sizes.push_back(nparts);
    while (t.sizes.back() > 1) t.sizes.push_back((int32_t)ceil_div(t.sizes.back(), fanout));
    if (t.sizes.size() == 1) t.sizes.push_back(1);   // a single block still gets a root step
    t.padded.resize(t.sizes.size());
    t.level.clear();
    t.level.resize(t.sizes.size());
    for (size_t l = 0; l < t.sizes.size(); l++) {
        t.padded[l] = (l + 1 < t.sizes.size()) ? (int32_t)round_up(t.sizes[l], fanout) : t.sizes[l];
        t.level[l].alloc((size_t)t.padded[l] * s_pad * k * esize);
        TSVD_HIP(hipMemsetAsync(t.level[l].p, 0, t.level[l].bytes(), st));
    }
    t.E.alloc((size_t)s_pad * k);
    TSVD_HIP(hipMemsetAsync(t.E.p, 0, t.E.bytes(), st));
    t.root_sigma.assign(k, 0.0);
    t.have_root = false;
    TSVD_HIP(hipStreamSynchronize(st));
}

int32_t tree_update(Tree &t, const std::vector<int32_t> &dirty_leaves, int transform, GemmRunner &gemm, Eig &eig,
                    hipStream_t st)
{
    const int32_t f = t.fanout, k = t.k, nc = f * k;
    const DT T = t.esize == 8 ? DT::F64 : DT::F32;
    const size_t node = (size_t)t.s_pad * k * t.esize;
    std::set<int32_t> touched;
    for (int32_t j : dirty_leaves) touched.insert(j / f);
    int32_t merged = 0;
    const size_t top = t.sizes.size() - 1;
    for (size_t lvl = 1; lvl <= top && !touched.empty(); lvl++) {
        std::vector<int32_t> parents(touched.begin(), touched.end());
        const int cnt = (int)parents.size();
        t.G.grow((size_t)cnt * nc * nc);
        t.lam.grow((size_t)cnt * nc);
        t.Wm.grow((size_t)cnt * nc * k);
        TSVD_HIP(hipMemsetAsync(t.G.p, 0, sizeof(double) * (size_t)cnt * nc * nc, st));
        std::vector<GemmProb> gp;
        for (int i = 0; i < cnt; i++) {
            GemmProb p{};
            p.A = t.level[lvl - 1].p + (size_t)parents[i] * f * node;
            p.lda = t.s_pad;
            p.transA = 1;
            p.B = p.A;
            p.ldb = t.s_pad;
            p.C = t.G.p + (size_t)i * nc * nc;
            p.ldc = nc;
            p.M = nc; p.N = nc; p.K = t.nsrc; p.alpha = 1.0; p.beta = 0.0; p.kchunk = 512;
            gp.push_back(p);
        }
        gemm.run(st, gp, GemmTypes{T, T, DT::F64, DT::F64}, true, "gemm_merge_gram");
        eig.solve(st, t.G.p, nc, cnt, t.lam.p, k);
        const bool root = lvl == top;
        const int mode = !root ? 3 : (transform == TSVD_TRANSFORM_NC ? 1 : 2);
        eig_select(t.Wm.p, t.G.p, t.lam.p, nc, k, cnt, mode, 0, root ? 1e-12 : 0.0, st);
        std::vector<GemmProb> pp;
        for (int i = 0; i < cnt; i++) {
            GemmProb p{};
            p.A = t.level[lvl - 1].p + (size_t)parents[i] * f * node;
            p.lda = t.s_pad;
            p.transA = 0;
            p.B = t.Wm.p + (size_t)i * nc * k;
            p.ldb = nc;
            p.C = root ? (void *)t.E.p : (void *)(t.level[lvl].p + (size_t)parents[i] * node);
            p.ldc = t.s_pad;
            p.M = t.nsrc; p.N = k; p.K = nc; p.alpha = 1.0; p.beta = 0.0;
            pp.push_back(p);
        }
        const bool bf3 = t.merge_bf3 && T == DT::F32;
        gemm.run(st, pp, GemmTypes{T, DT::F64, root ? DT::F64 : T, bf3 ? DT::BF3 : DT::F64}, k > 80,
                 bf3 ? "gemm_merge_bf16x3" : "gemm_merge");
        if (root) {
            std::vector<double> lam((size_t)nc);
            TSVD_HIP(hipMemcpyAsync(lam.data(), t.lam.p, sizeof(double) * nc, hipMemcpyDeviceToHost, st));
            TSVD_HIP(hipStreamSynchronize(st));
            for (int c = 0; c < k; c++) t.root_sigma[c] = std::sqrt(std::max(lam[(size_t)nc - 1 - c], 0.0));
            t.have_root = true;
        }
        merged += cnt;
        std::set<int32_t> up;
        for (int32_t p : parents) up.insert(p / f);
        touched.swap(up);
    }
    TSVD_HIP(hipStreamSynchronize(st));
    return merged;
}

}  // namespace tsvd
