// tsvd_pipeline.h -- column blocks, blockwise frPCA, merge tree, driver.
#pragma once
#include "tsvd_internal.h"
#include "tsvd_graph.h"
#include "tsvd_ppr.h"
#include "tsvd_gemm.h"
#include "tsvd_sparse.h"


#include <map>
#include <unordered_map>

namespace tsvd {

// Batched FP64 symmetric eigensolver (eig.hip): `count` n x n matrices at
// stride n*n (column-major, n <= EIG_NMAX); eigenvalues ascending into lam
// (stride n, all n), eigenvectors in the same order over the input, of which
// only the last `top` columns are written (top <= 0: all).  The workspace of
// eig_workspace_bytes(n, count) bytes is the caller's; GEMM descriptors go
// through desc_dev (cap slots) when given, else the runner's own buffer.
// info_host, when given, receives 1 if a QR segment did not converge (it
// synchronizes the stream).
constexpr int EIG_NMAX = 1024;
constexpr size_t eig_workspace_bytes(int n, int count)
{
    return 1024 + (size_t)(count > 0 ? count : 0) *
                      ((4 * (size_t)n * (size_t)n + 5 * (size_t)n) * sizeof(double) + sizeof(int));
}
void eig_batched(hipStream_t st, double *A, int n, int count, double *lam, int top, void *ws, size_t ws_bytes,
                 GemmRunner &gemm, GemmProb *desc_dev = nullptr, int desc_cap = 0, int *info_host = nullptr);

class Eig {
public:
    // eig_batched with a workspace that grows as needed
    void solve(hipStream_t st, double *A, int n, int count, double *w, int top = 0);

private:
    DBuf<uint8_t> ws_;
    GemmRunner gemm_;
};

// Column blocks of the |S| x n matrix (Tree_u_NC.cpp:396-432): common = n /
// nparts, block of v = min(v / common, nparts-1), the last block takes the
// remainder.  Block j's factored matrix is kept dense over its active columns
// (vertices with an edge, or sources), column-major with leading dimension
// s_pad, at A + start_j * s_pad.
struct Blocks {
    int32_t nparts = 0, n = 0, nsrc = 0, s_pad = 0, common = 0;
    int precision = TSVD_FP32, transform = TSVD_TRANSFORM_NC;
    size_t esize = 4;
    double rmax = 1e-7;
    std::vector<int64_t> start, width;
    std::vector<int32_t> ncols;             // active columns stored per block
    DBuf<uint8_t> A;                        // n * s_pad * esize
    DBuf<int32_t> slot, colmap, d_ncols;    // n, n, nparts
    DBuf<uint8_t> active, is_source;        // n
    DBuf<double> norms;                     // 2 * nparts: ||A_new||^2, ||A_new - A_old||^2
    DBuf<int32_t> d_dirty;                  // dirty block ids
    DBuf<int32_t> row_of;                   // s_pad: subset row of push lane s (a permutation of the
                                            // sources); stored blocks keep lane order, the full-matrix
                                            // slabs of transform_chunk are in subset-row order
    std::vector<double> a2, e2, resid;
    // Incremental norm pass: per (vertex, source group) item, the two sums of
    // its 64 lanes (||A_new||^2 part, ||A_new - A_old||^2 part), kept current
    // between passes; only items whose reserves changed are recomputed.
    DBuf<double> item_sums;                 // 2 * n * (s_pad / 64), made on first use
    DBuf<unsigned long long> item_done;     // items recomputed by the last pass (sharded)
    int64_t last_items = 0;
};

// Sparse leaf products (spmm.hip).  One dirty block factored from CSR / CSC
// copies of its stored matrix; its Q (m x l) and Qt (nc x l) are row-major.
struct SpItem {
    const void *A;     // stored block, m x nc column-major, leading dimension lda
    int64_t lda;
    int32_t m, nc;
    int64_t base;      // first of the item's m row counters, then its nc column counters
    void *Q, *Qt;
};
struct SparseLeaves {
    std::vector<SpItem> items;
    std::vector<SpItem> cand;       // candidates of the last density pass, with their count runs
    DBuf<SpItem> ditems;
    DBuf<unsigned long long> nnz;   // per item
    DBuf<int64_t> cnt, off;         // counts and their exclusive scan (one run per row, then per column)
    DBuf<int32_t> idx;              // CSR column ids / CSC row ids
    DBuf<int32_t> part;             // row counts per column range (item, range, row)
    DBuf<uint8_t> val, tmp;
    int64_t counts = 0, entries = 0;
};
// Non-zero share of each candidate block (items need A, lda, m, nc).
// Sets each candidate's `base` (its run of row then column counts) and keeps
// the column counts for leaf_sparse_build.
std::vector<double> leaf_densities(SparseLeaves &s, std::vector<SpItem> &cand, int precision, hipStream_t st);
// CSR and CSC of the listed blocks (their Q / Qt pointers are used by leaf_spmm).
void leaf_sparse_build(SparseLeaves &s, std::vector<SpItem> items, int precision, hipStream_t st);
// Q = A Qt (transposed = false) or Qt = A^T Q for every built item.
void leaf_spmm(SparseLeaves &s, int precision, int32_t l, bool transposed, hipStream_t st);

// frPCA workspace for all blocks (rank k, l = k + 5).
struct FrpcaWork {
    int32_t k = 0, l = 0, q = 6;
    DBuf<uint8_t> Qt;       // n * l * esize   (block j at start_j * l, ld = width_j)
    DBuf<uint8_t> Q, Q2;    // nparts * s_pad * l * esize
    DBuf<double> G, W, lam; // nparts * l * l, nparts * l * l, nparts * l
    SparseLeaves sp;        // sparse dirty blocks of the current factorisation
    double sparse_density = 0.08;   // blocks with fewer non-zeros than this share use the sparse products
    int64_t last_sparse = 0;        // dirty blocks factored sparse in the last call
};

// Merge tree: level 0 = leaves.  Level sizes from the leaves up; parent i of
// a level holds children [i*f, min((i+1)*f, size)) of the level below (the
// reference's heap tree when nparts is a power of f, ppr:225-245).  Node
// factors U*S are s_pad x k column-major, consecutive within a level, so a
// parent's children form one s_pad x (f*k) matrix.
struct Tree {
    int32_t fanout = 4, k = 0, s_pad = 0, nsrc = 0;
    std::vector<int32_t> sizes;         // nodes per level
    std::vector<int32_t> padded;        // sizes rounded up to a multiple of f
    std::vector<DBuf<uint8_t>> level;   // padded[l] * s_pad * k * esize
    DBuf<double> G, lam, Wm;            // Gram / eigen scratch per merge batch
    DBuf<double> E;                     // root embedding s_pad x k
    std::vector<double> root_sigma;
    size_t esize = 4;
    bool have_root = false;
    bool merge_bf3 = false;             // FP32 levels: merge products as bf16x3
};

void blocks_setup(Blocks &b, int32_t n, int32_t nparts, int32_t nsrc, int32_t s_pad, int precision, int transform,
                  double rmax, const std::vector<int32_t> &sources, hipStream_t st);
// ||A_new||^2 and ||A_new - A_old||^2 of every block (p0 [+ p1] transformed).
void blocks_norms(Blocks &b, const void *p0, const void *p1, const tsvd_graph *g, hipStream_t st);
// The same for blocks [j0, j1) only (the other blocks' norms are left as they are).
void blocks_norms_range(Blocks &b, const void *p0, const void *p1, const tsvd_graph *g, int32_t j0, int32_t j1,
                        hipStream_t st);
// The same for blocks [j0, j1), recomputing only the items flagged in
// `changed` (n * (s_pad / 64) bytes, cleared here), or every item when `all`
// or when no item sums exist yet.  Items of other blocks are brought up to
// date too, but only [j0, j1) of a2/e2 are written.
void blocks_norms_items(Blocks &b, const void *p0, const void *p1, const tsvd_graph *g, uint8_t *changed, bool all,
                        int32_t j0, int32_t j1, hipStream_t st);
// Rebuild the stored matrix of the listed blocks from p0 [+ p1].
void blocks_rebuild(Blocks &b, const std::vector<int32_t> &dirty, const void *p0, const void *p1, hipStream_t st);

// The same steps for the n >= 1e7-node path, one batch of sources at a time
// (FP32 forward/transpose reserves pf, pt of width sp = rows [row0, row0 +
// rows) of the matrix; rows >= keep excluded):
void blocks_active(Blocks &b, const tsvd_graph *g, hipStream_t st);   // active columns; norms zeroed
void blocks_norms_batch(Blocks &b, const float *pf, const float *pt, int32_t sp, int32_t row0, int32_t rows,
                        int32_t keep, hipStream_t st);
void blocks_norms_read(Blocks &b, hipStream_t st);
void blocks_slots(Blocks &b, const std::vector<int32_t> &dirty, hipStream_t st);
void blocks_gather_batch(Blocks &b, const std::vector<int32_t> &dirty, const float *pf, const float *pt, int32_t sp,
                         int32_t row0, int32_t rows, int32_t keep, hipStream_t st);
void blocks_ncols_read(Blocks &b, hipStream_t st);

void frpca_setup(FrpcaWork &w, const Blocks &b, int32_t k, int32_t q);
// Factor the listed blocks; leaf factors U*S go to `leaves` (s_pad x k per
// block, block j at j * s_pad * k).  Returns sum of the kept sigma^2 per block
// in sig2 (host, indexed like `dirty`).
void frpca_blocks(FrpcaWork &w, Blocks &b, const std::vector<int32_t> &dirty, void *leaves, GemmRunner &gemm,
                  Eig &eig, uint64_t seed, hipStream_t st, std::vector<double> &sig2);

void tree_setup(Tree &t, int32_t nparts, int32_t fanout, int32_t k, int32_t nsrc, int32_t s_pad, size_t esize,
                hipStream_t st);
// Recompute every ancestor of the dirty leaves, level by level, and the root
// embedding (mode NC: U*sqrt|S|, LP: U).  Returns the number of merged nodes.
int32_t tree_update(Tree &t, const std::vector<int32_t> &dirty_leaves, int transform, GemmRunner &gemm, Eig &eig,
                    hipStream_t st);

// Elementwise helpers (blocks.hip)
// Standard normal sketches: entry (i, c) of a rows x cols matrix is number
// c * rows + i of its seed, stored column-major (leading dimension ld) or
// row-major (entry (i, c) at X[i * cols + c]); one launch per GAUSS_BATCH.
struct GaussItem {
    void *X;
    int64_t rows, cols, ld;
    uint64_t seed;
    int32_t rowmajor;
};
constexpr int GAUSS_BATCH = 64;
struct GaussBatch {
    GaussItem it[GAUSS_BATCH];
};
void gaussian_fill(const std::vector<GaussItem> &items, int precision, hipStream_t st);
// W[:, c] = V[:, c] * scale(c) with scale = 1/sqrt(lam) (zeroed below the
// guard) for the orthonormalising rotation; batch of `count` l x l matrices.
void eig_rotation(double *W, const double *V, const double *lam, int l, int count, double eps_abs, double eps_rel,
                  hipStream_t st);
// Top-k selection: W[:, c] = V[:, src(c)] * f(lam[src(c)]) for c < k, with
// src(c) = l-1-c (descending) or l-k+c (ascending, the reference's frPCA
// order) and, for sigma = sqrt(lam), f = sigma (mode 0, U*S), sigma^-1/2
// (mode 1, NC root U*sqrt(S)), 1/sigma (mode 2, U), 1 (mode 3); f = 0 when
// sigma <= eps (modes 0-2).  Matrices at stride l*l (V) and l*k (W).
void eig_select(double *W, const double *V, const double *lam, int l, int k, int count, int mode, int ascending,
                double eps, hipStream_t st);

// LP link scores: scores[e] = sum_c U_j[c * s_pad + rows[e]] * V_j(slot(v), c)
// for v = dst[e] in block j (0 when v is not a stored column or j has no V);
// with dblk (the n >= 1e7-node path), times dblk[j][rows[e]] * vdeg[v]: the
// out-degree of the source and the in-degree of v when V_j was computed.
void lp_scores(const Blocks &b, const double *V, const double *const *ublk, const double *const *dblk,
               const double *vdeg, int32_t k, const int32_t *rows,
               const int32_t *dst, int64_t m, double *scores, hipStream_t st);

// FP64 vertex-major slab [v0, v0+nv) of the transformed |S| x n matrix
// (X[(v - v0) * s_pad + s]; padding rows zero).
void transform_chunk(const Blocks &b, const void *p0, const void *p1, double *X, int64_t v0, int64_t nv,
                     hipStream_t st);

// One column block (or one drop-in matrix) for the grouped frPCA core.
struct FrpcaItem {
    const void *A;    // m x nc, column-major, leading dimension lda
    int64_t lda;
    int32_t m, nc;
    void *Qt;         // nc x l, leading dimension ldqt
    int64_t ldqt;
    void *Q, *Q2;     // m x l each, leading dimension ldq
    int64_t ldq;
    uint64_t seed;
    int32_t sp = -1;  // >= 0: item of the SparseLeaves set; Q, Q2 and Qt are then row-major (ld l)
};
// Runs the passes of frPCA (frpca.c:239-316) for all items at once.  On
// return Q holds the final orthonormal basis, Qt = A^T Q, and G / lam (item i
// at i*l*l / i*l) the eigenvectors and ascending eigenvalues of Qt^T Qt.
void frpca_core(std::vector<FrpcaItem> &items, int precision, int32_t l, int32_t q, double *G, double *W,
                double *lam, GemmRunner &gemm, Eig &eig, hipStream_t st, SparseLeaves *sp = nullptr);

}  // namespace tsvd

// LP right embedding: V (n x k FP64; block j's rows in stored-column order
// at V + start_j * k, leading dimension width_j) and the root U each block's
// V was computed with (one FP64 s_pad x k copy per snapshot still in use).
struct RightEmbedding {
    tsvd::DBuf<double> V;
    std::vector<int32_t> ver;                    // per block: snapshot of its U, -1 none
    std::map<int32_t, tsvd::DBuf<double>> U;     // snapshot -> U
    tsvd::DBuf<const double *> ublk;             // per block: device pointer of its U (or null)
    // n >= 1e7-node path: scores carry the degrees of the snapshot that
    // computed them (ppr:7707, 7727): out-degrees of the sources per
    // snapshot, and each vertex's in-degree when its block's V was computed
    std::map<int32_t, tsvd::DBuf<double>> dsrc;
    tsvd::DBuf<const double *> dblk;
    tsvd::DBuf<double> vdeg;
    std::unordered_map<int32_t, int32_t> row_of; // subset vertex -> row
};

namespace tsvd {
// One snapshot in three parts, so that a server can run the back part of
// step t (frPCA, merges, root: stream `st`) while the front part of step t+1
// (ingest, refresh, push, block norms) runs on the graph's stream.  The
// decide part (dirty test, rebuild of the dirty blocks) must follow the
// previous back part: it reads the residuals that part writes and rewrites
// the blocks it factors.
struct StepPlan {
    int32_t it = 0;
    bool skipped = false;                  // only ingested (the lazy_start rule)
    std::vector<int32_t> rebuild, factor;  // dirty blocks, and those with a nonzero matrix
    std::vector<double> a2;                // ||A_new||^2 per block at this step
    tsvd_snapshot_stats S{};
    double t0 = 0.0;
};
struct PushTune {
    const tsvd_push_setting *cand = nullptr;
    int32_t ncand = 0;
    double *cand_ms = nullptr;   // per candidate, -1 when not timed
};
bool pipeline_front(tsvd_pipeline *p, const int32_t *src, const int32_t *dst, int64_t m, int on_device,
                    StepPlan &plan, const PushTune *tune = nullptr);
void pipeline_decide(tsvd_pipeline *p, StepPlan &plan);
void pipeline_back(tsvd_pipeline *p, StepPlan &plan, hipStream_t st);
}  // namespace tsvd

struct tsvd_pipeline {
    tsvd_config cfg{};
    std::vector<int32_t> sources;
    RightEmbedding re;
    tsvd_graph *g = nullptr;
    tsvd_ppr *ppr = nullptr;
    tsvd::Blocks blocks;
    tsvd::FrpcaWork fw;
    tsvd::Tree tree;
    tsvd::GemmRunner gemm;
    tsvd::Eig *eig = nullptr;
    tsvd::SparseState sp;                    // sparse_state pipelines only
    bool order_set = false;                  // push lanes ordered (at the first static snapshot)
    std::vector<int32_t> lane_of;            // subset row -> push lane (empty: the same)
    // Stored blocks, tree nodes and the root keep push-lane row order; this is
    // the device row that holds subset row i.
    int32_t lane_row(int32_t i) const { return lane_of.empty() ? i : lane_of[(size_t)i]; }
    int32_t snapshot = 0;
    std::vector<int32_t> last_dirty;
    ~tsvd_pipeline();
};
