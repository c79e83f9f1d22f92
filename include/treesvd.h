/*
 * treesvd.h -- C ABI of the MI355X-native Tree-SVD update->embedding path.
 *
 * The reference (DoThingYo/Tree-Embedding, directory Code/) is a CPU program
 * whose hot path is reached through C/C++ calls.  Each entry point below names
 * the reference interface it replaces (file:line).  HIP lives behind this
 * header only: no torch or HIP types cross it; every function returns an int
 * status (TSVD_OK = 0, negative = error; tsvd_last_error() gives the text of
 * the calling thread's last error).  Handles own device memory; one handle is
 * not re-entrant, distinct handles are independent.  Unless a function says
 * otherwise it is synchronous: outputs are complete when it returns.
 */
#ifndef TREESVD_H
#define TREESVD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TSVD_OK 0
#define TSVD_EINVAL (-1)   /* bad argument (ids out of range, sizes, modes) */
#define TSVD_ENOMEM (-2)   /* device or host allocation failed            */
#define TSVD_EHIP (-3)     /* a HIP / ROCm library call failed            */
#define TSVD_ESTATE (-4)   /* call out of sequence for this handle        */

#define TSVD_DIR_FWD 0        /* out-edges (undirected: the adjacency)   */
#define TSVD_DIR_TRANSPOSE 1  /* in-edges (directed graphs only)         */

#define TSVD_REFRESH_INIT 0     /* static: r = e_s, p = 0                */
#define TSVD_REFRESH_DYNAMIC 1  /* insertion update over the new slots   */

#define TSVD_FP32 0
#define TSVD_FP64 1

#define TSVD_TRANSFORM_NC 0   /* log10(x/rmax) if x > rmax else 0      */
#define TSVD_TRANSFORM_LP 1   /* log10(1 + x/rmax) on every non-zero   */

const char *tsvd_last_error(void);
int tsvd_device_info(char *name, int name_len, int64_t *hbm_total, int64_t *hbm_free);
/* Device-buffer plumbing for callers that stage inputs in HBM (no reference
 * counterpart).  kind: 0 host->device, 1 device->host, 2 device->device. */
int tsvd_device_alloc(int64_t bytes, void **out);
int tsvd_device_free(void *p);
int tsvd_memcpy(void *dst, const void *src, int64_t bytes, int kind);
/* Per-kernel timing (no reference counterpart; the reference prints phase
 * wall-clock seconds, e.g. Tree_u_NC.cpp:526-563).  When enabled, each library
 * launch records a HIP event pair on the stream it runs on; read returns a JSON
 * object {"kernel": [launches, total_ms], ...}. */
int tsvd_profile_enable(int on);
int tsvd_profile_reset(void);
int tsvd_profile_read(char *json, int len);
/* The project's autotuner for the MFMA products' tile width (no reference
 * counterpart).  While on, the first call of each product kind (label, types,
 * N, K class) whose problems can be run twice (no K split, beta = 0) is timed
 * with the 80- and the 144-column tile and the faster is kept, process-wide,
 * for every later call; off stops tuning new kinds and keeps the choices.
 * read returns {"kind": {"tile_n": .., "ms_tile80": .., "ms_tile144": ..,
 * "calls": ..}, ...}. */
int tsvd_gemm_tuning(int on);
int tsvd_gemm_tuning_read(char *json, int len);
/* Throughput ceiling of the push's memory operation, for its roofline: 64-lane
 * returning FP32 atomic adds on 256-byte rows at hashed row indices over a
 * `rows` x 256 B array, `iters` per wavefront.  Returns wave-level atomic
 * instructions per second (each updates 64 consecutive floats). */
int tsvd_probe_atomic_rate(int64_t rows, int32_t iters, double *wave_atomics_per_s);
/* The same probe with the returning (returning = 1) or non-returning form of
 * the atomic and only the first `lanes` lanes of each wavefront taking part
 * (32 lanes = one 128-byte line): the L2 atomic units are paced by the lines
 * an instruction touches, not by its active lanes. */
int tsvd_probe_atomic_lanes(int64_t rows, int32_t iters, int32_t returning, int32_t lanes,
                            double *wave_atomics_per_s);

/* ------------------------------------------------------------------ graph
 * Replaces UGraph::initializeDynamicUGraph / inputDynamicGraph
 * (Graph_dynamic.h:513-528, 530-605) and Graph::initializeDirectedDynamicGraph
 * / inputDirectedDynamicGraph (Graph_dynamic.h:90-116, 119-238).  A device CSR
 * (out-rows, plus in-rows when directed) with per-row append slack.  Append
 * keeps the reference's row order bit-exactly: each row gains its new
 * neighbours in batch order; undirected edge (u,v) appends v to row u then u
 * to row v.  `former` degrees are snapshotted at every append.            */
typedef struct tsvd_graph tsvd_graph;

int tsvd_graph_create(int32_t n, int directed, double slack, tsvd_graph **out);
int tsvd_graph_destroy(tsvd_graph *g);
/* src/dst: `m` int32 ids; on_device != 0 means they are device pointers. */
int tsvd_graph_append(tsvd_graph *g, const int32_t *src, const int32_t *dst, int64_t m, int on_device);
int64_t tsvd_graph_num_edges(const tsvd_graph *g);
/* A host batch is staged on the device; a batch of more than `edges` edges
 * (default 2^26) frees that staging once the append and the caller's
 * insertion update have read it, so an initial load of a large graph does
 * not keep ~40 B per edge for the small batches that follow.  No reference
 * counterpart (its reader holds the whole edge vector on the host). */
int tsvd_graph_set_staging_limit(tsvd_graph *g, int64_t edges);
/* Host copies of degree / former degree (n int32 each; either may be NULL). */
int tsvd_graph_degrees(const tsvd_graph *g, int dir, int32_t *deg, int32_t *former);
/* Compact host CSR of one direction: rowptr[n+1], col[rowptr[n]]; col may be
 * NULL to query the size through rowptr[n]. */
int tsvd_graph_export_csr(const tsvd_graph *g, int dir, int64_t *rowptr, int32_t *col);
/* The snapshot-file reader of Graph_dynamic.h:127-130 / 539-542 ("from to"
 * pairs); dup_last=1 keeps the reference's duplicate of the last edge of a
 * file that ends in a newline.  With src == NULL only *count is set.       */
int tsvd_read_edge_file(const char *path, int dup_last, int32_t *src, int32_t *dst, int64_t cap, int64_t *count);
/* The drivers' other files.  Subset file (Tree_u_NC.cpp:363-371): integers
 * read until extraction fails; ids may be NULL to count.  config.txt
 * (Tree_u_NC.cpp:338-351): getline lines, NUL-terminated one after another
 * in buf (buf may be NULL; *needed gives the size).  Embedding CSV
 * (Tree_u_NC.cpp:63, 898-909): Eigen's IOFormat(StreamPrecision,
 * DontAlignCols, ", ", "\n") of a row-major rows x cols matrix. */
int tsvd_read_subset_file(const char *path, int32_t *ids, int64_t cap, int64_t *count);
int tsvd_read_lines(const char *path, char *buf, int64_t len, int32_t *count, int64_t *needed);
int tsvd_write_embedding_csv(const char *path, const double *U, int64_t rows, int32_t cols);

/* -------------------------------------------------------------------- PPR
 * Replaces nodegree_DenseUndirected_Refresh_PPR_initialization
 * (ppr_computation_store_dynamic.h:2035-2121), nodegree_DenseDynamicForwardPush
 * (:1394-1498), nodegree_DenseDirected_Refresh_PPR_initialization
 * (:3811-3898) and its _Transpose (:4133-4228), nodegree_DenseDirected-
 * DynamicForwardPush (:2945-3030) and ...Transpose (:3356-3437), each called
 * by the drivers per source range [start, end) (e.g. Tree_u_NC.cpp:529-600).
 * Reserves/residues stay on the device, vertex-major (p[v][s]).            */
typedef struct tsvd_ppr tsvd_ppr;

int tsvd_ppr_create(tsvd_graph *g, const int32_t *sources, int32_t nsrc, double alpha, double rmax,
                    int precision, tsvd_ppr **out);
int tsvd_ppr_destroy(tsvd_ppr *s);
/* mode TSVD_REFRESH_INIT / _DYNAMIC for sources [src_begin, src_end). */
int tsvd_ppr_refresh(tsvd_ppr *s, int mode, int dir, int32_t src_begin, int32_t src_end);
int tsvd_ppr_push(tsvd_ppr *s, int dir, int32_t src_begin, int32_t src_end);
/* The insertion update of the n >= 1e7-node path (the refresh half of
 * All_initialization_and_push_final, ppr:6504-6533 forward, :6675-6704
 * transpose) for sources [src_begin, src_end), after tsvd_graph_append of the
 * same batch (src, dst: host ids, or device ids with on_device).  Edge by
 * edge in batch order the reference does: j = max(former degree, edges of
 * `from` seen so far); j == 0 -> r[from] = 0; else p *= (j+1)/j,
 * r[from] -= p/(j+1)/alpha, r[to] += (1-alpha) p/(j+1)/alpha (dir 1 swaps
 * the ends and uses in-degrees).  Same result, computed edge-parallel; the
 * vertices left with r > rmax are queued for tsvd_ppr_push.               */
int tsvd_ppr_refresh_edges(tsvd_ppr *s, int dir, const int32_t *src, const int32_t *dst, int64_t m, int on_device,
                           int32_t src_begin, int32_t src_end);
/* Host copy of p or r (what = 0: reserve p, 1: residue r) as FP64, source-
 * major [nsrc][n] like the reference's float** pi / residue.              */
int tsvd_ppr_read(tsvd_ppr *s, int dir, int what, double *out);
/* Threshold side of refresh and push: 0 (default) pushes r > rmax only, as the
 * reference does (ppr:1445, 2981, 3390); 1 pushes |r| > rmax, which also
 * drains the negative residues the insertion update leaves, so reserves stay
 * within the push error bound of the exact PPR after dynamic updates.  No
 * reference counterpart. */
int tsvd_ppr_set_two_sided(tsvd_ppr *s, int on);
/* Push rounds replayed from captured hipGraphs, six rounds per replay
 * (on = 1; TSVD_PUSH_GRAPHS=1 at creation sets it), or launched one by one
 * (0, the default, measured 0.4-2.8% faster on the batched and precision
 * workloads).  No counterpart in the reference, whose push is a host loop
 * (ppr:1426-1487); the results are the same either way within the push
 * bound.  tsvd_ppr_graph_stats: replays, captures and failed captures of the
 * handle so far (a failed capture makes the handle launch rounds one by one). */
int tsvd_ppr_set_graphs(tsvd_ppr *s, int on);
int tsvd_ppr_graph_stats(const tsvd_ppr *s, int64_t *launches, int64_t *captures, int64_t *failures);
/* Counters of the last push on `dir`: rounds, popped (vertex, source) pairs,
 * edge relaxations (pairs x degree), and the push kernel time in ms. */
int tsvd_ppr_last_push_stats(tsvd_ppr *s, int dir, int64_t *rounds, int64_t *pops, int64_t *relax,
                             double *kernel_ms);

/* ------------------------------------------------------ frPCA (drop-in)
 * The reference's own types (matrix_vector_functions_intel_mkl.h:20-23,
 * matrix_vector_functions_intel_mkl_ext.h:22-33): col-major FP64 `mat`,
 * 1-based CSR `mat_csr` with pointerB/pointerE.  Define TSVD_NO_MAT_TYPES
 * when the reference's headers already declared them.                     */
#ifndef TSVD_NO_MAT_TYPES
typedef struct { int nrows, ncols; double *d; } mat;
typedef struct { int nnz; int nrows, ncols; double *values; int *cols; int *pointerB, *pointerE; } mat_csr;
#endif
/* Replaces frPCA (frpca.c:239-326, declared matrix_vector_functions_intel_mkl_
 * ext.h:75).  Caller allocates *U (m x k), *S (k x 1), *V (n x k); S comes out
 * ascending as in the reference.  void, like the reference; failures are
 * reported through tsvd_last_error() and leave the outputs zero.          */
void frPCA(mat_csr *A, mat **U, mat **S, mat **V, int k, int q);
/* Same computation with a status code and an explicit sketch seed. */
int tsvd_frpca_csr(const mat_csr *A, mat *U, mat *S, mat *V, int k, int q, uint64_t seed);
/* `count` independent matrices in one grouped device run (item i with sketch
 * seed seed + i): replaces the drivers' serial frPCA loop over dirty blocks
 * (Tree_u_NC.cpp:679-710).  Per item the result equals tsvd_frpca_csr with
 * that seed.  Not re-entrant (one set of solver handles per process). */
int tsvd_frpca_csr_batch(const mat_csr *const *A, mat *const *U, mat *const *S, mat *const *V, int count, int k,
                         int q, uint64_t seed);

/* Replaces LAPACKE_dsyevd (frpca.c:102, eigSVD frpca.c:99-136) and the merge
 * SVDs' small dense problem (dgesvd, matrix_vector_functions_intel_mkl.c:1278,
 * taken through the FP64 Gram): `count` symmetric n x n host matrices
 * (n <= 1024) at stride n*n are solved on the device, eigenvalues ascending
 * into lam (stride n), eigenvectors as the columns of V (column-major, same
 * order).  top in [1, n] writes only the last `top` columns (the largest
 * eigenvalues' vectors) and leaves the others unspecified; top <= 0 writes
 * all.  *info (may be null) is 1 if an iteration limit was reached. */
int tsvd_eig_batched(const double *A, int32_t n, int32_t count, int32_t top, double *V, double *lam, int32_t *info);

/* No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The computation of tsvd_frpca_csr on a device-resident matrix:
 * values/cols/pointerB/pointerE are device arrays in mat_csr's 1-based
 * layout, U (m x k), S (k, ascending) and V (n x k) device arrays
 * (column-major) that the caller owns.  It calls no hipMalloc, hipFree or
 * allocating library routine: every temporary lives in one static device
 * array per GPU (about 45 MB, held until the process ends), so m <= 4096,
 * n <= 32768 and k + 5 <= 136; larger sizes are rejected with TSVD_EINVAL.
 * The matrix is checked on the device before use.  Calls in one process are
 * serialised; each call leaves the array's contents behind, and nothing in
 * it is assumed zero on entry.  With the same seed it draws the same sketch
 * as tsvd_frpca_csr, so the two agree to rounding. */
int tsvd_frpca_csr_noalloc(int32_t m, int32_t n, int32_t nnz, const double *values, const int32_t *cols,
                           const int32_t *pointerB, const int32_t *pointerE, double *U, double *S, double *V,
                           int32_t k, int32_t q, uint64_t seed);

/* ------------------------------------------------------------- pipeline
 * The per-snapshot driver of Tree_u_NC.cpp:507-866, Tree_d_NC.cpp:611-1132,
 * Tree_u_LP.cpp:664-925 (the "plain GPU method"): ingest -> refresh -> push
 * -> block transform + lazy dirty test (ppr:2527-2826) -> frPCA of every
 * dirty block (Tree_u_NC.cpp:75-226) -> dirty-path merges up the fan-out
 * tree (Tree_u_NC.cpp:238-262, 723-850) -> root embedding.                */
typedef struct tsvd_pipeline tsvd_pipeline;

typedef struct {
    int32_t n;            /* vertices                                      */
    int32_t directed;     /* 0 undirected (UGraph), 1 directed (Graph)     */
    int32_t nsrc;         /* |S|                                           */
    const int32_t *sources;
    double alpha, rmax;   /* teleport probability, push threshold          */
    int32_t precision;    /* TSVD_FP32 / TSVD_FP64 for p, r and the matrix */
    int32_t nparts;       /* column blocks                                 */
    int32_t fanout;       /* merge-tree fan-out                            */
    int32_t rank;         /* k (the reference's d)                         */
    int32_t q;            /* frPCA passes (even; 6 = two power iterations) */
    int32_t transform;    /* TSVD_TRANSFORM_NC / _LP                        */
    double delta;         /* lazy-update threshold                         */
    int32_t dyn_start;    /* first snapshot (1-based) with dynamic updates */
    int32_t lazy_start;   /* first snapshot with the dirty test            */
    double slack;         /* CSR append slack fraction                     */
    uint64_t seed;        /* sketch seed                                   */
    int32_t two_sided;    /* 0: push r > rmax (reference), 1: |r| > rmax   */
    int32_t right_embedding; /* LP only: keep V_j = A_j^T U of dirty blocks */
    /* 1: the n >= 1e7-node path of Tree_d_LP.cpp:1377-2024 (directed, LP):
     * reserves and residues kept between snapshots as positive-only pair
     * lists (All_initialization_and_push_final, ppr:6188-6832; negative
     * residues dropped), the insertion update in batch order (first edge of
     * a vertex of former degree 0 clears its residue), blocks from p + p^T
     * with rows >= 30 floor(|S|/30) left empty (ppr:6898-7229).  The state
     * is pushed a batch of source groups at a time. */
    int32_t sparse_state;
    int32_t sparse_groups;   /* groups of 64 sources per batch (0: all of them resident
                              * when they fit in HBM, else as HBM allows)          */
    /* 1: the merge products C W_k of an FP32 pipeline run as bf16x3 on MFMA
     * (each FP32 operand split into bf16 hi + lo, hi*hi + hi*lo + lo*hi,
     * FP32 accumulation, about 16 significant bits); the merge Grams and
     * eigensolves stay FP64.  0: FP64 accumulation. */
    int32_t merge_bf16x3;
} tsvd_config;

typedef struct {
    int32_t snapshot;     /* 1-based index of the snapshot just processed   */
    int32_t ndirty;       /* leaf blocks refactored                         */
    int32_t nmerged;      /* internal tree nodes recomputed (incl. root)    */
    int32_t reused;       /* 1: no dirty block, previous embedding kept     */
    int64_t push_rounds, push_pops, push_relax;
    int64_t push_items;       /* (vertex, 64-source group) work items popped   */
    int64_t push_wave_relax;  /* neighbour visits summed over those items      */
    double ms_total, ms_ingest, ms_refresh, ms_push, ms_blocks, ms_frpca, ms_merge;
    int64_t state_entries;    /* sparse_state: (vertex, value) pairs stored, all lists */
    int32_t state_batches;    /* sparse_state: source batches pushed per snapshot (0: one
                               * resident batch, the state kept in HBM without lists)  */
    int64_t push_line_relax;  /* 128-byte lines the relaxation atomics touched: neighbour
                               * visits x lines of the group holding a pushing lane    */
    int64_t norm_items;       /* (vertex, 64-source group) items whose block-norm sums
                               * were recomputed (the others had unchanged reserves)  */
} tsvd_snapshot_stats;

int tsvd_pipeline_create(const tsvd_config *cfg, tsvd_pipeline **out);
int tsvd_pipeline_destroy(tsvd_pipeline *p);
/* One snapshot: append the batch (host or device ids), update, embed. */
int tsvd_pipeline_snapshot(tsvd_pipeline *p, const int32_t *src, const int32_t *dst, int64_t m, int on_device,
                           tsvd_snapshot_stats *stats);
/* Root embedding (nsrc x rank, row-major FP64): NC U*sqrt|S|, LP U. */
int tsvd_pipeline_embedding(tsvd_pipeline *p, double *E);
/* Root singular values (rank, descending). */
int tsvd_pipeline_root_sigma(tsvd_pipeline *p, double *sigma);
/* Handles owned by the pipeline (valid until it is destroyed). */
tsvd_graph *tsvd_pipeline_graph(tsvd_pipeline *p);
tsvd_ppr *tsvd_pipeline_ppr(tsvd_pipeline *p);
/* The project's autotuner for the push's launch shape (no reference
 * counterpart).  Setting: rows longer than hub_chunk neighbours are split
 * across wavefronts (without a setting from the caller or TSVD_HUB_CHUNK,
 * the first push picks 64 or 128 when the worst-case chunk lists fit in
 * 4 GiB, else keeps 256); lanes whose residue is above ride_frac *
 * rmax (but not above rmax) push along with a popped item (default 0.5; >= 1:
 * off).  A tuned snapshot
 * is tsvd_pipeline_snapshot, except that its push is first timed under each
 * of the ncand candidates from the same saved state (p, r and the pending
 * lists), then run for real with the fastest, which stays for later
 * snapshots.  cand_ms (may be NULL) gets each candidate's push time.  A
 * snapshot that cannot be timed so (static, sparse_state, or too little free
 * HBM to save the state) runs untuned: the handle's own setting, cand_ms -1. */
typedef struct {
    int32_t hub_chunk;
    double ride_frac;
    int32_t items_per_wave;   /* items a wave takes per push round (>= 1; default 8;
                               * fewer in rounds too small to fill the device):
                               * the round grid shrinks by this factor             */
} tsvd_push_setting;
/* Read or set the push setting of a push state (e.g. tsvd_pipeline_ppr's). */
int tsvd_ppr_get_push_setting(const tsvd_ppr *s, tsvd_push_setting *out);
int tsvd_ppr_set_push_setting(tsvd_ppr *s, const tsvd_push_setting *in);
int tsvd_pipeline_snapshot_tuned(tsvd_pipeline *p, const int32_t *src, const int32_t *dst, int64_t m, int on_device,
                                 const tsvd_push_setting *cand, int32_t ncand, tsvd_push_setting *chosen,
                                 double *cand_ms, tsvd_snapshot_stats *stats);
/* Host copy of leaf block j's factored matrix (nsrc x width, col-major FP64)
 * and of the per-block norms ||A||, ||A_new - A_old||, residual. */
int tsvd_pipeline_block_matrix(tsvd_pipeline *p, int32_t j, double *A);
int tsvd_pipeline_block_norms(tsvd_pipeline *p, double *a_norm, double *e_norm, double *resid, int32_t *dirty);
/* Replaces Log_sparse_matrix_entries_with_norm_computation[_LP]
 * (ppr_computation_store_dynamic.h:2649-2732, 2742-2826), which the drivers
 * run per block on threads (Tree_u_NC.cpp:646-666), and with test = 0 the
 * plain Log_sparse_matrix_entries[_LP] (:2527-2561, 2598-2638) used before
 * lazy_start, over blocks [j0, j1): block j's new matrix is the pipeline's
 * transform (NC or LP) of the current reserves (p, or p + p^T when
 * directed); a_norm[j - j0] = ||A_new||_F and e_norm[j - j0] =
 * ||A_new - A_old||_F against the stored matrix; record[j - j0] = -1 when
 * test = 1 and resid_j + e < delta sqrt(2) a (clean, resid_j the block's
 * residual from its last factorisation), else iter (the reference's
 * update_mat_tree_record).  Blocks recorded iter get A_new as their stored
 * matrix, as the reference replaces mat_mapping[i].  Output arrays may be
 * NULL.  Dense-state pipelines only. */
int tsvd_pipeline_block_transform(tsvd_pipeline *p, int32_t j0, int32_t j1, int32_t iter, double delta, int32_t test,
                                  double *a_norm, double *e_norm, int32_t *record);
/* Vertex ids of block j's stored columns (the active columns at its last
 * rebuild, ascending; *count of them; cols may be NULL).  frPCA's sketch of
 * block j has one row per stored column, in this order. */
int tsvd_pipeline_block_columns(tsvd_pipeline *p, int32_t j, int32_t *cols, int32_t *count);
/* Host copy of one tree node's factor U*S (nsrc x rank, col-major FP64);
 * level 0 = leaves. */
int tsvd_pipeline_node(tsvd_pipeline *p, int32_t level, int32_t index, double *out);
/* FP64 Gram A A^T (nsrc x nsrc, row-major) of the current transformed matrix
 * over all n columns, computed on the device.  For accuracy reports against a
 * dense FP64 SVD; the reference has no counterpart (it never forms it).     */
int tsvd_pipeline_dense_gram(tsvd_pipeline *p, double *G);
/* LP right embedding and link scores (Tree_u_LP.cpp:291-404, 1040-1122).
 * With cfg.right_embedding (transform LP), each snapshot computes, for every
 * dirty block j, V_j = A_j^T U with that snapshot's root U, and remembers
 * which snapshot's U it used; a clean block keeps its V_j and its U.  The
 * score of a pair (i, v), i a subset vertex, is U_j[row(i)] . V[v] with U_j
 * the U that block(v)'s V was computed with (the reference's
 * left_matrix_U_cache / left_matrix_index_mapping). */
/* Host copy of V rows [v0, v1) (row-major, rank columns; inactive vertices 0). */
int tsvd_pipeline_right_embedding(tsvd_pipeline *p, int32_t v0, int32_t v1, double *V);
/* scores[e] for pairs (src[e], dst[e]); src must be subset vertices.  With
 * sparse_state the score also carries out_degree(i) * in_degree(v) of the
 * snapshot that computed block(v)'s V, as the n >= 1e7-node path's scores do
 * (ppr:7604-7730). */
int tsvd_pipeline_lp_scores(tsvd_pipeline *p, const int32_t *src, const int32_t *dst, int64_t m, double *scores);
/* Z = A A^T Y (nsrc x l, row-major FP64 in and out) for the same matrix,
 * without forming A A^T: two FP64 products per slab of columns.  Subspace
 * iteration on it gives the exact top singular values where the Gram is too
 * large to form; no reference counterpart.                                */
int tsvd_pipeline_gram_apply(tsvd_pipeline *p, const double *Y, int32_t l, double *Z);
/* Reserves and residues (n values each, FP64; either pointer may be NULL) of
 * subset source i (0 <= i < nsrc) in push direction dir after the last
 * snapshot: the device state, or with sparse_state the stored pair lists
 * (the reference's tripletList / residue_pair_list of that source).        */
int tsvd_pipeline_source_state(tsvd_pipeline *p, int32_t dir, int32_t i, double *reserve, double *residue);

/* Seeded synthetic directed R-MAT graph drawn in HBM (no reference
 * counterpart; for the measured configurations, whose host generation would
 * outlast the run).  Quadrant probabilities a, b, c, 1-a-b-c per level over
 * `scale` levels, ids modulo n, no self-loops, m distinct edges in a seeded
 * random order.  *src and *dst are device arrays of m ids, released with
 * tsvd_device_free. */
int tsvd_gen_rmat(int32_t scale, int32_t n, int64_t m, double a, double b, double c, uint64_t seed, int32_t **src,
                  int32_t **dst);
/* Seeded synthetic directed Chung-Lu graph drawn in HBM (no reference
 * counterpart; same purpose).  u is drawn with probability proportional to
 * (u+1)^(-1/(out_exp-1)), v proportional to (pi(v)+1)^(-1/(in_exp-1)) for a
 * seeded random permutation pi; no self-loops; m distinct edges in a seeded
 * random order.  Outputs as tsvd_gen_rmat. */
int tsvd_gen_chung_lu(int32_t n, int64_t m, double out_exp, double in_exp, uint64_t seed, int32_t **src,
                      int32_t **dst);

/* ---------------------------------------------------------- served form
 * No reference counterpart: the reference's drivers run the whole update once
 * per snapshot file (Tree_u_NC.cpp:507-866).  A server owns one pipeline and a
 * worker thread.  Clients submit insertion batches (requests) from any
 * thread; each step the worker admits the queued requests in submit order,
 * up to max_step_edges (0: half of the free HBM at 64 B per edge) and
 * max_step_requests (0: no cap; 1: the plain per-snapshot method), appends
 * them as one snapshot and completes them with that snapshot's embedding.
 * Requests arriving during a step wait for the next one.  Steps overlap on
 * the device: the frPCA, merges and root of step t run on a second stream
 * while step t+1 appends, updates and pushes; step t+1's dirty test waits for
 * step t (no_overlap = 1, or right_embedding, runs steps one at a time).   */
typedef struct tsvd_server tsvd_server;

typedef struct {
    int64_t max_step_edges;
    int32_t max_step_requests;
    int32_t no_overlap;
} tsvd_server_config;

typedef struct {
    int64_t ticket;
    int32_t snapshot;        /* pipeline snapshot that applied the request      */
    int32_t step_requests;   /* requests applied by that snapshot               */
    int64_t step_edges;
    int64_t last_ticket;     /* state version: every ticket <= this is applied  */
    double submit_ms, start_ms, done_ms, latency_ms;   /* server clock, ms      */
    tsvd_snapshot_stats step;
} tsvd_request_result;

int tsvd_server_create(const tsvd_config *cfg, const tsvd_server_config *scfg, tsvd_server **out);
/* Completes every queued request, then frees the server and its pipeline. */
int tsvd_server_destroy(tsvd_server *s);
/* Copies the batch (host ids, validated here) and returns its ticket. */
int tsvd_server_submit(tsvd_server *s, const int32_t *src, const int32_t *dst, int64_t m, int64_t *ticket);
/* Blocks until the ticket's step is done; E (nsrc x rank row-major, may be
 * NULL) receives that step's embedding.  Each ticket is collected once.   */
int tsvd_server_wait(tsvd_server *s, int64_t ticket, double *E, tsvd_request_result *res);
/* The server's pipeline, for read-only accessors while no request is queued. */
tsvd_pipeline *tsvd_server_pipeline(tsvd_server *s);
int64_t tsvd_server_edge_budget(const tsvd_server *s);

#ifdef __cplusplus
}
#endif
#endif
