/*
 * treesvd_oracle.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * A plain-C, single-address-space CPU restatement of the reference's graph
 * ingest, dynamic PPR update and FIFO forward push.  It exists to check the
 * HIP product path and to time the reference algorithm on host cores.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it;
 * the product library never links or calls it.
 *
 * Every function cites the reference routine it restates (paths relative to
 * the reference's Code/ directory).  The reference keeps residues/reserves in
 * `float` and does its arithmetic in `double`; `fp32=1` reproduces that by
 * rounding every stored p/r value to float (ST()), `fp32=0` keeps FP64.
 *
 * Pinning (see tests/test_oracle_*.py):
 *   ingest        -- golden CSR rows written by the reference's own
 *                    Graph_dynamic.h (oracle/_ref) and committed fixtures;
 *   refresh+push  -- the PPR invariant  pi_s = p_s + sum_v r_s(v) pi_v  against
 *                    an exact FP64 linear solve, the termination condition,
 *                    and hand-computed single-step vectors;
 *   file reader   -- the reference's duplicate-last-edge behaviour (oracle/_ref);
 *   edge-order update (the n >= 1e7-node path) -- equal to the row update and
 *                    the exact invariant where no vertex gets its first edge;
 *                    with a first edge, the invariant's defect is exactly the
 *                    dropped residue times pi_w, in batch order;
 *   negative drop -- keeps exactly the positive entries.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

#define ORC_EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------ graph */

typedef struct {
    int n, directed, ndir;      /* ndir: 1 undirected, 2 directed (0=out,1=in) */
    int *deg[2], *former[2], *cap[2];
    int **adj[2];
    long long m;
} orc_graph;

ORC_EXPORT orc_graph *orc_graph_new(int n, int directed)
{
    /* Graph::initializeDirectedDynamicGraph (Graph_dynamic.h:90-116),
     * UGraph::initializeDynamicUGraph (Graph_dynamic.h:513-528).  Row arrays
     * start empty (the reference leaves the pointer table uninitialised and
     * compares it to NULL; here it is zero-initialised). */
    orc_graph *g = (orc_graph *)calloc(1, sizeof(orc_graph));
    if (!g) return NULL;
    g->n = n; g->directed = directed ? 1 : 0; g->ndir = directed ? 2 : 1;
    for (int d = 0; d < g->ndir; d++) {
        g->deg[d] = (int *)calloc((size_t)n, sizeof(int));
        g->former[d] = (int *)calloc((size_t)n, sizeof(int));
        g->cap[d] = (int *)calloc((size_t)n, sizeof(int));
        g->adj[d] = (int **)calloc((size_t)n, sizeof(int *));
    }
    return g;
}

ORC_EXPORT void orc_graph_free(orc_graph *g)
{
    if (!g) return;
    for (int d = 0; d < g->ndir; d++) {
        for (int i = 0; i < g->n; i++) free(g->adj[d][i]);
        free(g->adj[d]); free(g->deg[d]); free(g->former[d]); free(g->cap[d]);
    }
    free(g);
}

static void row_push(orc_graph *g, int d, int v, int x)
{
    if (g->deg[d][v] == g->cap[d][v]) {
        int nc = g->cap[d][v] ? 2 * g->cap[d][v] : 4;
        g->adj[d][v] = (int *)realloc(g->adj[d][v], sizeof(int) * (size_t)nc);
        g->cap[d][v] = nc;
    }
    g->adj[d][v][g->deg[d][v]++] = x;
}

/* UGraph::inputDynamicGraph (Graph_dynamic.h:530-605) and
 * Graph::inputDirectedDynamicGraph (Graph_dynamic.h:119-238), minus the file
 * parse: former_degree <- degree (:545-547 / :135-138), then every edge is
 * appended in batch order -- undirected: (u->v) to row u, then (v->u) to row v
 * (:591-600); directed: to out-row u and in-row v (:219-232).
 * Returns 0, or -1 if an id is out of range (the reference does not check). */
ORC_EXPORT int orc_graph_append(orc_graph *g, const int *src, const int *dst, long long m)
{
    for (long long e = 0; e < m; e++)
        if (src[e] < 0 || src[e] >= g->n || dst[e] < 0 || dst[e] >= g->n) return -1;
    for (int d = 0; d < g->ndir; d++)
        memcpy(g->former[d], g->deg[d], sizeof(int) * (size_t)g->n);
    for (long long e = 0; e < m; e++) {
        int u = src[e], v = dst[e];
        if (g->directed) { row_push(g, 0, u, v); row_push(g, 1, v, u); }
        else { row_push(g, 0, u, v); row_push(g, 0, v, u); }
        g->m++;
    }
    return 0;
}

ORC_EXPORT void orc_graph_degrees(const orc_graph *g, int d, int *deg, int *former)
{
    if (deg) memcpy(deg, g->deg[d], sizeof(int) * (size_t)g->n);
    if (former) memcpy(former, g->former[d], sizeof(int) * (size_t)g->n);
}

/* Export direction d as CSR: rowptr[n+1] (exclusive scan of degrees), col[]. */
ORC_EXPORT long long orc_graph_csr(const orc_graph *g, int d, long long *rowptr, int *col)
{
    long long off = 0;
    for (int i = 0; i < g->n; i++) {
        if (rowptr) rowptr[i] = off;
        if (col && g->deg[d][i]) memcpy(col + off, g->adj[d][i], sizeof(int) * (size_t)g->deg[d][i]);
        off += g->deg[d][i];
    }
    if (rowptr) rowptr[g->n] = off;
    return off;
}

/* ---------------------------------------------------------- edge-file read */

/* One `infile >> x` on an int (libstdc++ semantics): skip whitespace; EOF
 * before a digit -> fail (value untouched); digits up to EOF -> success with
 * eofbit.  Returns 0 ok, 1 ok+eof, 2 fail. */
static int read_int(FILE *f, int *out)
{
    int c;
    do { c = fgetc(f); } while (c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\v' || c == '\f');
    if (c == EOF) return 2;
    int neg = 0;
    if (c == '-' || c == '+') { neg = (c == '-'); c = fgetc(f); }
    if (c < '0' || c > '9') { if (c != EOF) ungetc(c, f); return 2; }
    long long v = 0;
    while (c >= '0' && c <= '9') { v = v * 10 + (c - '0'); c = fgetc(f); }
    *out = (int)(neg ? -v : v);
    if (c == EOF) return 1;
    ungetc(c, f);
    return 0;
}

/* The snapshot-file loop `while (infile.good()) { infile >> from >> to;
 * edge_vec.push_back(...); }` (Graph_dynamic.h:127-130 and :539-542).  A file
 * that ends in a newline yields a duplicate of its last edge; with
 * dup_last=0 that duplicate (and any pair read after a failed extraction) is
 * dropped.  Two calls: with src==NULL it only counts.  Returns the edge
 * count, -1 if the file cannot be opened, -2 if the first read fails (the
 * reference would push an uninitialised pair there). */
ORC_EXPORT long long orc_read_edge_file(const char *path, int dup_last, int *src, int *dst, long long cap)
{
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    long long cnt = 0;
    int a = 0, b = 0, first = 1;
    for (;;) {
        int ra = read_int(f, &a);
        int rb = 2;
        if (ra == 0) rb = read_int(f, &b);
        int failed = (ra == 2) || (ra == 0 && rb == 2) || ra == 1;
        if (failed && first) { fclose(f); return -2; }
        if (!failed || dup_last) {
            if (src && cnt < cap) { src[cnt] = a; dst[cnt] = b; }
            cnt++;
        }
        first = 0;
        if (failed || rb == 1) break;
    }
    fclose(f);
    return cnt;
}

/* -------------------------------------------------------------- PPR state */

typedef struct {
    int n, nsrc, ndir, fp32, quirk;
    int two;                    /* 1: push |r| > rmax (negative residues too) */
    double alpha, rmax;
    int *src;
    double *p[2], *r[2];        /* [nsrc * n], row k = source k */
    double *M;                  /* [nsrc * n] FP64 matrix accumulator (dense_mat_mapping) */
    unsigned char *flag[2];     /* [nsrc * n] */
    int *q[2];                  /* per-source ring, capacity n + 1 */
    long long *qh[2], *qt[2];
    long long pops, relax;      /* pushes with r > rmax and their edge relaxations */
} orc_ppr;

#define ST(x) (s->fp32 ? (double)(float)(x) : (x))
/* Residue beyond the push threshold: r > rmax (the reference, positive side
 * only), or |r| > rmax with two-sided pushing. */
#define OUT(x) ((x) > rmax || (s->two && (x) < -rmax))
#define IDX(k, v) ((size_t)(k) * (size_t)s->n + (size_t)(v))

ORC_EXPORT orc_ppr *orc_ppr_new(int n, int nsrc, const int *src, int directed,
                                double alpha, double rmax, int fp32, int quirk_int_ratio)
{
    orc_ppr *s = (orc_ppr *)calloc(1, sizeof(orc_ppr));
    if (!s) return NULL;
    s->n = n; s->nsrc = nsrc; s->ndir = directed ? 2 : 1; s->fp32 = fp32; s->quirk = quirk_int_ratio;
    s->alpha = alpha; s->rmax = rmax;
    s->src = (int *)malloc(sizeof(int) * (size_t)nsrc);
    memcpy(s->src, src, sizeof(int) * (size_t)nsrc);
    size_t N = (size_t)nsrc * (size_t)n;
    s->M = (double *)calloc(N, sizeof(double));
    for (int d = 0; d < s->ndir; d++) {
        s->p[d] = (double *)calloc(N, sizeof(double));
        s->r[d] = (double *)calloc(N, sizeof(double));
        s->flag[d] = (unsigned char *)calloc(N, 1);
        s->q[d] = (int *)malloc(sizeof(int) * (size_t)nsrc * (size_t)(n + 1));
        s->qh[d] = (long long *)calloc((size_t)nsrc, sizeof(long long));
        s->qt[d] = (long long *)calloc((size_t)nsrc, sizeof(long long));
        if (!s->p[d] || !s->r[d] || !s->flag[d] || !s->q[d]) return NULL;
    }
    return s;
}

ORC_EXPORT void orc_ppr_free(orc_ppr *s)
{
    if (!s) return;
    for (int d = 0; d < s->ndir; d++) {
        free(s->p[d]); free(s->r[d]); free(s->flag[d]); free(s->q[d]); free(s->qh[d]); free(s->qt[d]);
    }
    free(s->M); free(s->src); free(s);
}

/* dir: 0 = forward (out-edges / undirected), 1 = transpose (in-edges) */
ORC_EXPORT double *orc_ppr_p(orc_ppr *s, int dir) { return s->p[dir]; }
ORC_EXPORT double *orc_ppr_r(orc_ppr *s, int dir) { return s->r[dir]; }
ORC_EXPORT double *orc_ppr_M(orc_ppr *s) { return s->M; }

/* Queue of my_queue.h:8-101 (FIFO of source k); the ring never holds a vertex
 * twice because the flag guards every enqueue, so n + 1 slots suffice. */
static inline void q_push(orc_ppr *s, int d, int k, int v)
{
    int *q = s->q[d] + (size_t)k * (size_t)(s->n + 1);
    q[s->qt[d][k]] = v;
    s->qt[d][k] = (s->qt[d][k] + 1) % (s->n + 1);
}

typedef struct {
    orc_ppr *s; const orc_graph *g; int dir, k0, k1;
    const int *es, *ed; long long m;      /* edge batch (refresh_edges only) */
} job_t;

static void run_parallel_edges(void *(*fn)(void *), orc_ppr *s, const orc_graph *g, int dir, int nthreads,
                               const int *es, const int *ed, long long m)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > s->nsrc) nthreads = s->nsrc > 0 ? s->nsrc : 1;
    pthread_t th[256];
    job_t jobs[256];
    if (nthreads > 256) nthreads = 256;
    /* source ranges as in the drivers' fork loop (Tree_u_NC.cpp:529-557) */
    for (int t = 0; t < nthreads; t++) {
        jobs[t].s = s; jobs[t].g = g; jobs[t].dir = dir;
        jobs[t].es = es; jobs[t].ed = ed; jobs[t].m = m;
        jobs[t].k0 = t * (s->nsrc / nthreads);
        jobs[t].k1 = (t == nthreads - 1) ? s->nsrc : (t + 1) * (s->nsrc / nthreads);
    }
    if (nthreads == 1) { fn(&jobs[0]); return; }
    for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, fn, &jobs[t]);
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
}

static void run_parallel(void *(*fn)(void *), orc_ppr *s, const orc_graph *g, int dir, int nthreads)
{
    run_parallel_edges(fn, s, g, dir, nthreads, NULL, NULL, 0);
}

/* Static branch of nodegree_DenseUndirected_Refresh_PPR_initialization
 * (ppr_computation_store_dynamic.h:2056-2068) and of the directed forms
 * (:3832-3845, :4154-4167): zero r, p, flags, queue; r[src] = 1; enqueue src.
 * The matrix accumulator is NOT cleared, as in the reference. */
ORC_EXPORT void orc_ppr_init_static(orc_ppr *s, int dir)
{
    for (int k = 0; k < s->nsrc; k++) {
        memset(s->r[dir] + IDX(k, 0), 0, sizeof(double) * (size_t)s->n);
        memset(s->p[dir] + IDX(k, 0), 0, sizeof(double) * (size_t)s->n);
        memset(s->flag[dir] + IDX(k, 0), 0, (size_t)s->n);
        s->qh[dir][k] = s->qt[dir][k] = 0;
        int v = s->src[k];
        s->r[dir][IDX(k, v)] = 1.0;
        s->flag[dir][IDX(k, v)] = 1;
        q_push(s, dir, k, v);
    }
}

/* Dynamic branch of nodegree_DenseUndirected_Refresh_PPR_initialization
 * (ppr_computation_store_dynamic.h:2072-2118), nodegree_DenseDirected_Refresh_
 * PPR_initialization (:3848-3893) and its _Transpose (:4170-4223): for every
 * vertex in id order, every new slot j in [former_deg, deg), every source k:
 *   j == 0 -> skip;  M += p/j;  p *= (j+1)/j;  r[from] -= p/(j+1)/alpha;
 *   r[to] += (1-alpha) p/(j+1)/alpha;  enqueue from / to if r > rmax.
 * quirk_int_ratio=1 keeps the reference's integer (j+1)/j. */
static void *refresh_job(void *arg)
{
    job_t *J = (job_t *)arg;
    orc_ppr *s = J->s; const orc_graph *g = J->g; int d = J->dir;
    double a = s->alpha, rmax = s->rmax;
    for (int i = 0; i < g->n; i++) {
        for (int j = g->former[d][i]; j < g->deg[d][i]; j++) {
            int to = g->adj[d][i][j];
            for (int k = J->k0; k < J->k1; k++) {
                if (j == 0) continue;
                double *p = s->p[d] + IDX(k, 0), *r = s->r[d] + IDX(k, 0);
                unsigned char *fl = s->flag[d] + IDX(k, 0);
                s->M[IDX(k, i)] += p[i] * 1 / j;
                if (s->quirk) p[i] = ST(p[i] * (double)((j + 1) / j));
                else p[i] = ST(p[i] * ((double)(j + 1) / (double)j));
                r[i] = ST(r[i] - p[i] / (j + 1) / a);
                r[to] = ST(r[to] + (1 - a) * p[i] / (j + 1) / a);
                if (OUT(r[i]) && !fl[i]) { q_push(s, d, k, i); fl[i] = 1; }
                if (OUT(r[to]) && !fl[to]) { q_push(s, d, k, to); fl[to] = 1; }
            }
        }
    }
    return NULL;
}

ORC_EXPORT void orc_ppr_refresh_dynamic(orc_ppr *s, const orc_graph *g, int dir, int nthreads)
{
    run_parallel(refresh_job, s, g, dir, nthreads);
}

/* Dynamic branch of All_initialization_and_push_final, the n >= 1e7-node path
 * (ppr_computation_store_dynamic.h:6504-6533 forward, :6675-6704 transpose):
 * for every source, every edge of the batch in batch order, (from, to) =
 * (src, dst) forward or (dst, src) transpose, with degrees of that direction:
 *   j = max(former_deg[from], seen[from]);
 *   j == 0 -> r[from] = 0, seen[from] = 1, next edge            (Q8)
 *   seen[from] = j + 1; p[from] *= (j+1)/j; r[from] -= p[from]/(j+1)/alpha;
 *   r[to] += (1-alpha) p[from]/(j+1)/alpha; enqueue from / to if r > rmax.
 * No matrix accumulator: this path builds its blocks from p itself. */
static void *refresh_edges_job(void *arg)
{
    job_t *J = (job_t *)arg;
    orc_ppr *s = J->s; const orc_graph *g = J->g; int d = J->dir;
    double a = s->alpha, rmax = s->rmax;
    int *seen = (int *)calloc((size_t)g->n, sizeof(int));
    for (int k = J->k0; k < J->k1; k++) {
        double *p = s->p[d] + IDX(k, 0), *r = s->r[d] + IDX(k, 0);
        unsigned char *fl = s->flag[d] + IDX(k, 0);
        for (long long e = 0; e < J->m; e++) {
            int from = d == 0 ? J->es[e] : J->ed[e];
            int to = d == 0 ? J->ed[e] : J->es[e];
            int f0 = g->former[d][from];
            double j = f0 > seen[from] ? f0 : seen[from];
            if (j == 0) {
                r[from] = 0.0;
                seen[from] = 1;
                continue;
            }
            seen[from] = (int)j + 1;
            p[from] = ST(p[from] * ((j + 1.0) / j));
            r[from] = ST(r[from] - p[from] / (j + 1.0) / a);
            r[to] = ST(r[to] + (1 - a) * p[from] / (j + 1.0) / a);
            if (OUT(r[from]) && !fl[from]) { q_push(s, d, k, from); fl[from] = 1; }
            if (OUT(r[to]) && !fl[to]) { q_push(s, d, k, to); fl[to] = 1; }
        }
        for (long long e = 0; e < J->m; e++) seen[d == 0 ? J->es[e] : J->ed[e]] = 0;
    }
    free(seen);
    return NULL;
}

ORC_EXPORT void orc_ppr_refresh_edges(orc_ppr *s, const orc_graph *g, int dir, const int *src, const int *dst,
                                      long long m, int nthreads)
{
    run_parallel_edges(refresh_edges_job, s, g, dir, nthreads, src, dst, m);
}

/* Storing the state as pair lists keeps only positive entries (ppr:6598-6626,
 * :6780-6808): reserves and residues <= 0 become 0, so the negative residues
 * left by the insertion update are dropped (Q4). */
ORC_EXPORT void orc_ppr_drop_negative(orc_ppr *s, int dir)
{
    size_t N = (size_t)s->nsrc * (size_t)s->n;
    for (size_t t = 0; t < N; t++) {
        if (s->p[dir][t] < 0.0) s->p[dir][t] = 0.0;
        if (s->r[dir][t] < 0.0) s->r[dir][t] = 0.0;
    }
}

/* nodegree_DenseDynamicForwardPush (ppr_computation_store_dynamic.h:1394-1498),
 * nodegree_DenseDirectedDynamicForwardPush (:2945-3030) and ...Transpose
 * (:3356-3437): FIFO per source until the queue is empty.  Degree-0 vertex:
 * undirected p = alpha*r (assignment, no matrix update, r kept, :1431-1441);
 * directed: dequeued untouched (:2971-2976).  Threshold r > rmax, not
 * degree-normalised, positive side only. */
static void *push_job(void *arg)
{
    job_t *J = (job_t *)arg;
    orc_ppr *s = J->s; const orc_graph *g = J->g; int d = J->dir;
    double a = s->alpha, rmax = s->rmax;
    const int *deg = g->deg[d];
    long long pops = 0, relax = 0;
    for (int k = J->k0; k < J->k1; k++) {
        double *p = s->p[d] + IDX(k, 0), *r = s->r[d] + IDX(k, 0), *M = s->M + IDX(k, 0);
        unsigned char *fl = s->flag[d] + IDX(k, 0);
        int *q = s->q[d] + (size_t)k * (size_t)(s->n + 1);
        int src = s->src[k];
        while (s->qh[d][k] != s->qt[d][k]) {
            int v = q[s->qh[d][k]];
            s->qh[d][k] = (s->qh[d][k] + 1) % (s->n + 1);
            if (deg[v] == 0) {
                fl[v] = 0;
                if (!g->directed) p[v] = ST(a * r[v]);
                continue;
            }
            if (OUT(r[v])) {
                pops++;
                relax += deg[v];
                const int *nb = g->adj[d][v];
                for (int j = 0; j < deg[v]; j++) {
                    int u = nb[j];
                    r[u] = ST(r[u] + (1 - a) * r[v] / deg[v]);
                    if (deg[u] == 0) {
                        if (!g->directed) {   /* unreachable for undirected graphs (:1450-1454) */
                            p[u] = ST(p[u] + a * r[u]);
                            r[src] = ST(r[src] + (1 - a) * r[u]);
                        }
                        continue;
                    }
                    if (OUT(r[u]) && !fl[u]) { q_push(s, d, k, u); fl[u] = 1; }
                }
                p[v] = ST(p[v] + a * r[v]);
                M[v] += a * r[v];
                r[v] = 0;
            }
            fl[v] = 0;
        }
        s->qh[d][k] = s->qt[d][k] = 0;
    }
    __sync_fetch_and_add(&s->pops, pops);
    __sync_fetch_and_add(&s->relax, relax);
    return NULL;
}

/* Two-sided pushing (|r| > rmax): an option beside the reference's positive-
 * only threshold (ppr:1445); the dynamic update leaves negative residues,
 * which only a two-sided push drains. */
ORC_EXPORT void orc_ppr_set_two_sided(orc_ppr *s, int on)
{
    s->two = on ? 1 : 0;
}

/* Pushes (vertices popped with r > rmax) and edge relaxations since the last
 * call; the counters are reset by the read. */
ORC_EXPORT void orc_ppr_counts(orc_ppr *s, long long *pops, long long *relax)
{
    *pops = s->pops;
    *relax = s->relax;
    s->pops = s->relax = 0;
}

ORC_EXPORT void orc_ppr_push(orc_ppr *s, const orc_graph *g, int dir, int nthreads)
{
    run_parallel(push_job, s, g, dir, nthreads);
}

/* Counts for the roofline bookkeeping: total pushes and edge relaxations of
 * the last push cannot be recovered afterwards, so this helper recomputes
 * sum_v deg(v) over the current queues (pending work). */
ORC_EXPORT long long orc_ppr_pending(const orc_ppr *s, int dir)
{
    long long c = 0;
    for (int k = 0; k < s->nsrc; k++) {
        long long h = s->qh[dir][k], t = s->qt[dir][k];
        c += (t - h + (s->n + 1)) % (s->n + 1);
    }
    return c;
}
