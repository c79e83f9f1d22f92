// graph.hip -- device CSR with per-row append slack (the dynamic graph store).
//
// Replaces UGraph::inputDynamicGraph (Graph_dynamic.h:530-605) and
// Graph::inputDirectedDynamicGraph (Graph_dynamic.h:119-238).  Those re-allocate
// every grown row to its exact new size and append the batch's edges in batch
// order.  Here each direction is one CSR whose rows carry slack
// (capacity >= degree); an append
//   1. expands the batch to (row, neighbour) entries in batch order,
//   2. stable-radix-sorts them by row (equal rows keep batch order),
//   3. finds each row's segment in the sorted list,
//   4. rebuilds the CSR with fresh slack if any row would overflow,
//   5. scatters every entry to rowoff[row] + deg[row] + rank-in-segment.
// The result is bit-identical to the reference's rows.  Rows touched by the
// batch are listed (ascending) for the dynamic PPR update.
#include "tsvd_internal.h"
#include "tsvd_graph.h"
#include "tsvd_profile.h"

#include <hipcub/hipcub.hpp>
#include <algorithm>
#include <chrono>
#include <cmath>

namespace tsvd {

namespace {

constexpr int TPB = 256;

__global__ void k_validate(const int32_t *__restrict__ src, const int32_t *__restrict__ dst, int64_t m, int32_t n,
                           int32_t *__restrict__ bad)
{
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    int32_t u = src[e], v = dst[e];
    if (u < 0 || u >= n || v < 0 || v >= n) atomicOr(bad, 1);
}

// Expand the batch into (row, value) entries of one direction, batch order.
// mode 0: undirected (entries 2e: u->v, 2e+1: v->u); 1: out-rows (u->v);
// 2: in-rows (v->u).
__global__ void k_expand(const int32_t *__restrict__ src, const int32_t *__restrict__ dst, int64_t m, int mode,
                         int32_t *__restrict__ key, int32_t *__restrict__ val)
{
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    int32_t u = src[e], v = dst[e];
    if (mode == 0) {
        key[2 * e] = u; val[2 * e] = v;
        key[2 * e + 1] = v; val[2 * e + 1] = u;
    } else if (mode == 1) {
        key[e] = u; val[e] = v;
    } else {
        key[e] = v; val[e] = u;
    }
}

// Segment start of every row present in the sorted keys.
__global__ void k_seg_start(const int32_t *__restrict__ key, int64_t ne, int32_t *__restrict__ segstart)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ne) return;
    int32_t k = key[i];
    if (i == 0 || key[i - 1] != k) segstart[k] = (int32_t)i;
}

// cnt[row] = length of the row's segment (written at the segment end).
__global__ void k_seg_len(const int32_t *__restrict__ key, int64_t ne, const int32_t *__restrict__ segstart,
                          int32_t *__restrict__ cnt)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ne) return;
    int32_t k = key[i];
    if (i == ne - 1 || key[i + 1] != k) cnt[k] = (int32_t)(i + 1 - segstart[k]);
}

__global__ void k_check_cap(const int64_t *__restrict__ rowoff, const int32_t *__restrict__ deg,
                            const int32_t *__restrict__ cnt, int32_t n, int32_t *__restrict__ overflow)
{
    int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    int64_t cap = rowoff[v + 1] - rowoff[v];
    if ((int64_t)deg[v] + cnt[v] > cap) atomicOr(overflow, 1);
}

// New capacity of every row: need*(1+slack) rounded up, at least need+2 --
// empty rows included, so that a vertex's first edges (common on power-law
// graphs, where many vertices have no in- or out-edge yet) do not force a
// rebuild of the whole CSR.
__global__ void k_new_cap(const int32_t *__restrict__ deg, const int32_t *__restrict__ cnt, int32_t n, double slack,
                          int64_t *__restrict__ cap)
{
    int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    int64_t need = (int64_t)deg[v] + cnt[v];
    int64_t c = need + (int64_t)ceil((double)need * slack);
    if (c < need + 2) c = need + 2;
    cap[v] = c;
}

// Move every row's current contents to the new layout.
__global__ void k_move_rows(const int64_t *__restrict__ old_off, const int32_t *__restrict__ old_adj,
                            const int64_t *__restrict__ new_off, int32_t *__restrict__ new_adj,
                            const int32_t *__restrict__ deg, int32_t n)
{
    // one wave per row, lanes over the row
    int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int32_t v = (int32_t)(gt / 64);
    int lane = threadIdx.x & 63;
    if (v >= n) return;
    int32_t d = deg[v];
    const int32_t *s = old_adj + old_off[v];
    int32_t *t = new_adj + new_off[v];
    for (int32_t j = lane; j < d; j += 64) t[j] = s[j];
}

__global__ void k_scatter(const int32_t *__restrict__ key, const int32_t *__restrict__ val, int64_t ne,
                          const int32_t *__restrict__ segstart, const int64_t *__restrict__ rowoff,
                          const int32_t *__restrict__ deg, int32_t *__restrict__ adj)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ne) return;
    int32_t k = key[i];
    adj[rowoff[k] + deg[k] + (i - segstart[k])] = val[i];
}

__global__ void k_add_deg(int32_t *__restrict__ deg, const int32_t *__restrict__ cnt, int32_t n)
{
    int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) deg[v] += cnt[v];
}

// Distinct rows of the sorted keys, ascending: position = exclusive scan of
// the boundary flags.
__global__ void k_boundary_flags(const int32_t *__restrict__ key, int64_t ne, int32_t *__restrict__ flag)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ne) return;
    flag[i] = (i == 0 || key[i - 1] != key[i]) ? 1 : 0;
}

__global__ void k_boundary_list(const int32_t *__restrict__ key, int64_t ne, const int32_t *__restrict__ pos,
                                int32_t *__restrict__ touched)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ne) return;
    if (i == 0 || key[i - 1] != key[i]) touched[pos[i]] = key[i];
}

int bits_for(int32_t n)
{
    int b = 1;
    while (b < 31 && ((int64_t)1 << b) < (int64_t)n) b++;
    return b;
}

}  // namespace

void graph_init(tsvd_graph *g)
{
    for (int d = 0; d < g->ndir; d++) {
        GraphDir &D = g->dir[d];
        D.rowoff.alloc((size_t)g->n + 1);
        D.deg.alloc((size_t)g->n);
        D.former.alloc((size_t)g->n);
        D.cnt.alloc((size_t)g->n);
        D.touched.alloc((size_t)g->n);
        TSVD_HIP(hipMemsetAsync(D.rowoff.p, 0, D.rowoff.bytes(), g->stream));
        TSVD_HIP(hipMemsetAsync(D.deg.p, 0, D.deg.bytes(), g->stream));
        TSVD_HIP(hipMemsetAsync(D.former.p, 0, D.former.bytes(), g->stream));
        D.cap_total = 0;
        D.ntouched = 0;
    }
    g->segstart.alloc((size_t)g->n);
    TSVD_HIP(hipMallocHost((void **)&g->h_scalar, 64));
    TSVD_HIP(hipStreamSynchronize(g->stream));
}

static void append_dir(tsvd_graph *g, int d, const int32_t *dsrc, const int32_t *ddst, int64_t m)
{
    GraphDir &D = g->dir[d];
    hipStream_t st = g->stream;
    const int mode = g->directed ? (d == 0 ? 1 : 2) : 0;
    const int64_t ne = g->directed ? m : 2 * m;
    TSVD_REQUIRE(ne < (int64_t)INT32_MAX, "batch of %lld entries exceeds the int32 sort range", (long long)ne);

    g->k_in.grow((size_t)ne);
    g->v_in.grow((size_t)ne);
    g->k_out.grow((size_t)ne);
    g->v_out.grow((size_t)ne);
    g->scan_flag.grow((size_t)ne);
    g->scan_pos.grow((size_t)ne);

    D.appended = ne;
    // former <- degree (Graph_dynamic.h:545-547 / :135-138)
    TSVD_HIP(hipMemcpyAsync(D.former.p, D.deg.p, D.deg.bytes(), hipMemcpyDeviceToDevice, st));
    TSVD_HIP(hipMemsetAsync(D.cnt.p, 0, D.cnt.bytes(), st));
    if (ne == 0) {
        D.ntouched = 0;
        return;
    }
    const unsigned gm = (unsigned)ceil_div(m, TPB), ge = (unsigned)ceil_div(ne, TPB), gn = (unsigned)ceil_div(g->n, TPB);
    hipLaunchKernelGGL(k_expand, dim3(gm), dim3(TPB), 0, st, dsrc, ddst, m, mode, g->k_in.p, g->v_in.p);

    // stable sort by row
    size_t tmp = 0;
    const int bits = bits_for(g->n);
    TSVD_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, g->k_in.p, g->k_out.p, g->v_in.p, g->v_out.p, (int)ne, 0,
                                                bits, st));
    g->cub_tmp.grow(tmp + 16);
    TSVD_HIP(hipcub::DeviceRadixSort::SortPairs(g->cub_tmp.p, tmp, g->k_in.p, g->k_out.p, g->v_in.p, g->v_out.p,
                                                (int)ne, 0, bits, st));

    hipLaunchKernelGGL(k_seg_start, dim3(ge), dim3(TPB), 0, st, g->k_out.p, ne, g->segstart.p);
    hipLaunchKernelGGL(k_seg_len, dim3(ge), dim3(TPB), 0, st, g->k_out.p, ne, g->segstart.p, D.cnt.p);

    // touched rows (ascending)
    hipLaunchKernelGGL(k_boundary_flags, dim3(ge), dim3(TPB), 0, st, g->k_out.p, ne, g->scan_flag.p);
    tmp = 0;
    TSVD_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, g->scan_flag.p, g->scan_pos.p, (int)ne, st));
    g->cub_tmp.grow(tmp + 16);
    TSVD_HIP(hipcub::DeviceScan::ExclusiveSum(g->cub_tmp.p, tmp, g->scan_flag.p, g->scan_pos.p, (int)ne, st));
    hipLaunchKernelGGL(k_boundary_list, dim3(ge), dim3(TPB), 0, st, g->k_out.p, ne, g->scan_pos.p, D.touched.p);

    // capacity check
    int32_t *hflag = (int32_t *)g->h_scalar;
    int32_t *dflag = g->d_scalar();
    TSVD_HIP(hipMemsetAsync(dflag, 0, sizeof(int32_t) * 4, st));
    hipLaunchKernelGGL(k_check_cap, dim3(gn), dim3(TPB), 0, st, D.rowoff.p, D.deg.p, D.cnt.p, g->n, dflag);
    // count of touched rows = last pos + last flag
    TSVD_HIP(hipMemcpyAsync(hflag, dflag, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    int32_t hpos_last = 0, hflag_last = 0;
    TSVD_HIP(hipMemcpyAsync(&hflag[1], g->scan_pos.p + (ne - 1), sizeof(int32_t), hipMemcpyDeviceToHost, st));
    TSVD_HIP(hipMemcpyAsync(&hflag[2], g->scan_flag.p + (ne - 1), sizeof(int32_t), hipMemcpyDeviceToHost, st));
    TSVD_HIP(hipStreamSynchronize(st));
    const int overflow = hflag[0];
    hpos_last = hflag[1];
    hflag_last = hflag[2];
    D.ntouched = hpos_last + hflag_last;

    if (overflow) {
        // rebuild with fresh slack (exclusive scan of the new capacities);
        // its wall time, allocations and frees included, is profiled as
        // "graph_rebuild"
        const auto t_rb = std::chrono::steady_clock::now();
        DBuf<int64_t> cap((size_t)g->n + 1);
        DBuf<int64_t> newoff((size_t)g->n + 1);
        hipLaunchKernelGGL(k_new_cap, dim3(gn), dim3(TPB), 0, st, D.deg.p, D.cnt.p, g->n, g->slack, cap.p);
        TSVD_HIP(hipMemsetAsync(cap.p + g->n, 0, sizeof(int64_t), st));
        tmp = 0;
        TSVD_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, cap.p, newoff.p, g->n + 1, st));
        g->cub_tmp.grow(tmp + 16);
        TSVD_HIP(hipcub::DeviceScan::ExclusiveSum(g->cub_tmp.p, tmp, cap.p, newoff.p, g->n + 1, st));
        int64_t total = 0;
        TSVD_HIP(hipMemcpyAsync(&total, newoff.p + g->n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
        TSVD_HIP(hipStreamSynchronize(st));
        DBuf<int32_t> nadj((size_t)std::max<int64_t>(total, 1));
        if (D.cap_total > 0) {
            const int64_t threads = (int64_t)g->n * 64;
            hipLaunchKernelGGL(k_move_rows, dim3((unsigned)ceil_div(threads, TPB)), dim3(TPB), 0, st, D.rowoff.p,
                               D.adj.p, newoff.p, nadj.p, D.deg.p, g->n);
        }
        TSVD_HIP(hipStreamSynchronize(st));   // the old arrays are freed below
        D.adj = std::move(nadj);
        D.rowoff = std::move(newoff);
        D.cap_total = total;
        g->rebuilds++;
        prof_add_time("graph_rebuild",
                      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_rb).count());
    }
    hipLaunchKernelGGL(k_scatter, dim3(ge), dim3(TPB), 0, st, g->k_out.p, g->v_out.p, ne, g->segstart.p, D.rowoff.p,
                       D.deg.p, D.adj.p);
    hipLaunchKernelGGL(k_add_deg, dim3(gn), dim3(TPB), 0, st, D.deg.p, D.cnt.p, g->n);
}

void graph_append(tsvd_graph *g, const int32_t *src, const int32_t *dst, int64_t m, int on_device)
{
    TSVD_REQUIRE(m >= 0, "negative edge count");
    TSVD_REQUIRE(m == 0 || (src && dst), "null edge arrays");
    hipStream_t st = g->stream;
    const int32_t *dsrc = src, *ddst = dst;
    if (m > 0 && !on_device) {
        g->e_src.grow((size_t)m);
        g->e_dst.grow((size_t)m);
        TSVD_HIP(hipMemcpyAsync(g->e_src.p, src, sizeof(int32_t) * m, hipMemcpyHostToDevice, st));
        TSVD_HIP(hipMemcpyAsync(g->e_dst.p, dst, sizeof(int32_t) * m, hipMemcpyHostToDevice, st));
        dsrc = g->e_src.p;
        ddst = g->e_dst.p;
    }
    if (m > 0) {
        int32_t *dflag = g->d_scalar();
        TSVD_HIP(hipMemsetAsync(dflag, 0, sizeof(int32_t), st));
        hipLaunchKernelGGL(k_validate, dim3((unsigned)ceil_div(m, TPB)), dim3(TPB), 0, st, dsrc, ddst, m, g->n, dflag);
        int32_t *h = (int32_t *)g->h_scalar;
        TSVD_HIP(hipMemcpyAsync(h, dflag, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        TSVD_HIP(hipStreamSynchronize(st));
        TSVD_REQUIRE(h[0] == 0, "edge batch holds a vertex id outside [0, %d)", g->n);
    }
    {
        PROF("graph_append", st);
        for (int d = 0; d < g->ndir; d++) append_dir(g, d, dsrc, ddst, m);
    }
    g->m += m;
    g->version++;
    TSVD_HIP(hipStreamSynchronize(st));
    if (m > g->release_above) {
        // an initial load of a large graph: its sort staging (keys, values,
        // sort and scan buffers, about 32 B per edge) is not kept for the
        // small batches that follow.  The staged edge ids stay: the caller's
        // insertion update may still read them (graph_release_edges frees
        // them once it has).
        g->k_in.release(); g->v_in.release(); g->k_out.release(); g->v_out.release();
        g->scan_flag.release(); g->scan_pos.release(); g->cub_tmp.release();
    }
}

void graph_release_edges(tsvd_graph *g)
{
    if ((int64_t)g->e_src.n > g->release_above) {
        TSVD_HIP(hipStreamSynchronize(g->stream));
        g->e_src.release();
        g->e_dst.release();
    }
}

}  // namespace tsvd

using namespace tsvd;

extern "C" int tsvd_graph_create(int32_t n, int directed, double slack, tsvd_graph **out)
{
    return guard([&] {
        TSVD_REQUIRE(out, "null output handle");
        TSVD_REQUIRE(n > 0, "vertex count must be positive");
        TSVD_REQUIRE(slack >= 0.0 && slack < 16.0, "slack out of range");
        tsvd_graph *g = new tsvd_graph();
        g->n = n;
        g->directed = directed ? 1 : 0;
        g->ndir = directed ? 2 : 1;
        g->slack = slack;
        try {
            TSVD_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
            g->scalar.alloc(64);
            graph_init(g);
        } catch (...) {
            delete g;
            throw;
        }
        *out = g;
    });
}

extern "C" int tsvd_graph_destroy(tsvd_graph *g)
{
    return guard([&] { delete g; });
}

extern "C" int tsvd_graph_append(tsvd_graph *g, const int32_t *src, const int32_t *dst, int64_t m, int on_device)
{
    return guard([&] {
        TSVD_REQUIRE(g, "null graph");
        graph_append(g, src, dst, m, on_device);
        graph_release_edges(g);
    });
}

extern "C" int tsvd_graph_set_staging_limit(tsvd_graph *g, int64_t edges)
{
    return guard([&] {
        TSVD_REQUIRE(g, "null graph");
        TSVD_REQUIRE(edges >= 0, "negative edge count");
        g->release_above = edges;
    });
}

extern "C" int64_t tsvd_graph_num_edges(const tsvd_graph *g) { return g ? g->m : -1; }

extern "C" int tsvd_graph_degrees(const tsvd_graph *g, int dir, int32_t *deg, int32_t *former)
{
    return guard([&] {
        TSVD_REQUIRE(g && dir >= 0 && dir < g->ndir, "bad graph or direction");
        const GraphDir &D = g->dir[dir];
        if (deg) TSVD_HIP(hipMemcpyAsync(deg, D.deg.p, D.deg.bytes(), hipMemcpyDeviceToHost, g->stream));
        if (former) TSVD_HIP(hipMemcpyAsync(former, D.former.p, D.former.bytes(), hipMemcpyDeviceToHost, g->stream));
        TSVD_HIP(hipStreamSynchronize(g->stream));
    });
}

extern "C" int tsvd_graph_export_csr(const tsvd_graph *g, int dir, int64_t *rowptr, int32_t *col)
{
    return guard([&] {
        TSVD_REQUIRE(g && dir >= 0 && dir < g->ndir && rowptr, "bad graph, direction or rowptr");
        const GraphDir &D = g->dir[dir];
        std::vector<int32_t> deg((size_t)g->n);
        TSVD_HIP(hipMemcpyAsync(deg.data(), D.deg.p, D.deg.bytes(), hipMemcpyDeviceToHost, g->stream));
        std::vector<int64_t> off((size_t)g->n + 1);
        TSVD_HIP(hipMemcpyAsync(off.data(), D.rowoff.p, D.rowoff.bytes(), hipMemcpyDeviceToHost, g->stream));
        std::vector<int32_t> adj((size_t)std::max<int64_t>(D.cap_total, 1));
        if (col && D.cap_total > 0)
            TSVD_HIP(hipMemcpyAsync(adj.data(), D.adj.p, sizeof(int32_t) * D.cap_total, hipMemcpyDeviceToHost,
                                    g->stream));
        TSVD_HIP(hipStreamSynchronize(g->stream));
        int64_t pos = 0;
        for (int32_t v = 0; v < g->n; v++) {
            rowptr[v] = pos;
            if (col)
                for (int32_t j = 0; j < deg[v]; j++) col[pos + j] = adj[off[v] + j];
            pos += deg[v];
        }
        rowptr[g->n] = pos;
    });
}
