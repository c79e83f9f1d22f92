// blocks.hip -- column-block transform, lazy dirty test and block rebuild;
// small elementwise kernels of the factorisation (sketch, eigen scaling).
//
// Replaces Log_sparse_matrix_entries[_LP] (ppr_computation_store_dynamic.h:
// 2527-2561, 2598-2638) and ..._with_norm_computation[_LP] (:2649-2732,
// :2742-2826), run there on one thread per block over a dense |S| x w_j FP64
// copy of the reserves.  Here the block matrix is never formed twice: one
// pass over the reserves (p, or p + p^T for directed graphs, the matrix of
// Tree_d_NC.cpp:689/816) computes ||A_new||_F^2 and ||A_new - A_old||_F^2 of
// every block, reading the stored A_old through a per-vertex slot map; only
// blocks found dirty are then re-gathered.  Stored columns keep the reserves'
// push-lane order (row s of a stored block is push lane s), so both passes
// read and write contiguous 64-lane segments; the lane -> subset-row
// permutation is applied where results leave the device.  The stored matrix keeps only
// active columns (a vertex with an edge, or a source); every other column is
// identically zero because its reserve is zero.
#include "tsvd_pipeline.h"

#include <cmath>

namespace tsvd {

namespace {

constexpr int TPB = 256;
constexpr int VPB = 32;   // vertices per workgroup in the norm pass

template <class T>
__device__ inline T xform(T x, T rm, int transform)
{
    if (transform == TSVD_TRANSFORM_NC) return x > rm ? log10(x / rm) : T(0);   // ppr:2546-2553
    return x > T(0) ? log10(T(1) + x / rm) : T(0);                              // ppr:2617-2630
}

__device__ inline double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <class T>
__global__ void __launch_bounds__(TPB) k_block_norms(const T *__restrict__ p0, const T *__restrict__ p1,
                                                     const T *__restrict__ A, const int32_t *__restrict__ slot,
                                                     int32_t n, int32_t nsrc,
                                                     int32_t s_pad, int32_t nparts, int32_t common, int transform,
                                                     T rm, double *__restrict__ norms, int32_t j0)
{
    const int32_t j = j0 + (int32_t)blockIdx.y;
    const int64_t start = (int64_t)j * common;
    const int64_t width = (j == nparts - 1) ? n - start : common;
    const int64_t v0 = start + (int64_t)blockIdx.x * VPB;
    if (v0 >= start + width) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double a2 = 0.0, e2 = 0.0;
    for (int vv = wave; vv < VPB; vv += TPB / 64) {
        const int64_t v = v0 + vv;
        if (v >= start + width) break;
        const int32_t sl = slot[v];
        const T *Ao = sl >= 0 ? A + (start + sl) * s_pad : nullptr;
        const T *x0 = p0 + v * s_pad;
        const T *x1 = p1 ? p1 + v * s_pad : nullptr;
        for (int s = lane; s < nsrc; s += 64) {
            T x = x0[s];
            if (x1) x += x1[s];
            const T a = xform<T>(x, rm, transform);
            const T ao = Ao ? Ao[s] : T(0);
            a2 += (double)a * (double)a;
            const double d = (double)a - (double)ao;
            e2 += d * d;
        }
    }
    a2 = wave_sum(a2);
    e2 = wave_sum(e2);
    __shared__ double red[2][TPB / 64];
    if (lane == 0) {
        red[0][wave] = a2;
        red[1][wave] = e2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0, se = 0;
        for (int w = 0; w < TPB / 64; w++) {
            sa += red[0][w];
            se += red[1][w];
        }
        atomicAdd(norms + 2 * j, sa);
        atomicAdd(norms + 2 * j + 1, se);
    }
}

// The same sums per (vertex, source group) item, for the items flagged in
// `changed` (or every item when `all`).  One wave takes 64 consecutive items:
// each lane reads one flag (and clears it), then the wave walks the flagged
// ones, lanes over the item's 64 sources, and writes the item's two sums.
// The transform and the stored-column lookup are those of k_block_norms, so
// a block's sums over its items are that kernel's sums in another order.
constexpr int DONE_SHARDS = 32;
constexpr int IB = 8;   // flagged items a wave loads together (the sums' reduce-scatter assumes 8)

template <class T>
__global__ void __launch_bounds__(TPB) k_item_norms(const T *__restrict__ p0, const T *__restrict__ p1,
                                                    const T *__restrict__ A, const int32_t *__restrict__ slot,
                                                    uint8_t *__restrict__ changed, int all, int64_t nitems,
                                                    int32_t ngroups, int32_t nsrc, int32_t s_pad, int32_t nparts,
                                                    int32_t common, int transform, T rm, double *__restrict__ sums,
                                                    unsigned long long *__restrict__ done)
{
    __shared__ unsigned long long sdone;
    if (threadIdx.x == 0) sdone = 0ull;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int64_t nw = (int64_t)gridDim.x * (TPB / 64);
    unsigned long long mine_done = 0;
    for (int64_t w = (int64_t)blockIdx.x * (TPB / 64) + (threadIdx.x >> 6); w * 64 < nitems; w += nw) {
        const int64_t it0 = w * 64, mine = it0 + lane;
        const int64_t vw = it0 / ngroups;   // the window's first item as (vertex, group)
        const int32_t gw = (int32_t)(it0 - vw * ngroups);
        const bool want = mine < nitems && (all || changed[mine] != 0);
        if (mine < nitems && (all || want)) changed[mine] = 0;
        unsigned long long bal = __ballot(want);
        // IB flagged items at a time, so that their loads are in flight together
        while (bal != 0ull) {
            int64_t its[IB];
#pragma unroll
            for (int u = 0; u < IB; u++) {
                its[u] = -1;
                if (bal != 0ull) {
                    its[u] = it0 + __builtin_ctzll(bal);
                    bal &= bal - 1ull;
                }
            }
            // every load is issued unconditionally at an in-bounds address
            // (an absent item reads item it0, a lane past nsrc reads the row's
            // padding, an unstored vertex reads its block's first stored
            // column) and masked afterwards, so all are in flight together
            int64_t vv[IB];
            int32_t ss[IB], sl[IB];
#pragma unroll
            for (int u = 0; u < IB; u++) {
                // item = (vw, gw) + k, k < 64: 32-bit arithmetic from the window's start
                const uint32_t idx = (uint32_t)gw + (uint32_t)((its[u] < 0 ? it0 : its[u]) - it0);
                const uint32_t dv = idx / (uint32_t)ngroups;
                vv[u] = vw + dv;
                ss[u] = (int32_t)(idx - dv * (uint32_t)ngroups) * 64 + lane;
                sl[u] = slot[vv[u]];
            }
            T x[IB], x1[IB], ao[IB];
#pragma unroll
            for (int u = 0; u < IB; u++) {
                const int64_t j = min((int64_t)((uint32_t)vv[u] / (uint32_t)common), (int64_t)nparts - 1);
                x[u] = p0[vv[u] * s_pad + ss[u]];
                x1[u] = p1 ? p1[vv[u] * s_pad + ss[u]] : T(0);
                ao[u] = A[(j * common + max(sl[u], 0)) * s_pad + ss[u]];
            }
#pragma unroll
            for (int u = 0; u < IB; u++) {
                const bool live = its[u] >= 0 && ss[u] < nsrc;
                x[u] = live ? x[u] + x1[u] : T(0);
                ao[u] = (live && sl[u] >= 0) ? ao[u] : T(0);
            }
            // the 2 IB = 16 sums of the batch, value 2u + 0/1 = a2/e2 of item u
            double sv[2 * IB];
#pragma unroll
            for (int u = 0; u < IB; u++) {
                const T a = xform<T>(x[u], rm, transform);
                sv[2 * u] = (double)a * (double)a;
                const double d = (double)a - (double)ao[u];
                sv[2 * u + 1] = d * d;
            }
            // reduce-scatter over the wave: at the step of mask m the lane keeps
            // the lower or upper half of its values (by its bit m) and adds the
            // partner's copy of that half; after masks 32..4 lane l holds, for
            // value (l >> 2) & 15, the sum over the 16 lanes that share l's bits
            // 0-1, and two plain steps complete it
#pragma unroll
            for (int m = 32, c = 2 * IB; m >= 4; m >>= 1, c >>= 1) {
                const bool upper = (lane & m) != 0;
#pragma unroll
                for (int q = 0; q < c / 2; q++) {
                    const double lo = sv[q], hi = sv[q + c / 2];
                    const double recv = __shfl_xor(upper ? lo : hi, m, 64);
                    sv[q] = (upper ? hi : lo) + recv;
                }
            }
            double tot = sv[0];
            tot += __shfl_xor(tot, 2, 64);
            tot += __shfl_xor(tot, 1, 64);
            const int vix = (lane >> 2) & 15;   // value held by this lane
            int64_t myit = its[0];
#pragma unroll
            for (int u = 1; u < IB; u++)
                if ((vix >> 1) == u) myit = its[u];
            if ((lane & 3) == 0 && myit >= 0) sums[2 * myit + (vix & 1)] = tot;
#pragma unroll
            for (int u = 0; u < IB; u++)
                if (its[u] >= 0) mine_done++;
        }
    }
    if (lane == 0 && mine_done) atomicAdd(&sdone, mine_done);
    __syncthreads();
    if (threadIdx.x == 0 && sdone) atomicAdd(done + blockIdx.x % DONE_SHARDS, sdone);
}

// Block sums over the item sums: `split` workgroups per block, each adding
// its part into norms[2j], norms[2j + 1] (zeroed before).
__global__ void __launch_bounds__(TPB) k_block_sums(const double *__restrict__ sums, int32_t n, int32_t ngroups,
                                                    int32_t nparts, int32_t common, int32_t j0,
                                                    double *__restrict__ norms)
{
    const int32_t j = j0 + (int32_t)blockIdx.y;
    const int64_t start = (int64_t)j * common;
    const int64_t width = (j == nparts - 1) ? n - start : common;
    const int64_t i0 = start * ngroups, i1 = (start + width) * ngroups;
    double a2 = 0.0, e2 = 0.0;
    for (int64_t i = i0 + (int64_t)blockIdx.x * TPB + threadIdx.x; i < i1; i += (int64_t)gridDim.x * TPB) {
        a2 += sums[2 * i];
        e2 += sums[2 * i + 1];
    }
    a2 = wave_sum(a2);
    e2 = wave_sum(e2);
    __shared__ double red[2][TPB / 64];
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][wave] = a2;
        red[1][wave] = e2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0.0, se = 0.0;
        for (int w = 0; w < TPB / 64; w++) {
            sa += red[0][w];
            se += red[1][w];
        }
        atomicAdd(norms + 2 * j, sa);
        atomicAdd(norms + 2 * j + 1, se);
    }
}

// A rebuilt block stores the transform of the reserves it was built from, so
// the difference sums of its items are zero until they change again.
__global__ void __launch_bounds__(TPB) k_zero_diff(const int32_t *__restrict__ blocks, int32_t n, int32_t ngroups,
                                                   int32_t nparts, int32_t common, double *__restrict__ sums)
{
    const int32_t j = blocks[blockIdx.y];
    const int64_t start = (int64_t)j * common;
    const int64_t width = (j == nparts - 1) ? n - start : common;
    const int64_t i0 = start * ngroups, i1 = (start + width) * ngroups;
    for (int64_t i = i0 + (int64_t)blockIdx.x * TPB + threadIdx.x; i < i1; i += (int64_t)gridDim.x * TPB)
        sums[2 * i + 1] = 0.0;
}

// Compact slots of the active columns of each listed block (one workgroup per
// block; vertices in id order keep their order).
__global__ void __launch_bounds__(TPB) k_slots(const int32_t *__restrict__ blocks, const uint8_t *__restrict__ active,
                                               int32_t n, int32_t nparts, int32_t common, int32_t *__restrict__ slot,
                                               int32_t *__restrict__ colmap, int32_t *__restrict__ ncols)
{
    const int32_t j = blocks[blockIdx.x];
    const int64_t start = (int64_t)j * common;
    const int64_t width = (j == nparts - 1) ? n - start : common;
    const int64_t per = (width + TPB - 1) / TPB;
    const int64_t a = start + threadIdx.x * per;
    const int64_t b = min(start + width, a + per);
    int32_t cnt = 0;
    for (int64_t v = a; v < b; v++) cnt += active[v] ? 1 : 0;
    __shared__ int32_t sc[TPB];
    sc[threadIdx.x] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t run = 0;
        for (int t = 0; t < TPB; t++) {
            const int32_t c = sc[t];
            sc[t] = run;
            run += c;
        }
        ncols[j] = run;
    }
    __syncthreads();
    int32_t pos = sc[threadIdx.x];
    for (int64_t v = a; v < b; v++) {
        if (active[v]) {
            slot[v] = pos;
            colmap[start + pos] = (int32_t)v;
            pos++;
        } else {
            slot[v] = -1;
        }
    }
}

// A[:, c] of each listed block = transform of the reserves of its c-th active
// vertex; one wave per column, lanes over sources (padding rows set to 0).
template <class T>
__global__ void __launch_bounds__(TPB) k_gather(const int32_t *__restrict__ blocks, const T *__restrict__ p0,
                                                const T *__restrict__ p1, T *__restrict__ A,
                                                const int32_t *__restrict__ colmap, const int32_t *__restrict__ ncols,
                                                int32_t nsrc, int32_t s_pad,
                                                int32_t common, int transform, T rm)
{
    const int32_t j = blocks[blockIdx.y];
    const int64_t start = (int64_t)j * common;
    const int64_t c = (int64_t)blockIdx.x * (TPB / 64) + (threadIdx.x >> 6);
    if (c >= ncols[j]) return;
    const int lane = threadIdx.x & 63;
    const int64_t v = colmap[start + c];
    T *col = A + (start + c) * s_pad;
    const T *x0 = p0 + v * s_pad;
    const T *x1 = p1 ? p1 + v * s_pad : nullptr;
    for (int s = lane; s < s_pad; s += 64) {
        T a = T(0);
        if (s < nsrc) {
            T x = x0[s];
            if (x1) x += x1[s];
            a = xform<T>(x, rm, transform);
        }
        col[s] = a;
    }
}

__global__ void k_active(const int32_t *__restrict__ deg0, const int32_t *__restrict__ deg1,
                         const uint8_t *__restrict__ is_source, int32_t n, uint8_t *__restrict__ active)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    int32_t d = deg0[v] + (deg1 ? deg1[v] : 0);
    active[v] = (d > 0 || is_source[v]) ? 1 : 0;
}

__global__ void k_mark_sources(const int32_t *__restrict__ src, int32_t nsrc, uint8_t *__restrict__ is_source)
{
    const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nsrc) is_source[src[s]] = 1;
}

__device__ inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Standard normal sketch entries from a counter-based hash (Box-Muller), for
// a batch of matrices in one launch (blockIdx.y = matrix).  Entry (i, c) of a
// rows x cols matrix is number t = c * rows + i of its seed, stored at
// X[c * ld + i] (column-major) or X[i * cols + c] (row-major).
template <class T>
__global__ void k_gaussian(const GaussBatch b)
{
    const GaussItem &g = b.it[blockIdx.y];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.rows * g.cols) return;
    const int64_t c = t / g.rows, i = t - c * g.rows;
    const uint64_t h = mix64(g.seed ^ mix64((uint64_t)t));
    const double u1 = ((double)(h >> 40) + 1.0) / 16777217.0;           // (0, 1]
    const double u2 = (double)(h & 0xFFFFFFull) / 16777216.0;           // [0, 1)
    const T v = (T)(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
    T *X = (T *)g.X;
    if (g.rowmajor) X[i * g.cols + c] = v;
    else X[c * g.ld + i] = v;
}

__global__ void k_eig_rotation(double *__restrict__ W, const double *__restrict__ V, const double *__restrict__ lam,
                               int l, double eps_abs, double eps_rel)
{
    const int b = blockIdx.x;
    const double *Vb = V + (int64_t)b * l * l;
    double *Wb = W + (int64_t)b * l * l;
    const double *lb = lam + (int64_t)b * l;
    const double smax = sqrt(fmax(lb[l - 1], 0.0));
    const double guard = fmax(eps_abs, eps_rel * smax);
    for (int e = threadIdx.x; e < l * l; e += blockDim.x) {
        const int c = e / l;
        const double s = sqrt(fmax(lb[c], 0.0));
        Wb[e] = s > guard ? Vb[e] / s : 0.0;
    }
}

__global__ void k_eig_select(double *__restrict__ W, const double *__restrict__ V, const double *__restrict__ lam,
                             int l, int k, int mode, int ascending, double eps)
{
    const int b = blockIdx.x;
    const double *Vb = V + (int64_t)b * l * l;
    double *Wb = W + (int64_t)b * l * k;
    const double *lb = lam + (int64_t)b * l;
    for (int e = threadIdx.x; e < l * k; e += blockDim.x) {
        const int c = e / l, r = e - c * l;
        const int src = ascending ? l - k + c : l - 1 - c;
        const double s = sqrt(fmax(lb[src], 0.0));
        double f = 1.0;
        if (mode == 0) f = s > eps ? s : 0.0;
        else if (mode == 1) f = s > eps ? 1.0 / sqrt(s) : 0.0;
        else if (mode == 2) f = s > eps ? 1.0 / s : 0.0;
        Wb[e] = Vb[(int64_t)src * l + r] * f;
    }
}

}  // namespace

void blocks_setup(Blocks &b, int32_t n, int32_t nparts, int32_t nsrc, int32_t s_pad, int precision, int transform,
                  double rmax, const std::vector<int32_t> &sources, hipStream_t st)
{
    TSVD_REQUIRE(nparts >= 1 && nparts <= n, "nparts must be in [1, n]");
    b.nparts = nparts;
    b.n = n;
    b.nsrc = nsrc;
    b.s_pad = s_pad;
    b.common = n / nparts;
    b.precision = precision;
    b.esize = precision == TSVD_FP64 ? 8 : 4;
    b.transform = transform;
    b.rmax = rmax;
    b.start.resize(nparts);
    b.width.resize(nparts);
    for (int32_t j = 0; j < nparts; j++) {
        b.start[j] = (int64_t)j * b.common;
        b.width[j] = (j == nparts - 1) ? n - b.start[j] : b.common;
    }
    b.ncols.assign(nparts, 0);
    b.a2.assign(nparts, 0.0);
    b.e2.assign(nparts, 0.0);
    b.resid.assign(nparts, 0.0);
    b.A.alloc((size_t)n * s_pad * b.esize);
    b.slot.alloc((size_t)n);
    b.colmap.alloc((size_t)n);
    b.d_ncols.alloc((size_t)nparts);
    b.active.alloc((size_t)n);
    b.is_source.alloc((size_t)n);
    b.norms.alloc((size_t)2 * nparts);
    b.d_dirty.alloc((size_t)nparts);
    b.row_of.alloc((size_t)s_pad);
    {
        std::vector<int32_t> id((size_t)s_pad);
        for (int32_t s = 0; s < s_pad; s++) id[(size_t)s] = s;
        TSVD_HIP(hipMemcpyAsync(b.row_of.p, id.data(), sizeof(int32_t) * s_pad, hipMemcpyHostToDevice, st));
        TSVD_HIP(hipStreamSynchronize(st));
    }
    TSVD_HIP(hipMemsetAsync(b.A.p, 0, b.A.bytes(), st));
    TSVD_HIP(hipMemsetAsync(b.slot.p, 0xFF, b.slot.bytes(), st));   // -1: no stored column
    TSVD_HIP(hipMemsetAsync(b.d_ncols.p, 0, b.d_ncols.bytes(), st));
    TSVD_HIP(hipMemsetAsync(b.is_source.p, 0, b.is_source.bytes(), st));
    DBuf<int32_t> ds((size_t)nsrc);
    TSVD_HIP(hipMemcpyAsync(ds.p, sources.data(), sizeof(int32_t) * nsrc, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_mark_sources, dim3((unsigned)ceil_div(nsrc, TPB)), dim3(TPB), 0, st, ds.p, nsrc,
                       b.is_source.p);
    TSVD_HIP(hipStreamSynchronize(st));
}

void blocks_norms(Blocks &b, const void *p0, const void *p1, const tsvd_graph *g, hipStream_t st)
{
    blocks_norms_range(b, p0, p1, g, 0, b.nparts, st);
}

void blocks_norms_range(Blocks &b, const void *p0, const void *p1, const tsvd_graph *g, int32_t j0, int32_t j1,
                        hipStream_t st)
{
    if (j1 <= j0) return;
    hipLaunchKernelGGL(k_active, dim3((unsigned)ceil_div(b.n, TPB)), dim3(TPB), 0, st, g->dir[0].deg.p,
                       g->ndir > 1 ? g->dir[1].deg.p : nullptr, b.is_source.p, b.n, b.active.p);
    TSVD_HIP(hipMemsetAsync(b.norms.p, 0, b.norms.bytes(), st));
    int64_t maxw = 0, cols = 0, stored = 0;
    for (int32_t j = j0; j < j1; j++) {
        maxw = std::max(maxw, b.width[j]);
        cols += b.width[j];
        stored += b.ncols[j];
    }
    dim3 grid((unsigned)ceil_div(maxw, VPB), (unsigned)(j1 - j0));
    if (g_profile) {
        // reserves of every vertex of the range (one or two directions) + the stored matrices
        const double dirs = p1 ? 2.0 : 1.0;
        prof_add_work("blocks_norms", (dirs * (double)cols + (double)stored) * b.nsrc * (double)b.esize, 0.0);
    }
    PROF("blocks_norms", st);
    if (b.precision == TSVD_FP64)
        hipLaunchKernelGGL(k_block_norms<double>, grid, dim3(TPB), 0, st, (const double *)p0, (const double *)p1,
                           (const double *)b.A.p, b.slot.p, b.n, b.nsrc, b.s_pad, b.nparts, b.common,
                           b.transform, b.rmax, b.norms.p, j0);
    else
        hipLaunchKernelGGL(k_block_norms<float>, grid, dim3(TPB), 0, st, (const float *)p0, (const float *)p1,
                           (const float *)b.A.p, b.slot.p, b.n, b.nsrc, b.s_pad, b.nparts, b.common,
                           b.transform, (float)b.rmax, b.norms.p, j0);
    TSVD_HIP(hipGetLastError());
    std::vector<double> h((size_t)2 * b.nparts);
    TSVD_HIP(hipMemcpyAsync(h.data(), b.norms.p, b.norms.bytes(), hipMemcpyDeviceToHost, st));
    TSVD_HIP(hipStreamSynchronize(st));
    for (int32_t j = j0; j < j1; j++) {
        b.a2[j] = h[2 * j];
        b.e2[j] = h[2 * j + 1];
    }
}

void blocks_norms_items(Blocks &b, const void *p0, const void *p1, const tsvd_graph *g, uint8_t *changed, bool all,
                        int32_t j0, int32_t j1, hipStream_t st)
{
    if (j1 <= j0) return;
    TSVD_REQUIRE(changed, "no changed-item flags");
    const int32_t ngroups = b.s_pad / 64;
    const int64_t nitems = (int64_t)b.n * ngroups;
    if (b.item_sums.n == 0) {
        b.item_sums.alloc((size_t)(2 * nitems));
        b.item_done.alloc(DONE_SHARDS);
        all = true;
    }
    hipLaunchKernelGGL(k_active, dim3((unsigned)ceil_div(b.n, TPB)), dim3(TPB), 0, st, g->dir[0].deg.p,
                       g->ndir > 1 ? g->dir[1].deg.p : nullptr, b.is_source.p, b.n, b.active.p);
    TSVD_HIP(hipMemsetAsync(b.norms.p, 0, b.norms.bytes(), st));
    TSVD_HIP(hipMemsetAsync(b.item_done.p, 0, b.item_done.bytes(), st));
    const unsigned grid = (unsigned)std::min<int64_t>(ceil_div(ceil_div(nitems, 64), TPB / 64), 16384);
    int64_t span = 0, maxw = 0;
    for (int32_t j = j0; j < j1; j++) {
        span += b.width[j];
        maxw = std::max(maxw, b.width[j]);
    }
    const unsigned split = (unsigned)std::min<int64_t>(std::max<int64_t>(ceil_div(maxw * ngroups, TPB * 16), 1), 64);
    {
        PROF("blocks_norms", st);
        if (b.precision == TSVD_FP64)
            hipLaunchKernelGGL(k_item_norms<double>, dim3(grid), dim3(TPB), 0, st, (const double *)p0,
                               (const double *)p1, (const double *)b.A.p, b.slot.p, changed, all ? 1 : 0, nitems,
                               ngroups, b.nsrc, b.s_pad, b.nparts, b.common, b.transform, b.rmax, b.item_sums.p,
                               b.item_done.p);
        else
            hipLaunchKernelGGL(k_item_norms<float>, dim3(grid), dim3(TPB), 0, st, (const float *)p0,
                               (const float *)p1, (const float *)b.A.p, b.slot.p, changed, all ? 1 : 0, nitems,
                               ngroups, b.nsrc, b.s_pad, b.nparts, b.common, b.transform, (float)b.rmax,
                               b.item_sums.p, b.item_done.p);
        TSVD_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_block_sums, dim3(split, (unsigned)(j1 - j0)), dim3(TPB), 0, st, b.item_sums.p, b.n,
                           ngroups, b.nparts, b.common, j0, b.norms.p);
        TSVD_HIP(hipGetLastError());
    }
    std::vector<double> h((size_t)2 * b.nparts);
    unsigned long long done[DONE_SHARDS];
    TSVD_HIP(hipMemcpyAsync(h.data(), b.norms.p, b.norms.bytes(), hipMemcpyDeviceToHost, st));
    TSVD_HIP(hipMemcpyAsync(done, b.item_done.p, sizeof(done), hipMemcpyDeviceToHost, st));
    TSVD_HIP(hipStreamSynchronize(st));
    b.last_items = 0;
    for (int i = 0; i < DONE_SHARDS; i++) b.last_items += (int64_t)done[i];
    if (g_profile) {
        // recomputed items: their reserves (one or two directions) and stored
        // values; every item's flag; the item sums of the range
        const double dirs = p1 ? 2.0 : 1.0;
        prof_add_work("blocks_norms",
                      (double)b.last_items * 64.0 * (double)b.esize * (dirs + 1.0) + (double)nitems +
                          16.0 * (double)span * ngroups, 0.0);
    }
    for (int32_t j = j0; j < j1; j++) {
        b.a2[j] = h[2 * j];
        b.e2[j] = h[2 * j + 1];
    }
}

void blocks_rebuild(Blocks &b, const std::vector<int32_t> &dirty, const void *p0, const void *p1, hipStream_t st)
{
    if (dirty.empty()) return;
    const int nd = (int)dirty.size();
    TSVD_HIP(hipMemcpyAsync(b.d_dirty.p, dirty.data(), sizeof(int32_t) * nd, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_slots, dim3(nd), dim3(TPB), 0, st, b.d_dirty.p, b.active.p, b.n, b.nparts, b.common,
                       b.slot.p, b.colmap.p, b.d_ncols.p);
    int64_t maxw = 0;
    for (int32_t j : dirty) maxw = std::max(maxw, b.width[j]);
    dim3 grid((unsigned)ceil_div(maxw, TPB / 64), (unsigned)nd);
    PROF("blocks_gather", st);
    if (g_profile) {
        double w = 0.0;
        for (int32_t j : dirty) w += (double)b.width[j];
        prof_add_work("blocks_gather", w * b.nsrc * (double)b.esize * (p1 ? 3.0 : 2.0), 0.0);
    }
    if (b.precision == TSVD_FP64)
        hipLaunchKernelGGL(k_gather<double>, grid, dim3(TPB), 0, st, b.d_dirty.p, (const double *)p0,
                           (const double *)p1, (double *)b.A.p, b.colmap.p, b.d_ncols.p, b.nsrc,
                           b.s_pad, b.common, b.transform, b.rmax);
    else
        hipLaunchKernelGGL(k_gather<float>, grid, dim3(TPB), 0, st, b.d_dirty.p, (const float *)p0,
                           (const float *)p1, (float *)b.A.p, b.colmap.p, b.d_ncols.p, b.nsrc,
                           b.s_pad, b.common, b.transform, (float)b.rmax);
    TSVD_HIP(hipGetLastError());
    if (b.item_sums.n > 0) {
        const int32_t ngroups = b.s_pad / 64;
        const unsigned split = (unsigned)std::min<int64_t>(std::max<int64_t>(ceil_div(maxw * ngroups, TPB * 16), 1), 64);
        hipLaunchKernelGGL(k_zero_diff, dim3(split, (unsigned)nd), dim3(TPB), 0, st, b.d_dirty.p, b.n, ngroups,
                           b.nparts, b.common, b.item_sums.p);
        TSVD_HIP(hipGetLastError());
    }
    TSVD_HIP(hipMemcpyAsync(b.ncols.data(), b.d_ncols.p, sizeof(int32_t) * b.nparts, hipMemcpyDeviceToHost, st));
    TSVD_HIP(hipStreamSynchronize(st));
}

namespace {

// One thread per pair: the dot product of the source's row of its block's U
// with the target's row of V (Tree_u_LP.cpp:1058-1069).
__global__ void k_lp_scores(const int32_t *__restrict__ slot, int32_t n, int32_t nparts, int32_t common,
                            const double *__restrict__ V, const double *const *__restrict__ ublk, int32_t k,
                            int32_t s_pad, const int32_t *__restrict__ rows, const int32_t *__restrict__ dst, int64_t m,
                            const double *const *__restrict__ dblk, const double *__restrict__ vdeg,
                            double *__restrict__ scores)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    const int32_t v = dst[e];
    const int32_t j = min(v / common, nparts - 1);
    const int64_t start = (int64_t)j * common;
    const int64_t width = (j == nparts - 1) ? n - start : common;
    const int32_t c = slot[v];
    const double *U = ublk[j];
    double acc = 0.0;
    if (c >= 0 && U != nullptr) {
        const double *Vj = V + start * k + c;
        const int32_t r = rows[e];
        for (int32_t t = 0; t < k; t++) acc += U[(int64_t)t * s_pad + r] * Vj[(int64_t)t * width];
        if (dblk) acc *= dblk[j][r] * vdeg[v];
    }
    scores[e] = acc;
}

}  // namespace

void lp_scores(const Blocks &b, const double *V, const double *const *ublk, const double *const *dblk,
               const double *vdeg, int32_t k, const int32_t *rows,
               const int32_t *dst, int64_t m, double *scores, hipStream_t st)
{
    if (m <= 0) return;
    PROF("lp_scores", st);
    hipLaunchKernelGGL(k_lp_scores, dim3((unsigned)ceil_div(m, TPB)), dim3(TPB), 0, st, b.slot.p, b.n, b.nparts,
                       b.common, V, ublk, k, b.s_pad, rows, dst, m, dblk, vdeg, scores);
    TSVD_HIP(hipGetLastError());
}

void gaussian_fill(const std::vector<GaussItem> &items, int precision, hipStream_t st)
{
    // the batch travels as a kernel argument (GAUSS_BATCH items per launch)
    for (size_t b0 = 0; b0 < items.size(); b0 += GAUSS_BATCH) {
        GaussBatch gb{};
        const int n = (int)std::min<size_t>(GAUSS_BATCH, items.size() - b0);
        int64_t maxtot = 0;
        for (int k = 0; k < n; k++) {
            gb.it[k] = items[b0 + k];
            maxtot = std::max<int64_t>(maxtot, gb.it[k].rows * gb.it[k].cols);
        }
        if (maxtot <= 0) continue;
        const dim3 grid((unsigned)ceil_div(maxtot, TPB), (unsigned)n);
        if (precision == TSVD_FP64) hipLaunchKernelGGL(k_gaussian<double>, grid, dim3(TPB), 0, st, gb);
        else hipLaunchKernelGGL(k_gaussian<float>, grid, dim3(TPB), 0, st, gb);
        TSVD_HIP(hipGetLastError());
    }
}

void eig_rotation(double *W, const double *V, const double *lam, int l, int count, double eps_abs, double eps_rel,
                  hipStream_t st)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(k_eig_rotation, dim3(count), dim3(TPB), 0, st, W, V, lam, l, eps_abs, eps_rel);
    TSVD_HIP(hipGetLastError());
}

void eig_select(double *W, const double *V, const double *lam, int l, int k, int count, int mode, int ascending,
                double eps, hipStream_t st)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(k_eig_select, dim3(count), dim3(TPB), 0, st, W, V, lam, l, k, mode, ascending, eps);
    TSVD_HIP(hipGetLastError());
}

}  // namespace tsvd

namespace tsvd {

namespace {

// ---- Blocks of the n >= 1e7-node path -------------------------------------
// Log_sparse_matrix_entries_LP_Robinhood_two_with_transpose_nparts_triplet
// [_norm_computation] (ppr_computation_store_dynamic.h:6898-7003, 7080-7229)
// builds block j from the forward and transpose reserve lists: entries of
// both are inserted and duplicates summed in double (x = p + p^T of the FP32
// values), then log10(1 + x/rmax).  Rows are inserted in 30 intervals of
// floor(|S|/30) rows, so rows >= 30 floor(|S|/30) stay empty (Q11).  Here a
// batch of sources is at hand as dense FP32 state (width sp), rows
// [row0, row0 + rows) of the matrix.
__device__ inline double xform_lp(float pf, float pt, double rm)
{
    const double x = (double)pf + (double)pt;
    return x > 0.0 ? log10(1.0 + x / rm) : 0.0;
}

template <class T>
__global__ void __launch_bounds__(TPB) k_batch_norms(const float *__restrict__ pf, const float *__restrict__ pt,
                                                     const T *__restrict__ A, const int32_t *__restrict__ slot,
                                                     int32_t n, int32_t sp, int32_t row0, int32_t rows,
                                                     int32_t s_pad, int32_t keep, int32_t nparts, int32_t common,
                                                     double rm, double *__restrict__ norms)
{
    const int32_t j = blockIdx.y;
    const int64_t start = (int64_t)j * common;
    const int64_t width = (j == nparts - 1) ? n - start : common;
    const int64_t v0 = start + (int64_t)blockIdx.x * VPB;
    if (v0 >= start + width) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double a2 = 0.0, e2 = 0.0;
    for (int vv = wave; vv < VPB; vv += TPB / 64) {
        const int64_t v = v0 + vv;
        if (v >= start + width) break;
        const int32_t sl = slot[v];
        const T *Ao = sl >= 0 ? A + (start + sl) * s_pad + row0 : nullptr;
        for (int s = lane; s < rows; s += 64) {
            const T a = row0 + s < keep ? (T)xform_lp(pf[v * sp + s], pt[v * sp + s], rm) : T(0);
            const T ao = Ao ? Ao[s] : T(0);
            a2 += (double)a * (double)a;
            const double d = (double)a - (double)ao;
            e2 += d * d;
        }
    }
    a2 = wave_sum(a2);
    e2 = wave_sum(e2);
    __shared__ double red[2][TPB / 64];
    if (lane == 0) {
        red[0][wave] = a2;
        red[1][wave] = e2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0, se = 0;
        for (int w = 0; w < TPB / 64; w++) {
            sa += red[0][w];
            se += red[1][w];
        }
        atomicAdd(norms + 2 * j, sa);
        atomicAdd(norms + 2 * j + 1, se);
    }
}

// Rows [row0, row0 + width) of the stored columns of each listed block; rows
// past the batch's sources (padding) and past `keep` are written as zero.
template <class T>
__global__ void __launch_bounds__(TPB) k_batch_gather(const int32_t *__restrict__ blocks,
                                                      const float *__restrict__ pf, const float *__restrict__ pt,
                                                      T *__restrict__ A, const int32_t *__restrict__ colmap,
                                                      const int32_t *__restrict__ ncols, int32_t sp, int32_t row0,
                                                      int32_t rows, int32_t width, int32_t s_pad, int32_t common,
                                                      int32_t keep, double rm)
{
    const int32_t j = blocks[blockIdx.y];
    const int64_t start = (int64_t)j * common;
    const int64_t c = (int64_t)blockIdx.x * (TPB / 64) + (threadIdx.x >> 6);
    if (c >= ncols[j]) return;
    const int lane = threadIdx.x & 63;
    const int64_t v = colmap[start + c];
    T *col = A + (start + c) * s_pad + row0;
    for (int s = lane; s < width; s += 64)
        col[s] = (s < rows && row0 + s < keep) ? (T)xform_lp(pf[v * sp + s], pt[v * sp + s], rm) : T(0);
}

}  // namespace

void blocks_active(Blocks &b, const tsvd_graph *g, hipStream_t st)
{
    hipLaunchKernelGGL(k_active, dim3((unsigned)ceil_div(b.n, TPB)), dim3(TPB), 0, st, g->dir[0].deg.p,
                       g->ndir > 1 ? g->dir[1].deg.p : nullptr, b.is_source.p, b.n, b.active.p);
    TSVD_HIP(hipMemsetAsync(b.norms.p, 0, b.norms.bytes(), st));
}

void blocks_norms_batch(Blocks &b, const float *pf, const float *pt, int32_t sp, int32_t row0, int32_t rows,
                        int32_t keep, hipStream_t st)
{
    int64_t maxw = 0;
    for (auto w : b.width) maxw = std::max(maxw, w);
    dim3 grid((unsigned)ceil_div(maxw, VPB), (unsigned)b.nparts);
    if (g_profile) {
        double stored = 0.0;
        for (int32_t j = 0; j < b.nparts; j++) stored += (double)b.ncols[j];
        prof_add_work("blocks_norms", (2.0 * 4.0 * b.n + stored * b.esize) * rows, 0.0);
    }
    PROF("blocks_norms", st);
    if (b.precision == TSVD_FP64)
        hipLaunchKernelGGL(k_batch_norms<double>, grid, dim3(TPB), 0, st, pf, pt, (const double *)b.A.p, b.slot.p,
                           b.n, sp, row0, rows, b.s_pad, keep, b.nparts, b.common, b.rmax, b.norms.p);
    else
        hipLaunchKernelGGL(k_batch_norms<float>, grid, dim3(TPB), 0, st, pf, pt, (const float *)b.A.p, b.slot.p,
                           b.n, sp, row0, rows, b.s_pad, keep, b.nparts, b.common, b.rmax, b.norms.p);
    TSVD_HIP(hipGetLastError());
}

void blocks_norms_read(Blocks &b, hipStream_t st)
{
    std::vector<double> h((size_t)2 * b.nparts);
    TSVD_HIP(hipMemcpyAsync(h.data(), b.norms.p, b.norms.bytes(), hipMemcpyDeviceToHost, st));
    TSVD_HIP(hipStreamSynchronize(st));
    for (int32_t j = 0; j < b.nparts; j++) {
        b.a2[j] = h[2 * j];
        b.e2[j] = h[2 * j + 1];
    }
}

void blocks_slots(Blocks &b, const std::vector<int32_t> &dirty, hipStream_t st)
{
    if (dirty.empty()) return;
    const int nd = (int)dirty.size();
    TSVD_HIP(hipMemcpyAsync(b.d_dirty.p, dirty.data(), sizeof(int32_t) * nd, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_slots, dim3(nd), dim3(TPB), 0, st, b.d_dirty.p, b.active.p, b.n, b.nparts, b.common,
                       b.slot.p, b.colmap.p, b.d_ncols.p);
    TSVD_HIP(hipGetLastError());
}

void blocks_gather_batch(Blocks &b, const std::vector<int32_t> &dirty, const float *pf, const float *pt, int32_t sp,
                         int32_t row0, int32_t rows, int32_t keep, hipStream_t st)
{
    if (dirty.empty()) return;
    const int nd = (int)dirty.size();
    int64_t maxw = 0;
    for (int32_t j : dirty) maxw = std::max(maxw, b.width[j]);
    const int32_t width = std::min(sp, b.s_pad - row0);
    dim3 grid((unsigned)ceil_div(maxw, TPB / 64), (unsigned)nd);
    if (g_profile) {
        double w = 0.0;
        for (int32_t j : dirty) w += (double)b.width[j];
        prof_add_work("blocks_gather", w * width * (8.0 + (double)b.esize), 0.0);
    }
    PROF("blocks_gather", st);
    if (b.precision == TSVD_FP64)
        hipLaunchKernelGGL(k_batch_gather<double>, grid, dim3(TPB), 0, st, b.d_dirty.p, pf, pt, (double *)b.A.p,
                           b.colmap.p, b.d_ncols.p, sp, row0, rows, width, b.s_pad, b.common, keep, b.rmax);
    else
        hipLaunchKernelGGL(k_batch_gather<float>, grid, dim3(TPB), 0, st, b.d_dirty.p, pf, pt, (float *)b.A.p,
                           b.colmap.p, b.d_ncols.p, sp, row0, rows, width, b.s_pad, b.common, keep, b.rmax);
    TSVD_HIP(hipGetLastError());
}

void blocks_ncols_read(Blocks &b, hipStream_t st)
{
    TSVD_HIP(hipMemcpyAsync(b.ncols.data(), b.d_ncols.p, sizeof(int32_t) * b.nparts, hipMemcpyDeviceToHost, st));
    TSVD_HIP(hipStreamSynchronize(st));
}

namespace {

// X[(v - v0) * s_pad + s] = transform of the reserves at vertex v (FP64), the
// padding rows zero: a vertex-major slab of the full |S| x n matrix.
template <class T>
__global__ void k_transform_chunk(const T *__restrict__ p0, const T *__restrict__ p1, double *__restrict__ X,
                                  const int32_t *__restrict__ row_of, int64_t v0, int64_t nv, int32_t nsrc,
                                  int32_t s_pad, int transform, T rm)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nv * s_pad) return;
    const int64_t vv = t / s_pad;
    const int32_t s = (int32_t)(t - vv * s_pad);
    double a = 0.0;
    if (s < nsrc) {
        const int64_t at = (v0 + vv) * s_pad + s;
        T x = p0[at];
        if (p1) x += p1[at];
        a = (double)xform<T>(x, rm, transform);
    }
    X[vv * s_pad + row_of[s]] = a;
}

}  // namespace

void transform_chunk(const Blocks &b, const void *p0, const void *p1, double *X, int64_t v0, int64_t nv,
                     hipStream_t st)
{
    const int64_t tot = nv * b.s_pad;
    if (tot <= 0) return;
    if (b.precision == TSVD_FP64)
        hipLaunchKernelGGL(k_transform_chunk<double>, dim3((unsigned)ceil_div(tot, TPB)), dim3(TPB), 0, st,
                           (const double *)p0, (const double *)p1, X, b.row_of.p, v0, nv, b.nsrc, b.s_pad,
                           b.transform, b.rmax);
    else
        hipLaunchKernelGGL(k_transform_chunk<float>, dim3((unsigned)ceil_div(tot, TPB)), dim3(TPB), 0, st,
                           (const float *)p0, (const float *)p1, X, b.row_of.p, v0, nv, b.nsrc, b.s_pad,
                           b.transform, (float)b.rmax);
    TSVD_HIP(hipGetLastError());
}

}  // namespace tsvd
