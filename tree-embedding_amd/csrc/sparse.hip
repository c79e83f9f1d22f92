// sparse.hip -- state of the n >= 1e7-node path: per-source positive pair
// lists, rehydrated into the dense push state one batch of source groups at a
// time.
//
// Replaces the pair lists of All_initialization_and_push_final
// (ppr_computation_store_dynamic.h:6188-6832).  There, every source keeps
// tripletList / residue_pair_list (and their transposes): one
// vector<pair<int,float>> per (block, source), restored into hash maps before
// the update and the push (:6465-6486) and written back afterwards with only
// the positive entries (:6598-6626, :6780-6808), so negative residues are
// dropped between snapshots (Q4).
//
// Layout: a list is CSR over the sources of one batch -- off[s] .. off[s+1]
// index (vertex id, value) pairs with vertex ids ascending, so block j's
// pairs of a source are one contiguous run (the reference's per-block
// vector).  Storing a list is a compaction of the dense vertex-major state
// (one wave per (vertex chunk, source group), lane = source: coalesced reads
// of 64 consecutive sources per vertex); loading one is a scatter into a
// zeroed state.
#include "tsvd_sparse.h"

#include <hipcub/hipcub.hpp>

namespace tsvd {

namespace {

constexpr int TPB = 256;
constexpr int WPB = TPB / 64;
constexpr int32_t CHUNK = 4096;   // vertices per compaction wave
constexpr int SPT = 16;           // scatter entries per thread

template <class T>
__global__ void __launch_bounds__(TPB) k_list_count(const T *__restrict__ X, int32_t n, int32_t s_pad,
                                                    int32_t count, int32_t gcount, int32_t nchunk,
                                                    int64_t *__restrict__ cnt)
{
    const int64_t wave = (int64_t)blockIdx.x * WPB + (threadIdx.x >> 6);
    if (wave >= (int64_t)nchunk * gcount) return;
    const int32_t c = (int32_t)(wave / gcount);
    const int32_t s = (int32_t)(wave - (int64_t)c * gcount) * 64 + (threadIdx.x & 63);
    const int64_t v0 = (int64_t)c * CHUNK, v1 = min((int64_t)n, v0 + CHUNK);
    int64_t k = 0;
    for (int64_t v = v0; v < v1; v++) k += X[v * s_pad + s] > T(0) ? 1 : 0;
    if (s < count) cnt[(int64_t)s * nchunk + c] = k;
}

template <class T>
__global__ void __launch_bounds__(TPB) k_list_write(const T *__restrict__ X, int32_t n, int32_t s_pad,
                                                    int32_t count, int32_t gcount, int32_t nchunk,
                                                    const int64_t *__restrict__ pos, int32_t *__restrict__ idx,
                                                    T *__restrict__ val)
{
    const int64_t wave = (int64_t)blockIdx.x * WPB + (threadIdx.x >> 6);
    if (wave >= (int64_t)nchunk * gcount) return;
    const int32_t c = (int32_t)(wave / gcount);
    const int32_t s = (int32_t)(wave - (int64_t)c * gcount) * 64 + (threadIdx.x & 63);
    if (s >= count) return;
    const int64_t v0 = (int64_t)c * CHUNK, v1 = min((int64_t)n, v0 + CHUNK);
    int64_t at = pos[(int64_t)s * nchunk + c];
    for (int64_t v = v0; v < v1; v++) {
        const T x = X[v * s_pad + s];
        if (x > T(0)) {
            idx[at] = (int32_t)v;
            val[at] = x;
            at++;
        }
    }
}

__global__ void k_list_offsets(const int64_t *__restrict__ pos, int32_t count, int32_t nchunk,
                               int64_t *__restrict__ off)
{
    const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s <= count) off[s] = pos[(int64_t)s * nchunk];
}

// Block (x, s) scatters entries [x * TPB * SPT, (x+1) * TPB * SPT) of source s.
template <class T>
__global__ void __launch_bounds__(TPB) k_list_scatter(const int64_t *__restrict__ off,
                                                      const int32_t *__restrict__ idx, const T *__restrict__ val,
                                                      T *__restrict__ X, int32_t s_pad)
{
    const int32_t s = blockIdx.y;
    const int64_t b = off[s], e = off[s + 1];
    const int64_t base = b + (int64_t)blockIdx.x * TPB * SPT;
    for (int t = 0; t < SPT; t++) {
        const int64_t at = base + (int64_t)t * TPB + threadIdx.x;
        if (at >= e) return;
        X[(int64_t)idx[at] * s_pad + s] = val[at];
    }
}

template <class T>
void store_t(SparseWork &w, PairList &L, const void *Xv, int32_t n, int32_t s_pad, int32_t count, hipStream_t st)
{
    const T *X = (const T *)Xv;
    const int32_t nchunk = (int32_t)ceil_div(n, CHUNK);
    const int32_t gcount = (int32_t)ceil_div(count, 64);
    const int64_t cells = (int64_t)count * nchunk;
    w.cnt.grow((size_t)cells + 1);
    w.pos.grow((size_t)cells + 1);
    TSVD_HIP(hipMemsetAsync(w.cnt.p + cells, 0, sizeof(int64_t), st));
    const int64_t waves = (int64_t)nchunk * gcount;
    if (g_profile) prof_add_work("sparse_store", 2.0 * (double)n * gcount * 64 * sizeof(T), 0.0);
    PROF("sparse_store", st);
    hipLaunchKernelGGL(k_list_count<T>, dim3((unsigned)ceil_div(waves, WPB)), dim3(TPB), 0, st, X, n, s_pad, count,
                       gcount, nchunk, w.cnt.p);
    size_t tmp = 0;
    TSVD_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, w.cnt.p, w.pos.p, cells + 1, st));
    w.tmp.grow(tmp + 16);
    TSVD_HIP(hipcub::DeviceScan::ExclusiveSum(w.tmp.p, tmp, w.cnt.p, w.pos.p, cells + 1, st));
    int64_t total = 0;
    TSVD_HIP(hipMemcpyAsync(&total, w.pos.p + cells, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    TSVD_HIP(hipStreamSynchronize(st));
    if ((size_t)total > L.idx.n) {   // grow by a quarter beyond the need, contents are rewritten anyway
        const size_t want = (size_t)total + (size_t)total / 4 + 1024;
        L.idx.alloc(want);
        L.val.alloc(want * sizeof(T));
    }
    L.off.grow((size_t)count + 1);
    hipLaunchKernelGGL(k_list_write<T>, dim3((unsigned)ceil_div(waves, WPB)), dim3(TPB), 0, st, X, n, s_pad, count,
                       gcount, nchunk, w.pos.p, L.idx.p, (T *)L.val.p);
    hipLaunchKernelGGL(k_list_offsets, dim3((unsigned)ceil_div(count + 1, TPB)), dim3(TPB), 0, st, w.pos.p, count,
                       nchunk, L.off.p);
    TSVD_HIP(hipGetLastError());
    L.nnz = total;
    L.count = count;
    L.esize = sizeof(T);
    TSVD_HIP(hipStreamSynchronize(st));
}

template <class T>
void load_t(const PairList &L, void *X, int32_t s_pad, hipStream_t st)
{
    if (L.count == 0 || L.nnz == 0) return;
    std::vector<int64_t> off((size_t)L.count + 1);
    TSVD_HIP(hipMemcpyAsync(off.data(), L.off.p, sizeof(int64_t) * off.size(), hipMemcpyDeviceToHost, st));
    TSVD_HIP(hipStreamSynchronize(st));
    int64_t longest = 0;
    for (int32_t s = 0; s < L.count; s++) longest = std::max(longest, off[s + 1] - off[s]);
    if (longest == 0) return;
    if (g_profile) prof_add_work("sparse_load", (double)L.nnz * (8.0 + sizeof(T)), 0.0);
    PROF("sparse_load", st);
    dim3 grid((unsigned)ceil_div(longest, (int64_t)TPB * SPT), (unsigned)L.count);
    hipLaunchKernelGGL(k_list_scatter<T>, grid, dim3(TPB), 0, st, L.off.p, L.idx.p, (const T *)L.val.p, (T *)X,
                       s_pad);
    TSVD_HIP(hipGetLastError());
}

}  // namespace

void list_store(SparseWork &w, PairList &L, const void *X, size_t esize, int32_t n, int32_t s_pad, int32_t count,
                hipStream_t st)
{
    if (esize == 8) store_t<double>(w, L, X, n, s_pad, count, st);
    else store_t<float>(w, L, X, n, s_pad, count, st);
}

void list_load(const PairList &L, void *X, int32_t s_pad, hipStream_t st)
{
    if (L.esize == 8) load_t<double>(L, X, s_pad, st);
    else load_t<float>(L, X, s_pad, st);
}

void list_clear(PairList &L)
{
    L.nnz = 0;
    L.count = 0;
}

void list_read(const PairList &L, int32_t s, int32_t n, double *out)
{
    for (int32_t v = 0; v < n; v++) out[v] = 0.0;
    if (s >= L.count || L.nnz == 0) return;
    int64_t r[2];
    TSVD_HIP(hipMemcpy(r, L.off.p + s, sizeof(r), hipMemcpyDeviceToHost));
    const int64_t len = r[1] - r[0];
    if (len == 0) return;
    std::vector<int32_t> idx((size_t)len);
    std::vector<uint8_t> val((size_t)len * L.esize);
    TSVD_HIP(hipMemcpy(idx.data(), L.idx.p + r[0], sizeof(int32_t) * len, hipMemcpyDeviceToHost));
    TSVD_HIP(hipMemcpy(val.data(), L.val.p + (size_t)r[0] * L.esize, val.size(), hipMemcpyDeviceToHost));
    for (int64_t e = 0; e < len; e++)
        out[idx[(size_t)e]] = L.esize == 8 ? ((double *)val.data())[e] : (double)((float *)val.data())[e];
}

}  // namespace tsvd
