// spmm.hip -- sparse leaf products of the blockwise frPCA.
//
// Replaces the MKL sparse products of frPCA's passes (csr_matrix_matrix_mult
// and csr_matrix_transpose_matrix_mult, matrix_vector_functions_intel_mkl_
// ext.c:364-398: mkl_dcsrmm "N" and "T" over the 1-based CSR block built by
// csr_init_from_coo, ext.c:94-117, 319-359).  A dirty block whose stored
// matrix is sparse enough is factored from a CSR copy (for Q = A Qt) and a
// CSC copy (for Qt = A^T Q) built on the device from the stored dense
// columns; its Q and Qt are then row-major (leading dimension l), so both
// products are row gathers: one wavefront per output row, the row's column
// ids and values read 64 at a time and broadcast by shuffles, lanes over the
// l columns of the dense operand (l <= 192 in three registers per lane), no
// atomics.  Denser blocks keep the grouped MFMA GEMM.
//
// Building: (1) one wave per stored column counts its non-zero rows (ballot)
// and adds the count to its block's total; the host picks the sparse blocks
// by density; (2) for those, one wave per 64-row chunk counts each row's
// non-zeros (lanes = rows, walking the block's columns, which are contiguous
// per column); (3) one exclusive scan over every sparse block's row counts
// and column counts, laid out block by block (rows, then columns), gives each
// row and column its run in one entry array; (4) the same two walks write
// the entries (CSR: column ids ascending; CSC: row ids ascending).
#include "tsvd_pipeline.h"

#include <hipcub/hipcub.hpp>

namespace tsvd {

namespace {

constexpr int TPB = 256;
constexpr int WPB = TPB / 64;
constexpr int RU = 16;   // column loads in flight per thread in the row walks
constexpr int GU = 8;    // entries whose operand rows are in flight per wave in the gathers
constexpr int CU = 4;    // 64-row segments in flight per wave in the column walks

template <class T>
__global__ void __launch_bounds__(TPB) k_col_counts(const SpItem *__restrict__ items, int64_t *__restrict__ cnt,
                                                    unsigned long long *__restrict__ item_nnz, int with_cnt)
{
    const SpItem &it = items[blockIdx.y];
    const int64_t c = (int64_t)blockIdx.x * WPB + (threadIdx.x >> 6);
    if (c >= it.nc) return;
    const int lane = threadIdx.x & 63;
    const T *col = (const T *)it.A + c * it.lda;
    uint32_t n = 0;
    for (int32_t r0 = 0; r0 < it.m; r0 += 64 * CU) {
        T x[CU];
#pragma unroll
        for (int q = 0; q < CU; q++) {
            const int32_t r = r0 + q * 64 + lane;
            x[q] = r < it.m ? col[r] : T(0);
        }
#pragma unroll
        for (int q = 0; q < CU; q++) n += (uint32_t)__popcll(__ballot(x[q] != T(0)));
    }
    if (lane == 0) {
        if (with_cnt) cnt[it.base + it.m + c] = (int64_t)n;
        atomicAdd(item_nnz + blockIdx.y, (unsigned long long)n);
    }
}

template <class T>
__global__ void __launch_bounds__(TPB) k_row_counts(const SpItem *__restrict__ items, int64_t *__restrict__ cnt)
{
    const SpItem &it = items[blockIdx.y];
    const int32_t r = (int32_t)((int64_t)blockIdx.x * TPB + threadIdx.x);
    if ((int64_t)blockIdx.x * TPB >= it.m) return;
    if (r >= it.m) return;
    const T *__restrict__ A = (const T *)it.A + r;
    int64_t n = 0;
    // RU column loads in flight per thread (the loop is latency-bound otherwise)
    for (int64_t c0 = 0; c0 < it.nc; c0 += RU) {
        T x[RU];
#pragma unroll
        for (int u = 0; u < RU; u++) x[u] = c0 + u < it.nc ? A[(c0 + u) * it.lda] : T(0);
#pragma unroll
        for (int u = 0; u < RU; u++) n += x[u] != T(0);
    }
    cnt[it.base + r] = n;
}

// CSR entries: lane = row, walking the columns in order
template <class T>
__global__ void __launch_bounds__(TPB) k_row_fill(const SpItem *__restrict__ items, const int64_t *__restrict__ off,
                                                  int32_t *__restrict__ idx, T *__restrict__ val)
{
    const SpItem &it = items[blockIdx.y];
    const int32_t r = (int32_t)((int64_t)blockIdx.x * TPB + threadIdx.x);
    if (r >= it.m) return;
    int64_t p = off[it.base + r];
    const T *__restrict__ A = (const T *)it.A + r;
    for (int64_t c0 = 0; c0 < it.nc; c0 += RU) {
        T x[RU];
#pragma unroll
        for (int u = 0; u < RU; u++) x[u] = c0 + u < it.nc ? A[(c0 + u) * it.lda] : T(0);
#pragma unroll
        for (int u = 0; u < RU; u++) {
            if (x[u] != T(0)) {
                idx[p] = (int32_t)(c0 + u);
                val[p] = x[u];
                p++;
            }
        }
    }
}

// CSC entries: one wave per column, rows in order (ballot + prefix popcount)
template <class T>
__global__ void __launch_bounds__(TPB) k_col_fill(const SpItem *__restrict__ items, const int64_t *__restrict__ off,
                                                  int32_t *__restrict__ idx, T *__restrict__ val)
{
    const SpItem &it = items[blockIdx.y];
    const int64_t c = (int64_t)blockIdx.x * WPB + (threadIdx.x >> 6);
    if (c >= it.nc) return;
    const int lane = threadIdx.x & 63;
    const T *col = (const T *)it.A + c * it.lda;
    int64_t p = off[it.base + it.m + c];
    for (int32_t r0 = 0; r0 < it.m; r0 += 64 * CU) {
        T xs[CU];
#pragma unroll
        for (int q = 0; q < CU; q++) {
            const int32_t r = r0 + q * 64 + lane;
            xs[q] = r < it.m ? col[r] : T(0);
        }
#pragma unroll
        for (int q = 0; q < CU; q++) {
            const int32_t r = r0 + q * 64 + lane;
            const T x = xs[q];
            const bool nz = x != T(0);
            const unsigned long long b = __ballot(nz);
            const int before = __popcll(b & ((1ull << lane) - 1ull));
            if (nz) {
                idx[p + before] = r;
                val[p + before] = x;
            }
            p += __popcll(b);
        }
    }
}

// One output row per wave: out[r, :] = sum over the run of (j, a) of a * in[j, :]
// (row-major, leading dimension l): CSR rows give Q = A Qt, CSC columns give
// Qt = A^T Q.
template <class T, bool ROWS>
__global__ void __launch_bounds__(TPB) k_spmm_gather(const SpItem *__restrict__ items,
                                                     const int64_t *__restrict__ off,
                                                     const int32_t *__restrict__ idx, const T *__restrict__ val,
                                                     int32_t l)
{
    const SpItem &it = items[blockIdx.y];
    const int64_t r = (int64_t)blockIdx.x * WPB + (threadIdx.x >> 6);
    const int64_t nout = ROWS ? it.m : it.nc;
    if (r >= nout) return;
    const int lane = threadIdx.x & 63;
    const T *__restrict__ in = (const T *)(ROWS ? it.Qt : it.Q);
    T *__restrict__ out = (T *)(ROWS ? it.Q : it.Qt);
    const int64_t at = ROWS ? it.base + r : it.base + it.m + r;
    const int64_t b = off[at], e = off[at + 1];
    T a0 = T(0), a1 = T(0), a2 = T(0);
    for (int64_t t0 = b; t0 < e; t0 += 64) {
        const int64_t t = t0 + lane;
        int32_t j = 0;
        T v = T(0);
        if (t < e) {
            j = idx[t];
            v = val[t];
        }
        const int cnt = (int)min((int64_t)64, e - t0);
        int u = 0;
        // GU entries at a time: their operand rows are loaded before any is
        // used (the walk is latency-bound otherwise); the sums keep entry order
        for (; u + GU <= cnt; u += GU) {
            int32_t jj[GU];
            T a[GU], x0[GU], x1[GU], x2[GU];
#pragma unroll
            for (int q = 0; q < GU; q++) {
                jj[q] = __shfl(j, u + q);
                a[q] = __shfl(v, u + q);
            }
#pragma unroll
            for (int q = 0; q < GU; q++) {
                const T *__restrict__ x = in + (int64_t)jj[q] * l;
                x0[q] = lane < l ? x[lane] : T(0);
                x1[q] = lane + 64 < l ? x[lane + 64] : T(0);
                x2[q] = lane + 128 < l ? x[lane + 128] : T(0);
            }
#pragma unroll
            for (int q = 0; q < GU; q++) {
                if (lane < l) a0 += a[q] * x0[q];
                if (lane + 64 < l) a1 += a[q] * x1[q];
                if (lane + 128 < l) a2 += a[q] * x2[q];
            }
        }
        for (; u < cnt; u++) {
            const int32_t jj = __shfl(j, u);
            const T a = __shfl(v, u);
            const T *__restrict__ x = in + (int64_t)jj * l;
            if (lane < l) a0 += a * x[lane];
            if (lane + 64 < l) a1 += a * x[lane + 64];
            if (lane + 128 < l) a2 += a * x[lane + 128];
        }
    }
    T *__restrict__ y = out + r * l;
    if (lane < l) y[lane] = a0;
    if (lane + 64 < l) y[lane + 64] = a1;
    if (lane + 128 < l) y[lane + 128] = a2;
}

template <class T>
void build_t(SparseLeaves &s, hipStream_t st)
{
    const int cnt = (int)s.items.size();
    int64_t maxm = 0, maxnc = 0;
    for (auto &it : s.items) {
        maxm = std::max<int64_t>(maxm, it.m);
        maxnc = std::max<int64_t>(maxnc, it.nc);
    }
    // rows then columns of every candidate of the density pass, whose column
    // counts are in place (the runs of candidates that stay dense are zero)
    const int64_t len = s.counts;
    s.off.grow((size_t)len + 1);
    hipLaunchKernelGGL(k_row_counts<T>, dim3((unsigned)ceil_div(maxm, TPB), (unsigned)cnt), dim3(TPB), 0, st,
                       s.ditems.p, s.cnt.p);
    TSVD_HIP(hipGetLastError());
    size_t need = 0;
    TSVD_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, need, s.cnt.p, s.off.p, len + 1, st));
    s.tmp.grow(need + 256);
    need = s.tmp.n;
    TSVD_HIP(hipcub::DeviceScan::ExclusiveSum(s.tmp.p, need, s.cnt.p, s.off.p, len + 1, st));
    int64_t total = 0;
    TSVD_HIP(hipMemcpyAsync(&total, s.off.p + len, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    TSVD_HIP(hipStreamSynchronize(st));
    s.idx.grow((size_t)std::max<int64_t>(total, 1));
    s.val.grow((size_t)std::max<int64_t>(total, 1) * sizeof(T));
    hipLaunchKernelGGL(k_row_fill<T>, dim3((unsigned)ceil_div(maxm, TPB), (unsigned)cnt), dim3(TPB), 0, st,
                       s.ditems.p, s.off.p, s.idx.p, (T *)s.val.p);
    hipLaunchKernelGGL(k_col_fill<T>, dim3((unsigned)ceil_div(maxnc, WPB), (unsigned)cnt), dim3(TPB), 0, st,
                       s.ditems.p, s.off.p, s.idx.p, (T *)s.val.p);
    TSVD_HIP(hipGetLastError());
    s.entries = total;
}

}  // namespace

std::vector<double> leaf_densities(SparseLeaves &s, std::vector<SpItem> &cand, int precision, hipStream_t st)
{
    std::vector<double> dens(cand.size(), 1.0);
    s.cand.clear();
    s.counts = 0;
    if (cand.empty()) return dens;
    const int cnt = (int)cand.size();
    // every candidate's run of row counts and column counts in one array: the
    // column counts made here are the ones the CSC build of a sparse block uses
    int64_t base = 0;
    for (auto &it : cand) {
        it.base = base;
        base += (int64_t)it.m + it.nc;
    }
    s.counts = base;
    s.cand = cand;
    s.cnt.grow((size_t)base + 1);
    TSVD_HIP(hipMemsetAsync(s.cnt.p, 0, sizeof(int64_t) * (base + 1), st));
    s.ditems.grow(cand.size());
    s.nnz.grow(cand.size());
    TSVD_HIP(hipMemcpyAsync(s.ditems.p, cand.data(), sizeof(SpItem) * cand.size(), hipMemcpyHostToDevice, st));
    TSVD_HIP(hipMemsetAsync(s.nnz.p, 0, sizeof(unsigned long long) * cand.size(), st));
    int64_t maxnc = 0;
    for (auto &it : cand) maxnc = std::max<int64_t>(maxnc, it.nc);
    if (maxnc > 0) {
        PROF("spmm_count", st);
        if (precision == TSVD_FP64)
            hipLaunchKernelGGL(k_col_counts<double>, dim3((unsigned)ceil_div(maxnc, WPB), (unsigned)cnt), dim3(TPB),
                               0, st, s.ditems.p, s.cnt.p, s.nnz.p, 1);
        else
            hipLaunchKernelGGL(k_col_counts<float>, dim3((unsigned)ceil_div(maxnc, WPB), (unsigned)cnt), dim3(TPB),
                               0, st, s.ditems.p, s.cnt.p, s.nnz.p, 1);
        TSVD_HIP(hipGetLastError());
    }
    std::vector<unsigned long long> h(cand.size());
    TSVD_HIP(hipMemcpyAsync(h.data(), s.nnz.p, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost, st));
    TSVD_HIP(hipStreamSynchronize(st));
    for (size_t i = 0; i < cand.size(); i++) {
        const double cells = (double)cand[i].m * (double)cand[i].nc;
        dens[i] = cells > 0 ? (double)h[i] / cells : 1.0;
    }
    return dens;
}

void leaf_sparse_build(SparseLeaves &s, std::vector<SpItem> items, int precision, hipStream_t st)
{
    s.items = std::move(items);
    s.entries = 0;
    if (s.items.empty()) return;
    // the items are candidates of the last density pass (same runs); the
    // candidates that stay dense get their column counts cleared so that the
    // scan gives them empty runs
    std::vector<char> chosen(s.cand.size(), 0);
    for (auto &it : s.items) {
        bool found = false;
        for (size_t c = 0; c < s.cand.size() && !found; c++)
            if (s.cand[c].base == it.base && s.cand[c].A == it.A && s.cand[c].nc == it.nc) {
                chosen[c] = 1;
                found = true;
            }
        TSVD_REQUIRE(found, "sparse leaf item without a density pass");
    }
    for (size_t c = 0; c < s.cand.size(); c++)
        if (!chosen[c] && s.cand[c].nc > 0)
            TSVD_HIP(hipMemsetAsync(s.cnt.p + s.cand[c].base + s.cand[c].m, 0, sizeof(int64_t) * s.cand[c].nc, st));
    s.ditems.grow(s.items.size());
    TSVD_HIP(hipMemcpyAsync(s.ditems.p, s.items.data(), sizeof(SpItem) * s.items.size(), hipMemcpyHostToDevice, st));
    PROF("spmm_build", st);
    if (precision == TSVD_FP64) build_t<double>(s, st);
    else build_t<float>(s, st);
}

void leaf_spmm(SparseLeaves &s, int precision, int32_t l, bool transposed, hipStream_t st)
{
    if (s.items.empty()) return;
    // the items' Q buffers alternate between passes: send the current pointers
    TSVD_HIP(hipMemcpyAsync(s.ditems.p, s.items.data(), sizeof(SpItem) * s.items.size(), hipMemcpyHostToDevice, st));
    TSVD_HIP(hipStreamSynchronize(st));
    TSVD_REQUIRE(l >= 1 && l <= 192, "sparse leaf products take l <= 192 (got %d)", l);
    int64_t maxn = 0;
    double nnz = 0.0, rows = 0.0;
    for (auto &it : s.items) {
        maxn = std::max<int64_t>(maxn, transposed ? it.nc : it.m);
        rows += (double)(transposed ? it.nc : it.m) + (double)(transposed ? it.m : it.nc);
    }
    nnz = (double)s.entries / 2.0;
    const double es = precision == TSVD_FP64 ? 8.0 : 4.0;
    if (maxn <= 0) return;
    const char *label = transposed ? "spmm_t" : "spmm_n";
    if (g_profile)   // entries (id + value) + both dense operands once
        prof_add_work(label, nnz * (4.0 + es) + rows * l * es, 2.0 * nnz * l);
    PROF(label, st);
    const dim3 grid((unsigned)ceil_div(maxn, WPB), (unsigned)s.items.size());
    if (precision == TSVD_FP64) {
        if (transposed)
            hipLaunchKernelGGL((k_spmm_gather<double, false>), grid, dim3(TPB), 0, st, s.ditems.p, s.off.p, s.idx.p,
                               (const double *)s.val.p, l);
        else
            hipLaunchKernelGGL((k_spmm_gather<double, true>), grid, dim3(TPB), 0, st, s.ditems.p, s.off.p, s.idx.p,
                               (const double *)s.val.p, l);
    } else {
        if (transposed)
            hipLaunchKernelGGL((k_spmm_gather<float, false>), grid, dim3(TPB), 0, st, s.ditems.p, s.off.p, s.idx.p,
                               (const float *)s.val.p, l);
        else
            hipLaunchKernelGGL((k_spmm_gather<float, true>), grid, dim3(TPB), 0, st, s.ditems.p, s.off.p, s.idx.p,
                               (const float *)s.val.p, l);
    }
    TSVD_HIP(hipGetLastError());
}

}  // namespace tsvd
