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
// by density; (2) for those, each of 8 column ranges of a block has one
// thread per row count that row's non-zeros in the range (lanes = rows,
// walking the range's columns, which are contiguous per column), and the
// ranges' partial counts are summed per row; (3) one exclusive scan over
// every sparse block's row counts and column counts, laid out block by block
// (rows, then columns), gives each row and column its run in one entry
// array; (4) the same walks write the entries, a range's row entries after
// the earlier ranges' (CSR: column ids ascending; CSC: row ids ascending).
#include "tsvd_pipeline.h"

#include <hipcub/hipcub.hpp>

namespace tsvd {

namespace {

constexpr int TPB = 256;
constexpr int WPB = TPB / 64;
constexpr int RU = 16;   // column loads in flight per thread in the row walks
constexpr int RS = 8;    // column ranges of a block, each walked by its own threads in the row walks
constexpr int GU = 8;    // entries whose operand rows are in flight per wave in the gathers
constexpr int CU = 16;   // 64-row segments in flight per wave in the column walks
constexpr int CPW = 8;   // columns per wave in the column walks

// Column range s of nc columns, [c0, c1).
__device__ inline void col_range(int64_t nc, int s, int64_t &c0, int64_t &c1)
{
    c0 = nc * s / RS;
    c1 = nc * (s + 1) / RS;
}

template <class T>
__global__ void __launch_bounds__(TPB) k_col_counts(const SpItem *__restrict__ items, int64_t *__restrict__ cnt)
{
    const SpItem &it = items[blockIdx.y];
    const int lane = threadIdx.x & 63;
    // a wave walks every (gridDim.x * WPB)-th column of its item
    for (int64_t c = (int64_t)blockIdx.x * WPB + (threadIdx.x >> 6); c < it.nc; c += (int64_t)gridDim.x * WPB) {
        const gptr<const T> col = gbl((const T *)it.A) + c * it.lda;
        uint32_t n = 0;
        for (int32_t r0 = 0; r0 < it.m; r0 += 64 * CU) {
            T x[CU];
            // clamped (not conditional) loads, so that all CU are in flight together
#pragma unroll
            for (int q = 0; q < CU; q++) x[q] = col[min(r0 + q * 64 + lane, it.m - 1)];
#pragma unroll
            for (int q = 0; q < CU; q++)
                n += (uint32_t)__popcll(__ballot(r0 + q * 64 + lane < it.m && x[q] != T(0)));
        }
        if (lane == 0) cnt[it.base + it.m + c] = (int64_t)n;
    }
}

// Non-zeros of each item: the sum of its column counts, one workgroup per
// item (a fixed summation order, no atomics)
__global__ void __launch_bounds__(TPB) k_item_nnz(const SpItem *__restrict__ items, const int64_t *__restrict__ cnt,
                                                  unsigned long long *__restrict__ item_nnz)
{
    const SpItem &it = items[blockIdx.x];
    __shared__ unsigned long long part[WPB];
    unsigned long long n = 0;
    for (int64_t c = threadIdx.x; c < it.nc; c += TPB) n += (unsigned long long)cnt[it.base + it.m + c];
    for (int o = 32; o > 0; o >>= 1) n += __shfl_xor(n, o);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < WPB; w++) t += part[w];
        item_nnz[blockIdx.x] = t;
    }
}

// Row counts of one column range (blockIdx.z) per thread: lane = row,
// walking the range's columns (each column contiguous, so a wave's loads are
// coalesced); partial counts go to part[item][range][row]
template <class T>
__global__ void __launch_bounds__(TPB) k_row_counts(const SpItem *__restrict__ items, int32_t *__restrict__ part,
                                                    int64_t maxm)
{
    const SpItem &it = items[blockIdx.y];
    const int32_t r = (int32_t)((int64_t)blockIdx.x * TPB + threadIdx.x);
    if (r >= it.m) return;
    int64_t c0, c1;
    col_range(it.nc, blockIdx.z, c0, c1);
    const gptr<const T> __restrict__ A = gbl((const T *)it.A) + r;
    int32_t n = 0;
    // RU column loads in flight per thread (the loop is latency-bound otherwise)
    for (int64_t cb = c0; cb < c1; cb += RU) {
        T x[RU];
        // clamped (not conditional) loads, so that all RU are in flight together
#pragma unroll
        for (int u = 0; u < RU; u++) x[u] = A[min(cb + u, c1 - 1) * it.lda];
#pragma unroll
        for (int u = 0; u < RU; u++) n += (cb + u < c1 && x[u] != T(0));
    }
    part[((int64_t)blockIdx.y * RS + blockIdx.z) * maxm + r] = n;
}

// Each row's count: the sum of its ranges' partial counts
__global__ void __launch_bounds__(TPB) k_row_sum(const SpItem *__restrict__ items, const int32_t *__restrict__ part,
                                                 int64_t maxm, int64_t *__restrict__ cnt)
{
    const SpItem &it = items[blockIdx.y];
    const int32_t r = (int32_t)((int64_t)blockIdx.x * TPB + threadIdx.x);
    if (r >= it.m) return;
    int64_t n = 0;
#pragma unroll
    for (int q = 0; q < RS; q++) n += part[((int64_t)blockIdx.y * RS + q) * maxm + r];
    cnt[it.base + r] = n;
}

// CSR entries of one column range (blockIdx.z): lane = row, walking the
// range's columns in order from the row's offset plus the earlier ranges'
// counts, so every row's column ids come out ascending
template <class T>
__global__ void __launch_bounds__(TPB) k_row_fill(const SpItem *__restrict__ items, const int64_t *__restrict__ off,
                                                  const int32_t *__restrict__ part, int64_t maxm,
                                                  int32_t *__restrict__ idx, T *__restrict__ val)
{
    const SpItem &it = items[blockIdx.y];
    const int32_t r = (int32_t)((int64_t)blockIdx.x * TPB + threadIdx.x);
    if (r >= it.m) return;
    int64_t p = off[it.base + r];
    for (int q = 0; q < (int)blockIdx.z; q++) p += part[((int64_t)blockIdx.y * RS + q) * maxm + r];
    int64_t c0, c1;
    col_range(it.nc, blockIdx.z, c0, c1);
    const gptr<const T> __restrict__ A = gbl((const T *)it.A) + r;
    for (int64_t cb = c0; cb < c1; cb += RU) {
        T x[RU];
#pragma unroll
        for (int u = 0; u < RU; u++) x[u] = cb + u < c1 ? A[(cb + u) * it.lda] : T(0);
#pragma unroll
        for (int u = 0; u < RU; u++) {
            if (x[u] != T(0)) {
                idx[p] = (int32_t)(cb + u);
                val[p] = x[u];
                p++;
            }
        }
    }
}

// CSC entries: a wave per column at a time, rows in order (ballot + prefix popcount)
template <class T>
__global__ void __launch_bounds__(TPB) k_col_fill(const SpItem *__restrict__ items, const int64_t *__restrict__ off,
                                                  int32_t *__restrict__ idx, T *__restrict__ val)
{
    const SpItem &it = items[blockIdx.y];
    const int lane = threadIdx.x & 63;
    // a wave walks every (gridDim.x * WPB)-th column of its item
    for (int64_t c = (int64_t)blockIdx.x * WPB + (threadIdx.x >> 6); c < it.nc; c += (int64_t)gridDim.x * WPB) {
        const gptr<const T> col = gbl((const T *)it.A) + c * it.lda;
        int64_t p = off[it.base + it.m + c];
        for (int32_t r0 = 0; r0 < it.m; r0 += 64 * CU) {
            T xs[CU];
            // clamped (not conditional) loads, so that all CU are in flight together
#pragma unroll
            for (int q = 0; q < CU; q++) xs[q] = col[min(r0 + q * 64 + lane, it.m - 1)];
#pragma unroll
            for (int q = 0; q < CU; q++) {
                const int32_t r = r0 + q * 64 + lane;
                const T x = xs[q];
                const bool nz = r < it.m && x != T(0);
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
    const gptr<const T> __restrict__ in = gbl((const T *)(ROWS ? it.Qt : it.Q));
    const gptr<T> __restrict__ out = gbl((T *)(ROWS ? it.Q : it.Qt));
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
                const gptr<const T> __restrict__ x = in + (int64_t)jj[q] * l;
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
            const gptr<const T> __restrict__ x = in + (int64_t)jj * l;
            if (lane < l) a0 += a * x[lane];
            if (lane + 64 < l) a1 += a * x[lane + 64];
            if (lane + 128 < l) a2 += a * x[lane + 128];
        }
    }
    const gptr<T> __restrict__ y = out + r * l;
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
    s.part.grow((size_t)cnt * RS * maxm);
    const dim3 rgrid((unsigned)ceil_div(maxm, TPB), (unsigned)cnt, RS);
    hipLaunchKernelGGL(k_row_counts<T>, rgrid, dim3(TPB), 0, st, s.ditems.p, s.part.p, maxm);
    hipLaunchKernelGGL(k_row_sum, dim3((unsigned)ceil_div(maxm, TPB), (unsigned)cnt), dim3(TPB), 0, st, s.ditems.p,
                       s.part.p, maxm, s.cnt.p);
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
    hipLaunchKernelGGL(k_row_fill<T>, rgrid, dim3(TPB), 0, st, s.ditems.p, s.off.p, s.part.p, maxm, s.idx.p,
                       (T *)s.val.p);
    hipLaunchKernelGGL(k_col_fill<T>, dim3((unsigned)ceil_div(maxnc, WPB * CPW), (unsigned)cnt), dim3(TPB), 0, st,
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
            hipLaunchKernelGGL(k_col_counts<double>, dim3((unsigned)ceil_div(maxnc, WPB * CPW), (unsigned)cnt), dim3(TPB),
                               0, st, s.ditems.p, s.cnt.p);
        else
            hipLaunchKernelGGL(k_col_counts<float>, dim3((unsigned)ceil_div(maxnc, WPB * CPW), (unsigned)cnt), dim3(TPB),
                               0, st, s.ditems.p, s.cnt.p);
        hipLaunchKernelGGL(k_item_nnz, dim3((unsigned)cnt), dim3(TPB), 0, st, s.ditems.p, s.cnt.p, s.nnz.p);
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
