// eig.hip -- batched FP64 symmetric eigensolver on the device.
//
// Replaces LAPACKE_dsyevd in eigSVD (frpca.c:99-136) and, for the merges,
// the dgesvd of the f*k-column child matrix (matrix_vector_functions_intel_
// mkl.c:1259-1322 via Tree_u_NC.cpp:238-262), which the device computes from
// the FP64 Gram of that matrix.  `count` independent n x n problems
// (n <= EIG_NMAX), each in three stages:
//
//  1. Householder tridiagonalization T = Q^T A Q (Golub and Van Loan,
//     Matrix Computations, Alg. 8.3.1).  n <= 136: one workgroup per matrix
//     with the matrix in LDS.  Larger n: one workgroup per matrix on an
//     L2-resident copy in global memory, in panels of nb columns whose
//     reflectors are applied to the trailing matrix as one rank-2nb update
//     (the panel form of LAPACK's dsytrd/dlatrd: a panel column is brought up
//     to date and its matrix-vector product corrected from the panel's V, W).
//  2. Cuppen's divide and conquer on T.  T is torn at split points into
//     segments of <= SEG rows (T = diag(T1, T2) + rho u u^T, rho = |e_m|, the
//     two adjacent diagonals reduced by rho); every segment is solved by the
//     implicit symmetric QR algorithm with Wilkinson shifts (G&VL Alg.
//     8.3.2-8.3.3), one wave driving the bulge chase and the others applying
//     the previous chase's rotations to the eigenvectors in LDS.  Sibling
//     segments are merged level by level: D + rho z z^T with z = (last row of
//     Q1, sign(e_m) first row of Q2) is deflated (|rho z_i| <= tol, and close
//     d_i, d_j by a Givens rotation), each root of the secular equation
//     1 + rho sum z_i^2 / (d_i - lam) = 0 is found by the two-pole rational
//     model of Bunch, Nielsen and Sorensen (1978) kept inside a bisection
//     bracket, in coordinates relative to the nearer pole, z is recomputed
//     from the roots (Gu and Eisenstat, SIAM J. Matrix Anal. Appl. 16, 1995)
//     so that the vectors z_i / (d_i - lam_j) are orthogonal, and the merged
//     eigenvectors diag(Q1, Q2) Y are two grouped FP64 MFMA products.  The
//     last merge forms only the `top` largest eigenvectors.
//  3. Back-transformation of those eigenvectors by the Householder
//     reflectors, a slice of columns per workgroup held in LDS.
//
// Output: eigenvalues ascending (all n), eigenvectors as columns in the same
// order over the input matrix; with top < n only the last `top` columns are
// written.  The workspace (eig_workspace_bytes) is caller-provided.
#include "tsvd_pipeline.h"

#include <algorithm>
#include <array>
#include <functional>

namespace tsvd {

namespace {

constexpr int SEG = 48;          // leaf segment rows: one wave holds a segment's d, e
constexpr int LDS_N = 136;       // tridiagonalization in LDS up to this n
constexpr int TRI_T = 1024;
constexpr int TRG_T = 1024;
constexpr int PANEL_MAX = 16;
constexpr int VW_CAP = 8192;     // doubles of each panel matrix (V, W) in LDS (16 columns of n = 512)
constexpr int LEAF_T = 128;      // wave 0 drives the chase, wave 1 rotates the rows of Z
constexpr int MRG_T = 512;
constexpr int BACK_T = 1024;
constexpr int MAX_PIECES = 32;
constexpr double EPS = 2.220446049250313e-16;
constexpr double TINY = 1e-300;

typedef double f64x4 __attribute__((ext_vector_type(4)));

struct SegTable {
    int count;
    int a[MAX_PIECES], len[MAX_PIECES];
};

struct MergeTable {
    int count;
    int a[MAX_PIECES], n1[MAX_PIECES], n2[MAX_PIECES];
};

// v of the lane that DPP control CTRL selects, both 32-bit halves moved
// (every lane active at the call: the wave reductions below run with the
// full execution mask)
template <int CTRL>
__device__ inline double dpp_f64(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// lane l's value of v, for a constant l
template <int L>
__device__ inline double lane_f64(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, L);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), L);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Wave reductions, the same value in every lane: within each row of 16
// lanes by DPP (quad_perm [1,0,3,2] and [2,3,0,1] pair lanes l ^ 1 and
// l ^ 2; row_half_mirror pairs the two quads of an 8-lane half, row_mirror
// the two halves), the operation being commutative so that every lane of a
// row holds the same row total; then the four row totals in a fixed order.
// Four VALU steps and four lane reads, where a shuffle butterfly takes six
// dependent crossbar round trips.
__device__ inline double wave_sum(double v)
{
    v += dpp_f64<0xB1>(v);
    v += dpp_f64<0x4E>(v);
    v += dpp_f64<0x141>(v);
    v += dpp_f64<0x140>(v);
    return ((lane_f64<0>(v) + lane_f64<16>(v)) + lane_f64<32>(v)) + lane_f64<48>(v);
}

__device__ inline double wave_max(double v)
{
    v = fmax(v, dpp_f64<0xB1>(v));
    v = fmax(v, dpp_f64<0x4E>(v));
    v = fmax(v, dpp_f64<0x141>(v));
    v = fmax(v, dpp_f64<0x140>(v));
    return fmax(fmax(fmax(lane_f64<0>(v), lane_f64<16>(v)), lane_f64<32>(v)), lane_f64<48>(v));
}

// every thread of the block gets the same sum (fixed order)
template <int NT>
__device__ inline double block_sum(double v, double *red)
{
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < NT / 64; i++) s += red[i];
    return s;
}

// the same with one barrier: `red` must not be read by a previous call that
// some thread may still be in (each call site uses its own array, and a
// barrier separates its successive uses)
template <int NT>
__device__ inline double block_sum_once(double v, double *red)
{
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < NT / 64; i++) s += red[i];
    return s;
}

template <int NT>
__device__ inline double block_max(double v, double *red)
{
    v = wave_max(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < NT / 64; i++) s = fmax(s, red[i]);
    return s;
}

// Householder vector of x = (alpha, x_2..) with sigma = |x_2..|^2:
// (I - t v v^T) x = beta e_1, v = (1, x_2.. * scale).
__device__ inline void house(double alpha, double sigma, double &t, double &beta, double &scale)
{
    t = 0.0;
    beta = alpha;
    scale = 0.0;
    if (sigma > 0.0) {
        const double nrm = sqrt(alpha * alpha + sigma);
        beta = alpha >= 0.0 ? -nrm : nrm;
        t = (beta - alpha) / beta;
        scale = 1.0 / (alpha - beta);
    }
}

// ---------------------------------------------------------------------------
// 1a. tridiagonalization in LDS (n <= LDS_N)

__global__ void __launch_bounds__(TRI_T) k_tridiag_lds(const double *__restrict__ Aall, double *__restrict__ Hall,
                                                       double *__restrict__ dall, double *__restrict__ eall,
                                                       double *__restrict__ tall, int n, int ldh)
{
    __shared__ double A[LDS_N * LDS_N];
    __shared__ double v[LDS_N], y[LDS_N];
    __shared__ double redA[TRI_T / 64], redB[TRI_T / 64];
    constexpr int NW = TRI_T / 64;
    const int64_t m = blockIdx.x;
    const double *__restrict__ G = Aall + m * n * n;
    double *__restrict__ H = Hall + m * n * ldh;
    double *__restrict__ d = dall + m * n;
    double *__restrict__ e = eall + m * n;
    double *__restrict__ tau = tall + m * n;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // A = (G + G^T) / 2: the Grams are symmetric up to the order of their sums
    for (int c = wave; c < n; c += NW)
        for (int r = lane; r < n; r += 64) A[c * n + r] = 0.5 * (G[(int64_t)c * n + r] + G[(int64_t)r * n + c]);
    __syncthreads();
    for (int j = 0; j + 2 < n; j++) {
        const int r0 = j + 1, tt = n - r0;
        double sg = 0.0;
        for (int r = r0 + 1 + tid; r < n; r += TRI_T) {
            const double x = A[j * n + r];
            sg += x * x;
        }
        const double sigma = block_sum_once<TRI_T>(sg, redA);
        double t, beta, scale;
        house(A[j * n + r0], sigma, t, beta, scale);
        for (int r = r0 + tid; r < n; r += TRI_T) {
            const double vr = r == r0 ? 1.0 : A[j * n + r] * scale;
            v[r] = vr;
            y[r] = 0.0;
            if (r > r0) H[(int64_t)j * ldh + r] = vr;
        }
        if (tid == 0) {
            d[j] = A[j * n + j];
            e[j] = beta;
            tau[j] = t;
        }
        __syncthreads();
        if (t == 0.0) continue;   // uniform: the column is already reduced
        // y = A22 v: a wave takes one 64-row chunk and every ngrp-th column
        const int nch = (tt + 63) >> 6, ngrp = NW / nch;
        const int ch = wave % nch, grp = wave / nch;
        if (grp < ngrp) {
            const int r = r0 + ch * 64 + lane;
            const bool ok = r < n;
            double acc = 0.0;
            int c = r0 + grp;
            for (; c + 3 * ngrp < n; c += 4 * ngrp) {
                double h[4];
#pragma unroll
                for (int u = 0; u < 4; u++) h[u] = ok ? A[(c + u * ngrp) * n + r] : 0.0;
#pragma unroll
                for (int u = 0; u < 4; u++) acc += h[u] * v[c + u * ngrp];
            }
            for (; c < n; c += ngrp) acc += (ok ? A[c * n + r] : 0.0) * v[c];
            if (ok) atomicAdd(&y[r], acc);
        }
        __syncthreads();
        // p = t y, w = p - (t/2)(p.v) v
        double pv = 0.0;
        for (int r = r0 + tid; r < n; r += TRI_T) pv += y[r] * v[r];
        const double K = 0.5 * t * t * block_sum_once<TRI_T>(pv, redB);
        // A22 -= v w^T + w v^T (both triangles), w = t y - K v formed where it is
        // used, every element by the same explicit fma
        if (grp < ngrp) {
            const int r = r0 + ch * 64 + lane;
            if (r < n) {
                const double vr = v[r], wr = fma(t, y[r], -K * vr);
                int c = r0 + grp;
                for (; c + 3 * ngrp < n; c += 4 * ngrp) {
                    double h[4];
#pragma unroll
                    for (int u = 0; u < 4; u++) h[u] = A[(c + u * ngrp) * n + r];
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        const int cc = c + u * ngrp;
                        const double vc = v[cc], wc = fma(t, y[cc], -K * vc);
                        A[cc * n + r] = h[u] - (vr * wc + wr * vc);
                    }
                }
                for (; c < n; c += ngrp) {
                    const double vc = v[c], wc = fma(t, y[c], -K * vc);
                    A[c * n + r] -= vr * wc + wr * vc;
                }
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        if (n >= 2) {
            d[n - 2] = A[(n - 2) * n + n - 2];
            e[n - 2] = A[(n - 2) * n + n - 1];
            tau[n - 2] = 0.0;
        }
        d[n - 1] = A[(n - 1) * n + n - 1];
        e[n - 1] = 0.0;
        tau[n - 1] = 0.0;
    }
}

// ---------------------------------------------------------------------------
// 1b. tridiagonalization in global memory (n > LDS_N), one panel of nb columns
// per launch pair: k_tri_panel (one workgroup per matrix) reduces the panel's
// columns, each brought up to date with the panel's earlier reflectors and
// its matrix-vector product corrected from the panel's V, W (the panel form
// of LAPACK's dsytrd/dlatrd), and publishes V, W; k_tri_trail then applies
// the rank-2nb update to the trailing matrix as 16 x 16 FP64 MFMA tiles over
// the whole device.  H has an even leading dimension ldh, so the panel's
// products read rows in aligned pairs (16-byte loads, twice the rate one
// workgroup gets with 8-byte loads); a pair reaching above the rows in use is
// read whole and its extra row masked.

__global__ void __launch_bounds__(256) k_tri_init(const double *__restrict__ Aall, double *__restrict__ Hall, int n,
                                                  int ldh)
{
    const int64_t m = blockIdx.y;
    const double *__restrict__ G = Aall + m * n * n;
    double *__restrict__ H = Hall + m * n * ldh;
    const int64_t total = (int64_t)n * n;
    for (int64_t x = (int64_t)blockIdx.x * 256 + threadIdx.x; x < total; x += (int64_t)gridDim.x * 256) {
        const int c = (int)(x / n), r = (int)(x - (int64_t)c * n);
        H[(int64_t)c * ldh + r] = 0.5 * (G[(int64_t)c * n + r] + G[(int64_t)r * n + c]);
    }
}

__global__ void __launch_bounds__(TRG_T) k_tri_panel(double *__restrict__ Hall, double *__restrict__ dall,
                                                     double *__restrict__ eall, double *__restrict__ tall,
                                                     double *__restrict__ VWall, int n, int ldh, int j0, int np)
{
    __shared__ double V[VW_CAP], W[VW_CAP];   // panel column q at [q * n]
    __shared__ double xv[EIG_NMAX], y[EIG_NMAX];
    __shared__ double s1[PANEL_MAX], s2[PANEL_MAX];
    __shared__ double redA[TRG_T / 64], redB[TRG_T / 64];
    constexpr int NW = TRG_T / 64;
    constexpr int U = 16;   // columns of loads in flight per wave
    const int64_t m = blockIdx.x;
    double *__restrict__ H = Hall + m * n * ldh;
    double *__restrict__ d = dall + m * n;
    double *__restrict__ e = eall + m * n;
    double *__restrict__ tau = tall + m * n;
    double *__restrict__ VW = VWall + m * 2 * (int64_t)n * PANEL_MAX;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = 0; i < np; i++) {
        const int j = j0 + i, r0 = j + 1;
        // column j brought up to date with the panel's earlier reflectors (rows
        // j..n-1), its squared norm below row r0 summed on the way; the sum's
        // barrier also publishes xv and the cleared y
        double sg = 0.0;
        for (int r = j + tid; r < n; r += TRG_T) {
            double x = H[(int64_t)j * ldh + r];
            for (int q = 0; q < i; q++) x -= V[q * n + r] * W[q * n + j] + W[q * n + r] * V[q * n + j];
            xv[r] = x;
            y[r] = 0.0;
            if (r > r0) sg += x * x;
        }
        const double sigma = block_sum_once<TRG_T>(sg, redA);
        double t, beta, scale;
        house(xv[r0], sigma, t, beta, scale);
        for (int r = tid; r < n; r += TRG_T) {
            const double vr = r < r0 ? 0.0 : (r == r0 ? 1.0 : xv[r] * scale);
            V[i * n + r] = vr;
            if (r > r0) H[(int64_t)j * ldh + r] = vr;
        }
        if (tid == 0) {
            d[j] = xv[j];
            e[j] = beta;
            tau[j] = t;
        }
        __syncthreads();
        if (t != 0.0) {   // uniform
            // y = A22 v with A22 as of the panel start: a wave takes a 128-row chunk
            // (a pair per lane) and every ngrp-th column, U columns of loads at a time
            const int rb = r0 & ~1, tt = n - rb;
            const int nch = (tt + 127) >> 7, ngrp = NW / nch;
            const int ch = wave % nch, grp = wave / nch;
            if (grp < ngrp) {
                const int r = rb + ch * 128 + 2 * lane;
                const bool ok = r < n;   // r + 1 <= ldh - 1: the pair is in the column
                const double *__restrict__ Hr = H + r;
                double a0 = 0.0, a1 = 0.0;
                int c = r0 + grp;
                for (; c + (U - 1) * ngrp < n; c += U * ngrp) {
                    double2 h[U];
#pragma unroll
                    for (int u = 0; u < U; u++)
                        h[u] = ok ? *(const double2 *)(Hr + (int64_t)(c + u * ngrp) * ldh) : make_double2(0.0, 0.0);
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        const double vc = V[i * n + c + u * ngrp];
                        a0 += h[u].x * vc;
                        a1 += h[u].y * vc;
                    }
                }
                for (; c < n; c += ngrp) {
                    const double2 h = ok ? *(const double2 *)(Hr + (int64_t)c * ldh) : make_double2(0.0, 0.0);
                    const double vc = V[i * n + c];
                    a0 += h.x * vc;
                    a1 += h.y * vc;
                }
                if (ok && r >= r0) {
                    if (ngrp == 1) y[r] = a0;
                    else atomicAdd(&y[r], a0);
                }
                if (ok && r + 1 < n) {
                    if (ngrp == 1) y[r + 1] = a1;
                    else atomicAdd(&y[r + 1], a1);
                }
            }
            // s1_q = W_q . v, s2_q = V_q . v for the panel's earlier reflectors
            for (int q = wave; q < 2 * i; q += NW) {
                const double *X = q < i ? &W[q * n] : &V[(q - i) * n];
                double s = 0.0;
                for (int r = r0 + lane; r < n; r += 64) s += X[r] * V[i * n + r];
                s = wave_sum(s);
                if (lane == 0) {
                    if (q < i) s1[q] = s;
                    else s2[q - i] = s;
                }
            }
        }
        __syncthreads();
        // p = t (A22 v - V s1 - W s2); w = p - (t/2)(p.v) v
        double pv = 0.0;
        for (int r = r0 + tid; r < n; r += TRG_T) {
            double p = 0.0;
            if (t != 0.0) {
                p = y[r];
                for (int q = 0; q < i; q++) p -= V[q * n + r] * s1[q] + W[q * n + r] * s2[q];
                p *= t;
            }
            y[r] = p;
            pv += p * V[i * n + r];
        }
        const double K = 0.5 * t * block_sum_once<TRG_T>(pv, redB);
        for (int r = tid; r < n; r += TRG_T) W[i * n + r] = r < r0 ? 0.0 : y[r] - K * V[i * n + r];
        __syncthreads();
    }
    // the panel's V, W for the trailing update
    for (int x = tid; x < np * n; x += TRG_T) {
        VW[x] = V[x];
        VW[(int64_t)PANEL_MAX * n + x] = W[x];
    }
}

// trailing block [b0, n)^2 -= V W^T + W V^T (b0 = j0 + np) as 16 x 16 tiles,
// D = H - [V W] [W V]^T on FP64 MFMA (K = 2 np <= 32): one tile per wave;
// lane l holds A(l & 15, l >> 4), B(l >> 4, l & 15) and D rows (l >> 4) + 4r
// of column l & 15
__global__ void __launch_bounds__(256) k_tri_trail(double *__restrict__ Hall, const double *__restrict__ VWall, int n,
                                                   int ldh, int b0, int np, int ntile)
{
    const int64_t m = blockIdx.y;
    const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (tile >= ntile * ntile) return;
    double *__restrict__ H = Hall + m * n * ldh;
    const double *__restrict__ Vg = VWall + m * 2 * (int64_t)n * PANEL_MAX;
    const double *__restrict__ Wg = Vg + (int64_t)PANEL_MAX * n;
    const int lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
    const int tb = b0 & ~15;
    const int R0 = tb + 16 * (tile % ntile), C0 = tb + 16 * (tile / ntile);
    const int ar = R0 + li, cj = C0 + li;
    const bool arow = ar >= b0 && ar < n, col = cj >= b0 && cj < n;
    const int K = 2 * np;
    double *__restrict__ Hc = H + (int64_t)cj * ldh + R0 + lg;
    f64x4 acc;
#pragma unroll
    for (int r = 0; r < 4; r++) acc[r] = (col && R0 + lg + 4 * r < n) ? Hc[4 * r] : 0.0;
#pragma unroll
    for (int kb = 0; kb < PANEL_MAX / 2; kb++) {
        const int k = 4 * kb + lg;
        const bool kin = k < K;
        const double a = (arow && kin) ? -(k < np ? Vg[(int64_t)k * n + ar] : Wg[(int64_t)(k - np) * n + ar]) : 0.0;
        const double b = (col && kin) ? (k < np ? Wg[(int64_t)k * n + cj] : Vg[(int64_t)(k - np) * n + cj]) : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int row = R0 + lg + 4 * r;
        if (col && row >= b0 && row < n) Hc[4 * r] = acc[r];
    }
}

__global__ void k_tri_finish(const double *__restrict__ Hall, double *__restrict__ dall, double *__restrict__ eall,
                             double *__restrict__ tall, int n, int ldh, int count)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= count) return;
    const double *__restrict__ H = Hall + (int64_t)m * n * ldh;
    dall[(int64_t)m * n + n - 2] = H[(int64_t)(n - 2) * ldh + n - 2];
    eall[(int64_t)m * n + n - 2] = H[(int64_t)(n - 2) * ldh + n - 1];
    dall[(int64_t)m * n + n - 1] = H[(int64_t)(n - 1) * ldh + n - 1];
    tall[(int64_t)m * n + n - 2] = 0.0;
    eall[(int64_t)m * n + n - 1] = 0.0;
    tall[(int64_t)m * n + n - 1] = 0.0;
}

// ---------------------------------------------------------------------------
// 2a. leaf segments: implicit symmetric QR with Wilkinson shifts

__device__ inline bool negligible(double e, double a, double b)
{
    return fabs(e) <= EPS * (fabs(a) + fabs(b)) || fabs(e) <= TINY;
}

// lane l's value of v, for a wave-uniform l
__device__ inline double rdlane(double v, int l)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// v with lane l (wave-uniform) set to x
__device__ inline double wrlane(double v, double x, int l)
{
    return (int)(threadIdx.x & 63) == l ? x : v;
}

// 1/sqrt(q), q > 0 normal: the hardware estimate refined by two Newton steps
// y <- y (3 - q y^2) / 2 (each squares the relative error)
__device__ inline double rsqrt_nr(double q)
{
    double y = __builtin_amdgcn_rsq(q);
    const double h = 0.5 * q;
    y = y * fma(-h * y, y, 1.5);
    y = y * fma(-h * y, y, 1.5);
    return y;
}

// rotation [c -s; s c]^T (x, z) = (r, 0)
__device__ inline void givens(double x, double z, double &c, double &s, double &r)
{
    if (z == 0.0) {
        c = 1.0;
        s = 0.0;
        r = x;
        return;
    }
    const double q = x * x + z * z;
    if (q < 1e-280 || q > 1e280) {   // rescale to keep the squares representable
        const double sc = fmax(fabs(x), fabs(z));
        const double xs = x / sc, zs = z / sc;
        r = sc * sqrt(xs * xs + zs * zs);
        c = x / r;
        s = -z / r;
        return;
    }
    const double ir = rsqrt_nr(q);
    r = q * ir;
    c = x * ir;
    s = -z * ir;
}

// One wave runs the chase with the segment's d, e held one element per lane
// (reads and writes by lane index), the rotations gathered the same way and
// published to LDS; wave 1 applies the previous chase's rotations to the rows
// of Z in LDS meanwhile.
__global__ void __launch_bounds__(LEAF_T) k_eig_leaf(const double *__restrict__ dall, const double *__restrict__ eall,
                                                     double *__restrict__ lamw_all, double *__restrict__ Zall, int n,
                                                     SegTable seg, int *__restrict__ info)
{
    __shared__ double Z[SEG * SEG];
    __shared__ double dd[SEG];
    __shared__ double rc[2][SEG], rs[2][SEG];
    __shared__ int clo[2], clen[2];
    __shared__ int rank[SEG];
    const int64_t m = blockIdx.y;
    const int a = seg.a[blockIdx.x], len = seg.len[blockIdx.x];
    const double *__restrict__ d = dall + m * n;
    const double *__restrict__ e = eall + m * n;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // wave 0: lane i holds d_i, e_i of the torn segment
    double dv = 0.0, ev = 0.0;
    if (wave == 0 && lane < len) {
        dv = d[a + lane];
        if (lane == 0 && a > 0) dv -= fabs(e[a - 1]);
        if (lane == len - 1 && a + len < n) dv -= fabs(e[a + len - 1]);
        ev = lane + 1 < len ? e[a + lane] : 0.0;
    }
    for (int c = wave; c < len; c += LEAF_T / 64)
        for (int r = lane; r < len; r += 64) Z[c * len + r] = r == c ? 1.0 : 0.0;
    if (tid == 0) {
        clen[0] = 0;
        clen[1] = 0;
    }
    __syncthreads();
    int hi = len - 1, chases = 0;   // wave-0 state, uniform across its lanes
    for (int k = 0;; k++) {
        const int b = k & 1;
        if (wave == 0) {
            int L = 0, lo = 0;
            // negligible off-diagonals, found by every lane at once
            const double dn = __shfl_down(dv, 1);
            const bool ng = lane + 1 < len && negligible(ev, dv, dn);
            if (ng) ev = 0.0;
            const unsigned long long neg = __ballot(ng);
            const unsigned long long below = hi > 0 ? (~neg) & ((1ull << hi) - 1ull) : 0ull;
            hi = below ? 64 - __clzll(below) : 0;   // highest i with e_{i-1} not negligible
            if (hi > 0 && chases >= 30 * len) {     // no convergence: give up on this segment
                if (lane == 0) atomicOr(info + m, 1);
                hi = 0;
            }
            if (hi > 0) {
                const unsigned long long nb = neg & ((1ull << (hi - 1)) - 1ull);
                lo = nb ? 64 - __clzll(nb) : 0;      // just above the last negligible e below hi-1
                // Wilkinson shift: eigenvalue of the trailing 2x2 nearer d[hi]
                const double p = rdlane(dv, hi - 1), q = rdlane(ev, hi - 1), w = rdlane(dv, hi);
                const double dl = 0.5 * (p - w);
                const double hyp = sqrt(dl * dl + q * q);
                const double mu = w - q * q / (dl + (dl >= 0.0 ? hyp : -hyp));
                double ai = rdlane(dv, lo), bi = rdlane(ev, lo);
                double x = ai - mu, z = bi;
                double rcv = 0.0, rsv = 0.0;
                double ai1 = rdlane(dv, lo + 1), bnext = lo + 1 < hi ? rdlane(ev, lo + 1) : 0.0;
                for (int i = lo; i < hi; i++) {
                    // next pair, read ahead of the dependent chain
                    const double ai2 = i + 2 <= hi ? rdlane(dv, i + 2) : 0.0;
                    const double bn2 = i + 2 < hi ? rdlane(ev, i + 2) : 0.0;
                    double cs, sn, r;
                    givens(x, z, cs, sn, r);
                    if (i > lo) ev = wrlane(ev, r, i - 1);
                    const double cc = cs * cs, ss = sn * sn, c2 = cs * sn;
                    dv = wrlane(dv, cc * ai - 2.0 * c2 * bi + ss * ai1, i);
                    const double na1 = ss * ai + 2.0 * c2 * bi + cc * ai1;
                    const double nbv = c2 * (ai - ai1) + (cc - ss) * bi;
                    ev = wrlane(ev, nbv, i);
                    rcv = wrlane(rcv, cs, i - lo);
                    rsv = wrlane(rsv, sn, i - lo);
                    x = nbv;
                    z = -sn * bnext;
                    bi = cs * bnext;
                    ai = na1;
                    ai1 = ai2;
                    bnext = bn2;
                }
                dv = wrlane(dv, ai, hi);
                L = hi - lo;
                chases++;
                if (lane < L) {
                    rc[b][lane] = rcv;
                    rs[b][lane] = rsv;
                }
            }
            if (lane == 0) {
                clo[b] = lo;
                clen[b] = L;
            }
        } else if (k > 0) {   // rotations of the previous chase, one Z row per thread
            const int pb = b ^ 1, L = clen[pb], lo = clo[pb];
            const int r = tid - 64;
            if (L > 0 && r < len) {
                double zc = Z[lo * len + r];
                for (int u = 0; u < L; u++) {
                    const double cs = rc[pb][u], sn = rs[pb][u];
                    const double z1 = Z[(lo + u + 1) * len + r];
                    Z[(lo + u) * len + r] = cs * zc - sn * z1;
                    zc = sn * zc + cs * z1;
                }
                Z[(lo + L) * len + r] = zc;
            }
        }
        __syncthreads();
        if (clen[b] == 0) break;
    }
    if (wave == 0 && lane < len) dd[lane] = dv;
    __syncthreads();
    // ascending order (ties by index)
    for (int i = tid; i < len; i += LEAF_T) {
        const double x = dd[i];
        int rk = 0;
        for (int j = 0; j < len; j++) rk += (dd[j] < x) || (dd[j] == x && j < i);
        rank[i] = rk;
        lamw_all[m * n + a + rk] = x;
    }
    __syncthreads();
    double *__restrict__ Zo = Zall + m * n * n;
    for (int c = wave; c < len; c += LEAF_T / 64) {
        const int oc = a + rank[c];
        for (int r = lane; r < len; r += 64) Zo[(int64_t)oc * n + a + r] = Z[c * len + r];
    }
}

// ---------------------------------------------------------------------------
// 2b. merge of two sibling segments

// number of x[0..len) < v (strict) or <= v, x ascending
__device__ inline int count_below(const double *x, int len, double v, bool inclusive)
{
    int lo = 0, hi = len;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const bool below = inclusive ? x[mid] <= v : x[mid] < v;
        if (below) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// 1/x: the hardware estimate refined by two Newton steps y <- y (2 - x y)
__device__ inline double rcp_nr(double x)
{
    double y = __builtin_amdgcn_rcp(x);
    y = fma(fma(-x, y, 1.0), y, y);
    y = fma(fma(-x, y, 1.0), y, y);
    return y;
}

// root j of 1 + rho sum z_i^2 / (d_i - lam) over dk ascending: lam = dk[org] + tau
__device__ inline void secular_root(const double *dk, const double *z2, int K, double rho, int j, int &org,
                                    double &tau_out)
{
    const bool last = j == K - 1;
    double tlo, thi, D1 = 0.0, D2 = 0.0;
    int o;
    if (!last) {
        const double mid = 0.5 * (dk[j] + dk[j + 1]);
        double f = 1.0;
        // only the sign of f at the midpoint is used (mid lies strictly between
        // two distinct poles, so no divisor is zero)
        for (int i = 0; i < K; i++) f += rho * z2[i] * rcp_nr(dk[i] - mid);
        if (f >= 0.0) {
            o = j;
            tlo = 0.0;
            thi = mid - dk[j];
        } else {
            o = j + 1;
            tlo = mid - dk[j + 1];
            thi = 0.0;
        }
        D1 = dk[j] - dk[o];
        D2 = dk[j + 1] - dk[o];
    } else {
        o = K - 1;
        double zz = 0.0;
        for (int i = 0; i < K; i++) zz += z2[i];
        tlo = 0.0;
        thi = rho * zz;
    }
    const double dorg = dk[o];
    double tau = 0.5 * (tlo + thi);
    for (int it = 0; it < 64; it++) {
        double f = 1.0, fabs_sum = 1.0, psi = 0.0, dpsi = 0.0, phi = 0.0, dphi = 0.0;
        for (int i = 0; i < K; i++) {
            const double delta = (dk[i] - dorg) - tau;
            const double inv = rcp_nr(delta);
            const double t = rho * z2[i] * inv;
            const double tp = t * inv;
            f += t;
            fabs_sum += fabs(t);
            if (i <= j) {
                psi += t;
                dpsi += tp;
            } else {
                phi += t;
                dphi += tp;
            }
        }
        if (f > 0.0) thi = tau;
        else tlo = tau;
        if (fabs(f) <= 8.0 * EPS * fabs_sum || thi - tlo <= 4.0 * EPS * fmax(fabs(tlo), fabs(thi))) break;
        const double dj = (dk[j] - dorg) - tau;
        const double b1 = dpsi * dj * dj, a1 = psi - dpsi * dj;
        double tn = 0.5 * (tlo + thi);
        if (last) {
            const double c = 1.0 + a1;
            if (c > 0.0) {
                const double r = b1 / c;
                if (r > tlo && r < thi) tn = r;
            }
        } else {
            const double dj1 = (dk[j + 1] - dorg) - tau;
            const double b2 = dphi * dj1 * dj1, a2 = phi - dphi * dj1;
            const double A = 1.0 + a1 + a2;
            const double B = A * (D1 + D2) + b1 + b2;
            const double C = A * D1 * D2 + b1 * D2 + b2 * D1;
            const double sq = sqrt(fmax(B * B - 4.0 * A * C, 0.0));
            double r1, r2;
            if (B >= 0.0) {
                r1 = A != 0.0 ? (B + sq) / (2.0 * A) : NAN;
                r2 = B + sq != 0.0 ? 2.0 * C / (B + sq) : NAN;
            } else {
                r1 = 2.0 * C / (B - sq);
                r2 = A != 0.0 ? (B - sq) / (2.0 * A) : NAN;
            }
            if (r1 > tlo && r1 < thi) tn = r1;
            if (r2 > tlo && r2 < thi) tn = r2;
        }
        if (tn == tau) break;
        tau = tn;
    }
    org = o;
    tau_out = tau;
}

__global__ void __launch_bounds__(MRG_T) k_eig_merge(const double *__restrict__ eall, double *__restrict__ lamw_all,
                                                     const double *__restrict__ Zin_all, double *__restrict__ Yall,
                                                     int n, MergeTable mt, int top_final)
{
    __shared__ double dl[EIG_NMAX], zl[EIG_NMAX];   // block order, later reused
    __shared__ double ds[EIG_NMAX], zs[EIG_NMAX];   // ascending d
    __shared__ double dk[EIG_NMAX], z2[EIG_NMAX], tau[EIG_NMAX], zh[EIG_NMAX], invn[EIG_NMAX];
    __shared__ double rcs[EIG_NMAX], rsn[EIG_NMAX];
    __shared__ int src[EIG_NMAX], ndpos[EIG_NMAX], org[EIG_NMAX], code[EIG_NMAX], outcode[EIG_NMAX];
    __shared__ int rp[EIG_NMAX], rq[EIG_NMAX];
    __shared__ int s_nrot, s_K, s_ndef;
    __shared__ double red[MRG_T / 64];
    constexpr int NW = MRG_T / 64;
    const int64_t m = blockIdx.y;
    const int a = mt.a[blockIdx.x], n1 = mt.n1[blockIdx.x], n2 = mt.n2[blockIdx.x];
    const int s = n1 + n2;
    const int s_out = s == n ? top_final : s;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const double bnd = eall[m * n + a + n1 - 1];
    const double rho = fabs(bnd), sg = bnd >= 0.0 ? 1.0 : -1.0;
    double *__restrict__ lamw = lamw_all + m * n + a;
    const double *__restrict__ Zin = Zin_all + m * n * n;
    for (int i = tid; i < s; i += MRG_T) {
        dl[i] = lamw[i];
        zl[i] = i < n1 ? Zin[(int64_t)(a + i) * n + a + n1 - 1] : sg * Zin[(int64_t)(a + i) * n + a + n1];
    }
    __syncthreads();
    // both halves ascending: merged rank by binary search (ties: left first)
    for (int i = tid; i < s; i += MRG_T) {
        const double x = dl[i];
        const int rk = i < n1 ? i + count_below(dl + n1, n2, x, false) : (i - n1) + count_below(dl, n1, x, true);
        ds[rk] = x;
        zs[rk] = zl[i];
        src[rk] = i;
    }
    __syncthreads();
    double mx = 0.0, zz = 0.0;
    for (int i = tid; i < s; i += MRG_T) {
        mx = fmax(mx, fabs(ds[i]));
        zz += zs[i] * zs[i];
    }
    mx = block_max<MRG_T>(mx, red);
    zz = block_sum<MRG_T>(zz, red);
    const double tol = 8.0 * EPS * fmax(mx, rho * zz);
    if (tid == 0) {
        // deflation: negligible rho z_i, then close pairs rotated together
        int last = -1, nr = 0;
        for (int i = 0; i < s; i++) {
            if (fabs(rho * zs[i]) <= tol) {
                ndpos[i] = -1;
                continue;
            }
            ndpos[i] = 0;
            if (last >= 0 && ds[i] - ds[last] <= 2.0 * tol) {
                const double zi = zs[last], zj = zs[i];
                const double r = sqrt(zi * zi + zj * zj);
                const double c = zj / r, sn = -zi / r;
                if (fabs((ds[i] - ds[last]) * c * sn) <= tol) {
                    rp[nr] = src[last];
                    rq[nr] = src[i];
                    rcs[nr] = c;
                    rsn[nr] = sn;
                    nr++;
                    const double tt = ds[last] * c * c + ds[i] * sn * sn;
                    ds[i] = ds[last] * sn * sn + ds[i] * c * c;
                    ds[last] = tt;
                    zs[i] = r;
                    zs[last] = 0.0;
                    ndpos[last] = -1;
                }
            }
            last = i;
        }
        int K = 0, nd = 0;
        for (int i = 0; i < s; i++) {
            if (ndpos[i] >= 0) {
                ndpos[i] = K;
                dk[K] = ds[i];
                z2[K] = zs[i];   // z itself for now
                K++;
            } else {
                dl[nd] = ds[i];  // deflated eigenvalues (keys), block-order array reused
                code[nd] = i;
                nd++;
            }
        }
        s_nrot = nr;
        s_K = K;
        s_ndef = nd;
    }
    __syncthreads();
    const int K = s_K, ndef = s_ndef, nrot = s_nrot;
    // zl <- z of the non-deflated (signs), z2 <- z^2
    for (int i = tid; i < K; i += MRG_T) {
        zl[i] = z2[i];
        z2[i] = z2[i] * z2[i];
    }
    __syncthreads();
    for (int j = tid; j < K; j += MRG_T) {
        int o;
        double t;
        secular_root(dk, z2, K, rho, j, o, t);
        org[j] = o;
        tau[j] = t;
    }
    __syncthreads();
    // Gu-Eisenstat: z_i from the computed roots
    for (int i = tid; i < K; i += MRG_T) {
        const double di = dk[i];
        double val = ((dk[org[K - 1]] - di) + tau[K - 1]) / rho;
        for (int j = 0; j < K - 1; j++) {
            const double num = (dk[org[j]] - di) + tau[j];
            const double den = j < i ? dk[j] - di : dk[j + 1] - di;
            val *= num / den;
        }
        zh[i] = copysign(sqrt(fmax(val, 0.0)), zl[i]);
    }
    __syncthreads();
    for (int j = tid; j < K; j += MRG_T) {
        const double dj = dk[org[j]], tj = tau[j];
        double ss = 0.0;
        for (int i = 0; i < K; i++) {
            const double x = zh[i] / ((dk[i] - dj) - tj);
            ss += x * x;
        }
        invn[j] = 1.0 / sqrt(ss);
        dl[ndef + j] = dj + tj;
        code[ndef + j] = s + j;
    }
    __syncthreads();
    // all s eigenvalues ascending: rank by counting (ties by position); a
    // slot no item claims (NaN input) stays -1 and gives a zero column
    for (int p = tid; p < s; p += MRG_T) outcode[p] = -1;
    __syncthreads();
    for (int p = tid; p < s; p += MRG_T) {
        const double x = dl[p];
        int rk = 0;
        for (int q = 0; q < s; q++) rk += (dl[q] < x) || (dl[q] == x && q < p);
        outcode[rk] = code[p];
        lamw[rk] = x;
    }
    __syncthreads();
    // Y (s x s_out, rows in block order): the last s_out eigenvectors of D + rho z z^T
    double *__restrict__ Y = Yall + m * n * n;
    for (int col = wave; col < s_out; col += NW) {
        const int cd = outcode[s - s_out + col];
        double *__restrict__ Yc = Y + (int64_t)(a + col) * n + a;
        if (cd < s) {
            for (int i = lane; i < s; i += 64) Yc[src[i]] = i == cd ? 1.0 : 0.0;   // cd < 0: zeros
        } else {
            const int j = cd - s;
            const double dj = dk[org[j]], tj = tau[j], inn = invn[j];
            for (int i = lane; i < s; i += 64) {
                const int p = ndpos[i];
                Yc[src[i]] = p < 0 ? 0.0 : zh[p] / ((dk[p] - dj) - tj) * inn;
            }
        }
    }
    if (nrot > 0) {
        __syncthreads();
        // the deflation rotations, last first, on the rows of each column
        for (int col = tid; col < s_out; col += MRG_T) {
            double *__restrict__ Yc = Y + (int64_t)(a + col) * n + a;
            for (int r = nrot - 1; r >= 0; r--) {
                const int p = rp[r], q = rq[r];
                const double c = rcs[r], sn = rsn[r];
                const double yp = Yc[p], yq = Yc[q];
                Yc[p] = c * yp - sn * yq;
                Yc[q] = sn * yp + c * yq;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// 3. back-transformation V = H_0 H_1 ... H_{n-3} Z for a slice of cb <= 64
// columns, reflectors applied in blocks of nk (compact WY form, LAPACK's
// dlarft/dlarfb): the block's product is I - Vb T Vb^T with T upper
// triangular (T[q][q] = tau_q, T[p][q] = -tau_q sum_s T[p][s] (Vb^T Vb)[s][q]),
// so Z <- Z - Vb (T (Vb^T Z)), blocks from the last reflector to the first.
// The slice is row-major in LDS (lanes over its columns, waves over rows).

constexpr int ZS_CAP = 8192;   // doubles of the slice
constexpr int VB_CAP = 8192;   // doubles of a block's reflectors
constexpr int NBK_MAX = 16;

__global__ void __launch_bounds__(BACK_T) k_eig_back(const double *__restrict__ Hall, const double *__restrict__ tall,
                                                     const double *__restrict__ Zall, const double *__restrict__ lamw_all,
                                                     double *__restrict__ Aall, double *__restrict__ lam_all, int n,
                                                     int ldh, int top, int zcol0, int cb, int nbk)
{
    __shared__ double Zs[ZS_CAP];   // Zs[r * cb + c]
    __shared__ double Vb[VB_CAP];   // Vb[q * n + r]
    __shared__ double Gm[NBK_MAX][NBK_MAX], Tm[NBK_MAX][NBK_MAX];
    __shared__ double P[NBK_MAX][64], WZ[NBK_MAX][64];
    __shared__ double taus[NBK_MAX];
    constexpr int NW = BACK_T / 64;
    const int64_t m = blockIdx.y;
    const int c0 = blockIdx.x * cb, nc = min(cb, top - c0);
    const double *__restrict__ H = Hall + m * n * ldh;
    const double *__restrict__ tau = tall + m * n;
    const double *__restrict__ Z = Zall + m * n * n;
    double *__restrict__ A = Aall + m * n * n;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (blockIdx.x == 0)
        for (int i = tid; i < n; i += BACK_T) lam_all[m * n + i] = lamw_all[m * n + i];
    for (int c = wave; c < nc; c += NW)
        for (int r = lane; r < n; r += 64) Zs[r * cb + c] = Z[(int64_t)(zcol0 + c0 + c) * n + r];
    for (int jb = n - 3; jb >= 0; jb -= nbk) {
        const int ja = max(0, jb - nbk + 1), nk = jb - ja + 1, rlo = ja + 1;
        for (int x = tid; x < nk * n; x += BACK_T) {
            const int q = x / n, r = x - q * n, jq = ja + q;
            Vb[x] = r <= jq ? 0.0 : (r == jq + 1 ? 1.0 : H[(int64_t)jq * ldh + r]);
        }
        if (tid < nk) taus[tid] = tau[ja + tid];
        for (int x = tid; x < nk * 64; x += BACK_T) P[x / 64][x % 64] = 0.0;
        for (int x = tid; x < NBK_MAX * NBK_MAX; x += BACK_T) Gm[x / NBK_MAX][x % NBK_MAX] = 0.0;
        __syncthreads();
        // one FP64 MFMA pass over the rows: P = Vb^T Z and Gm = Vb^T Vb (16 x 16 per
        // tile, K = rows, 4 per step; each wave takes every NW-th step and adds its
        // partial tiles to LDS).  Lane l: A(l & 15, l >> 4), B(l >> 4, l & 15),
        // accumulator rows (l >> 4) + 4r of column l & 15.
        {
            const int li = lane & 15, lg = lane >> 4;
            const int ntc = (nc + 15) >> 4;
            f64x4 accP[4], accG;
#pragma unroll
            for (int u = 0; u < 4; u++) accP[u] = f64x4{0.0, 0.0, 0.0, 0.0};
            accG = f64x4{0.0, 0.0, 0.0, 0.0};
            for (int rb = rlo + 4 * wave; rb < n; rb += 4 * NW) {
                const int r = rb + lg;
                const bool rok = r < n;
                const double a = (rok && li < nk) ? Vb[li * n + r] : 0.0;
                const double bg = (rok && li < nk) ? Vb[li * n + r] : 0.0;
                accG = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bg, accG, 0, 0, 0);
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    if (u < ntc) {
                        const int c = 16 * u + li;
                        const double bz = (rok && c < nc) ? Zs[r * cb + c] : 0.0;
                        accP[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bz, accP[u], 0, 0, 0);
                    }
                }
            }
#pragma unroll
            for (int rr = 0; rr < 4; rr++) {
                const int q = lg + 4 * rr;
                if (q < nk) {
                    if (li < nk) atomicAdd(&Gm[q][li], accG[rr]);
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        const int c = 16 * u + li;
                        if (u < ntc && c < nc) atomicAdd(&P[q][c], accP[u][rr]);
                    }
                }
            }
        }
        __syncthreads();
        // T, column by column (wave 0; lanes over rows of T)
        if (wave == 0) {
            for (int q = 0; q < nk; q++) {
                double tv = 0.0;
                if (lane < q) {
                    for (int s2 = lane; s2 < q; s2++) tv += Tm[lane][s2] * Gm[s2][q];
                    tv *= -taus[q];
                }
                if (lane < q) Tm[lane][q] = tv;
                if (lane == q) Tm[q][q] = taus[q];
                if (lane > q && lane < nk) Tm[lane][q] = 0.0;
                __builtin_amdgcn_wave_barrier();
            }
        }
        __syncthreads();
        // WZ = T P
        for (int x = tid; x < nk * nc; x += BACK_T) {
            const int q = x / nc, c = x - q * nc;
            double sv = 0.0;
            for (int t2 = q; t2 < nk; t2++) sv += Tm[q][t2] * P[t2][c];
            WZ[q][c] = sv;
        }
        __syncthreads();
        // Z -= Vb WZ on FP64 MFMA: 16-row tiles per wave, K = nk in steps of 4
        {
            const int li = lane & 15, lg = lane >> 4;
            const int ntc = (nc + 15) >> 4;
            for (int R0 = (rlo & ~15) + 16 * wave; R0 < n; R0 += 16 * NW) {
                const int ar = R0 + li;
                double af[NBK_MAX / 4];
#pragma unroll
                for (int kb = 0; kb < NBK_MAX / 4; kb++) {
                    const int q = 4 * kb + lg;
                    af[kb] = (q < nk && ar >= rlo && ar < n) ? -Vb[q * n + ar] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    if (u >= ntc) continue;
                    const int c = 16 * u + li;
                    const bool cok = c < nc;
                    f64x4 acc;
#pragma unroll
                    for (int rr = 0; rr < 4; rr++) {
                        const int row = R0 + lg + 4 * rr;
                        acc[rr] = (cok && row < n) ? Zs[row * cb + c] : 0.0;
                    }
#pragma unroll
                    for (int kb = 0; kb < NBK_MAX / 4; kb++) {
                        const int q = 4 * kb + lg;
                        const double bv = (cok && q < nk) ? WZ[q][c] : 0.0;
                        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(af[kb], bv, acc, 0, 0, 0);
                    }
#pragma unroll
                    for (int rr = 0; rr < 4; rr++) {
                        const int row = R0 + lg + 4 * rr;
                        if (cok && row >= rlo && row < n) Zs[row * cb + c] = acc[rr];
                    }
                }
            }
        }
        __syncthreads();
    }
    for (int c = wave; c < nc; c += NW)
        for (int r = lane; r < n; r += 64) A[(int64_t)(n - top + c0 + c) * n + r] = Zs[r * cb + c];
}

struct Plan {
    SegTable leaves{};
    std::vector<MergeTable> levels;
};

// Every segment is split to the same depth (the smallest at which all leaves
// have <= SEG rows), so each merge level's two buffers hold every block: the
// levels alternate between Z0 and Z1 and a block that skipped a level would
// be left behind in the other one.
Plan make_plan(int n)
{
    Plan p;
    int depth = 0;
    while ((n + (1 << depth) - 1) >> depth > SEG) depth++;
    std::vector<std::vector<std::array<int, 3>>> lv(depth);
    std::function<void(int, int, int)> split = [&](int a, int len, int h) {
        if (h == 0) {
            TSVD_REQUIRE(p.leaves.count < MAX_PIECES, "eigensolver: too many segments");
            p.leaves.a[p.leaves.count] = a;
            p.leaves.len[p.leaves.count] = len;
            p.leaves.count++;
            return;
        }
        const int n1 = len / 2;
        split(a, n1, h - 1);
        split(a + n1, len - n1, h - 1);
        lv[h - 1].push_back({a, n1, len - n1});
    };
    split(0, n, depth);
    for (auto &l : lv) {
        MergeTable t{};
        for (auto &x : l) {
            TSVD_REQUIRE(t.count < MAX_PIECES, "eigensolver: too many merges");
            t.a[t.count] = x[0];
            t.n1[t.count] = x[1];
            t.n2[t.count] = x[2];
            t.count++;
        }
        p.levels.push_back(t);
    }
    return p;
}

}  // namespace

void eig_batched(hipStream_t st, double *A, int n, int count, double *lam, int top, void *ws, size_t ws_bytes,
                 GemmRunner &gemm, GemmProb *desc_dev, int desc_cap, int *info_host)
{
    if (count <= 0) return;
    TSVD_REQUIRE(n >= 1 && n <= EIG_NMAX, "eigensolver takes 1 <= n <= %d (got %d)", EIG_NMAX, n);
    if (top <= 0 || top > n) top = n;
    TSVD_REQUIRE(ws_bytes >= eig_workspace_bytes(n, count), "eigensolver workspace too small");
    const size_t nn = (size_t)n * n, c = (size_t)count;
    double *base = (double *)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
    const int ldh = n + (n & 1);   // even: aligned row pairs in every column
    double *H = base, *Z0 = H + c * (size_t)n * ldh, *Z1 = Z0 + c * nn, *Y = Z1 + c * nn;
    double *d = Y + c * nn, *e = d + c * n, *tau = e + c * n, *lamw = tau + c * n;
    int *info = (int *)(lamw + c * n);
    PROF("eig", st);
    TSVD_HIP(hipMemsetAsync(info, 0, sizeof(int) * c, st));
    {
    PROF("eig_tridiag", st);
    if (n <= LDS_N) {
        hipLaunchKernelGGL(k_tridiag_lds, dim3(count), dim3(TRI_T), 0, st, A, H, d, e, tau, n, ldh);
        TSVD_HIP(hipGetLastError());
    } else {
        const int nb = std::min(PANEL_MAX, VW_CAP / n);
        double *VW = Y;   // the merge workspace is free until the leaves are solved
        hipLaunchKernelGGL(k_tri_init, dim3(std::max(1, std::min(1024, (int)(nn / 256 / 8))), count), dim3(256), 0,
                           st, A, H, n, ldh);
        for (int j0 = 0; j0 + 2 < n; j0 += nb) {
            const int np = std::min(nb, n - 2 - j0);
            hipLaunchKernelGGL(k_tri_panel, dim3(count), dim3(TRG_T), 0, st, H, d, e, tau, VW, n, ldh, j0, np);
            const int b0 = j0 + np, ntile = (n - (b0 & ~15) + 15) >> 4;
            hipLaunchKernelGGL(k_tri_trail, dim3((ntile * ntile + 3) / 4, count), dim3(256), 0, st, H, VW, n, ldh, b0,
                               np, ntile);
        }
        hipLaunchKernelGGL(k_tri_finish, dim3((count + 63) / 64), dim3(64), 0, st, H, d, e, tau, n, ldh, count);
        TSVD_HIP(hipGetLastError());
    }
    }
    const Plan plan = make_plan(n);
    {
    PROF("eig_leaves", st);
    hipLaunchKernelGGL(k_eig_leaf, dim3(plan.leaves.count, count), dim3(LEAF_T), 0, st, d, e, lamw, Z0, n,
                       plan.leaves, info);
    TSVD_HIP(hipGetLastError());
    }
    double *zin = Z0, *zout = Z1;
    std::vector<GemmProb> probs;
    for (size_t L = 0; L < plan.levels.size(); L++) {
        const MergeTable &mt = plan.levels[L];
        {
        PROF("eig_merges", st);
        hipLaunchKernelGGL(k_eig_merge, dim3(mt.count, count), dim3(MRG_T), 0, st, e, lamw, zin, Y, n, mt, top);
        TSVD_HIP(hipGetLastError());
        }
        probs.clear();
        for (int mm = 0; mm < count; mm++) {
            const size_t off = (size_t)mm * nn;
            for (int x = 0; x < mt.count; x++) {
                const int a = mt.a[x], n1 = mt.n1[x], n2 = mt.n2[x];
                const int so = n1 + n2 == n ? top : n1 + n2;
                for (int half = 0; half < 2; half++) {
                    const int r = half ? a + n1 : a, len = half ? n2 : n1;
                    GemmProb p{};
                    p.A = zin + off + (size_t)r * n + r;
                    p.lda = n;
                    p.B = Y + off + (size_t)a * n + r;
                    p.ldb = n;
                    p.C = zout + off + (size_t)a * n + r;
                    p.ldc = n;
                    p.M = len;
                    p.N = so;
                    p.K = len;
                    p.alpha = 1.0;
                    p.beta = 0.0;
                    probs.push_back(p);
                }
            }
        }
        const GemmTypes t{DT::F64, DT::F64, DT::F64, DT::F64};
        if (desc_dev) gemm.run_into(st, probs.data(), (int)probs.size(), t, true, desc_dev, desc_cap, "eig_gemm");
        else gemm.run(st, probs, t, true, "eig_gemm");
        std::swap(zin, zout);
    }
    const int zcol0 = plan.levels.empty() ? n - top : 0;
    const int cb = std::min(64, ZS_CAP / n);
    const int nbk = std::min(NBK_MAX, VB_CAP / n);
    PROF("eig_back", st);
    hipLaunchKernelGGL(k_eig_back, dim3((top + cb - 1) / cb, count), dim3(BACK_T), 0, st, H, tau, zin, lamw, A, lam, n,
                       ldh, top, zcol0, cb, nbk);
    TSVD_HIP(hipGetLastError());
    if (info_host) {
        std::vector<int> h(c);
        TSVD_HIP(hipMemcpyAsync(h.data(), info, sizeof(int) * c, hipMemcpyDeviceToHost, st));
        TSVD_HIP(hipStreamSynchronize(st));
        int bad = 0;
        for (int x : h) bad |= x;
        *info_host = bad;
    }
}

void Eig::solve(hipStream_t st, double *A, int n, int count, double *w, int top)
{
    if (count <= 0) return;
    const size_t need = eig_workspace_bytes(n, count);
    ws_.ensure(need);
    eig_batched(st, A, n, count, w, top, ws_.p, need, gemm_);
}

}  // namespace tsvd

extern "C" int tsvd_eig_batched(const double *A, int32_t n, int32_t count, int32_t top, double *V, double *lam,
                                int32_t *info)
{
    return tsvd::guard([&] {
        TSVD_REQUIRE(A && V && lam && count >= 0, "null argument or negative count");
        if (count == 0) return;
        TSVD_REQUIRE(n >= 1 && n <= tsvd::EIG_NMAX, "eigensolver takes 1 <= n <= %d (got %d)", tsvd::EIG_NMAX, n);
        const size_t nn = (size_t)n * n * count;
        tsvd::DBuf<double> dA, dl;
        tsvd::DBuf<uint8_t> ws;
        dA.alloc(nn);
        dl.alloc((size_t)n * count);
        const size_t wb = tsvd::eig_workspace_bytes(n, count);
        ws.alloc(wb);
        tsvd::GemmRunner gemm;
        hipStream_t st = nullptr;
        TSVD_HIP(hipMemcpyAsync(dA.p, A, sizeof(double) * nn, hipMemcpyHostToDevice, st));
        int bad = 0;
        tsvd::eig_batched(st, dA.p, n, count, dl.p, top, ws.p, wb, gemm, nullptr, 0, &bad);
        TSVD_HIP(hipMemcpyAsync(V, dA.p, sizeof(double) * nn, hipMemcpyDeviceToHost, st));
        TSVD_HIP(hipMemcpyAsync(lam, dl.p, sizeof(double) * n * count, hipMemcpyDeviceToHost, st));
        TSVD_HIP(hipStreamSynchronize(st));
        if (info) *info = bad;
    });
}
