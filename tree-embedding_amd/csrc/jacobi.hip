// jacobi.hip -- batched FP64 symmetric eigensolver, one workgroup per matrix.
//
// Replaces the LAPACKE_dsyevd call of eigSVD (frpca.c:99-136) for the small
// l x l Grams (l = k + 5 <= 136) of the frPCA passes.  Parallel cyclic
// Jacobi: the matrix lives in LDS (one 64-wide wavefront group of 256
// threads), the eigenvector matrix in the caller's global array (it
// overwrites the input, which is read into LDS first).  Each sweep visits
// every (p, q) pair once in l' - 1 steps of l'/2 disjoint pairs (l' = l
// rounded up to even; round-robin tournament order, the extra player of an
// odd l sits out).  A step computes the l'/2 rotations from the current
// matrix (Golub and Van Loan's symmetric Schur 2x2: tau = (a_qq - a_pp) /
// 2 a_pq, t = sign(tau) / (|tau| + sqrt(1 + tau^2)), c = 1/sqrt(1 + t^2),
// s = t c), then applies them to rows, then to columns and to V.  Disjoint
// rotations commute, so a step is J^T A J for their product J.  Sweeps stop
// when the off-diagonal Frobenius norm is below 4e-16 of the whole norm.
// Output: eigenvalues ascending, eigenvectors as columns in the same order
// (column-major l x l), the contract of rocsolver's syevd / LAPACK dsyevd.
// No device memory beyond the caller's arrays is used.
#include "tsvd_pipeline.h"

namespace tsvd {

namespace {

constexpr int JTPB = 256;

__device__ inline double block_sum(double v, double *red)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < JTPB / 64; i++) s += red[i];
    return s;
}

__global__ void __launch_bounds__(JTPB) k_jacobi(double *__restrict__ Gall, double *__restrict__ lam_all, int l,
                                                 int max_sweeps)
{
    // row stride: odd, so a column of A spreads over the LDS banks
    __shared__ double A[JACOBI_LMAX * (JACOBI_LMAX + 1)];
    __shared__ int pp[JACOBI_LMAX / 2], pq[JACOBI_LMAX / 2];
    __shared__ double pc[JACOBI_LMAX / 2], ps[JACOBI_LMAX / 2];
    __shared__ double dv[JACOBI_LMAX];
    __shared__ double red[JTPB / 64];
    const int ld = l | 1;
    double *__restrict__ V = Gall + (int64_t)blockIdx.x * l * l;
    double *__restrict__ lam = lam_all + (int64_t)blockIdx.x * l;
    const int tid = threadIdx.x;
    const int ll = l * l;
    // A = (G + G^T) / 2 (the Grams are symmetric up to the order of their sums)
    double fro = 0.0;
    for (int e = tid; e < ll; e += JTPB) {
        const int r = e / l, c = e - r * l;
        const double x = 0.5 * (V[(int64_t)c * l + r] + V[(int64_t)r * l + c]);
        A[r * ld + c] = x;
        fro += x * x;
    }
    fro = block_sum(fro, red);
    for (int e = tid; e < ll; e += JTPB) {
        const int c = e / l, r = e - c * l;
        V[e] = r == c ? 1.0 : 0.0;
    }
    __syncthreads();
    const int np = l + (l & 1);
    const int half = np >> 1;
    const double tol2 = 1.6e-31 * fro;   // (4e-16)^2 * ||A||_F^2
    for (int sweep = 0; sweep < max_sweeps; sweep++) {
        double off = 0.0;
        for (int e = tid; e < ll; e += JTPB) {
            const int r = e / l, c = e - r * l;
            if (r != c) off += A[r * ld + c] * A[r * ld + c];
        }
        off = block_sum(off, red);
        if (off <= tol2) break;
        for (int step = 0; step < np - 1; step++) {
            if (tid < half) {
                int a, b;
                if (tid == 0) { a = step; b = np - 1; }
                else { a = (step + tid) % (np - 1); b = (step - tid + (np - 1)) % (np - 1); }
                const int p = a < b ? a : b, q = a < b ? b : a;
                double c = 1.0, s = 0.0;
                if (q < l) {
                    const double apq = A[p * ld + q];
                    if (apq != 0.0) {
                        const double tau = (A[q * ld + q] - A[p * ld + p]) / (2.0 * apq);
                        const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                        c = 1.0 / sqrt(1.0 + t * t);
                        s = t * c;
                    }
                }
                pp[tid] = p;
                pq[tid] = q < l ? q : -1;
                pc[tid] = c;
                ps[tid] = s;
            }
            __syncthreads();
            // rows p, q  <-  J^T A   (consecutive threads take consecutive columns)
            for (int e = tid; e < half * l; e += JTPB) {
                const int i = e / l, j = e - i * l;
                const int q = pq[i];
                if (q < 0 || ps[i] == 0.0) continue;
                const int p = pp[i];
                const double c = pc[i], s = ps[i];
                const double x = A[p * ld + j], y = A[q * ld + j];
                A[p * ld + j] = c * x - s * y;
                A[q * ld + j] = s * x + c * y;
            }
            __syncthreads();
            // columns p, q  <-  A J, and V <- V J (consecutive threads take consecutive pairs)
            for (int e = tid; e < half * l; e += JTPB) {
                const int r = e / half, i = e - r * half;
                const int q = pq[i];
                if (q < 0 || ps[i] == 0.0) continue;
                const int p = pp[i];
                const double c = pc[i], s = ps[i];
                const double x = A[r * ld + p], y = A[r * ld + q];
                A[r * ld + p] = c * x - s * y;
                A[r * ld + q] = s * x + c * y;
            }
            for (int e = tid; e < half * l; e += JTPB) {
                const int i = e / l, r = e - i * l;
                const int q = pq[i];
                if (q < 0 || ps[i] == 0.0) continue;
                const int p = pp[i];
                const double c = pc[i], s = ps[i];
                const double x = V[(int64_t)p * l + r], y = V[(int64_t)q * l + r];
                V[(int64_t)p * l + r] = c * x - s * y;
                V[(int64_t)q * l + r] = s * x + c * y;
            }
            __syncthreads();
        }
    }
    // ascending order: rank of each eigenvalue (ties by index), then permute V
    for (int i = tid; i < l; i += JTPB) dv[i] = A[i * ld + i];
    __syncthreads();
    double *Vs = A;   // A is no longer needed: stage V through LDS
    for (int e = tid; e < ll; e += JTPB) Vs[e] = V[e];
    __syncthreads();
    for (int i = tid; i < l; i += JTPB) {
        const double x = dv[i];
        int rank = 0;
        for (int j = 0; j < l; j++) rank += (dv[j] < x) || (dv[j] == x && j < i);
        lam[rank] = x;
        for (int r = 0; r < l; r++) V[(int64_t)rank * l + r] = Vs[i * l + r];
    }
}

}  // namespace

void jacobi_eig(hipStream_t st, double *G, int l, int count, double *lam)
{
    if (count <= 0) return;
    TSVD_REQUIRE(l >= 1 && l <= JACOBI_LMAX, "Jacobi eigensolver takes l <= %d (got %d)", JACOBI_LMAX, l);
    PROF("eig_jacobi", st);
    hipLaunchKernelGGL(k_jacobi, dim3(count), dim3(JTPB), 0, st, G, lam, l, 40);
    TSVD_HIP(hipGetLastError());
}

}  // namespace tsvd
