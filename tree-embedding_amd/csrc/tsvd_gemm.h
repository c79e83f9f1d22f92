// tsvd_gemm.h -- grouped dense products on MFMA (FP32 or FP64 accumulate).
#pragma once
#include "tsvd_internal.h"

#include <string>

namespace tsvd {

// BF3 is an accumulate type only: bf16x3 products with FP32 accumulation
enum class DT : int { F32 = 0, F64 = 1, BF3 = 2 };

// One product  C (M x N) = alpha * op(A) * B + beta * C, all column-major:
//   transA = 0: A is M x K,  A(i,k) = A[k*lda + i]
//   transA = 1: A is K x M,  op(A)(i,k) = A[i*lda + k]   (C = A^T B)
//   B is K x N,  B(k,j) = B[j*ldb + k];  C(i,j) = C[j*ldc + i].
// kchunk > 0 splits K into chunks whose partial products are added to C
// atomically (C must hold beta*C already; used for FP64 Grams).
struct GemmProb {
    const void *A;
    const void *B;
    void *C;
    int64_t lda, ldb, ldc;
    int32_t M, N, K, transA;
    int32_t transB;   // 1: B is stored N x K, B(k,j) = B[k*ldb + j]
    double alpha, beta;
    int32_t kchunk;
    int32_t tiles_m, tiles_n, tiles_k;
    int64_t tile_begin;
};

// Element types of A, B, C and the accumulate type (F32 -> v_mfma_f32_16x16x4f32,
// F64 -> v_mfma_f64_16x16x4f64).  `wide` selects the 144-column tile (else 80).
struct GemmTypes {
    DT a, b, c, acc;
};

// Tile-width autotuning (process-wide): while enabled, the first call of a
// (label, types, N, K class) whose problems are idempotent (no K split,
// beta = 0, C distinct from A and B) runs with both tile widths, timed with
// events, and the faster width is kept for every later call of that key.
void gemm_tuning_enable(bool on);
std::string gemm_tuning_json();

class GemmRunner {
public:
    void run(hipStream_t st, std::vector<GemmProb> &probs, GemmTypes t, bool wide, const char *label = "gemm");
    // The same, with the problem descriptors staged in a caller-provided
    // device array of `cap` slots (allocates nothing).
    void run_into(hipStream_t st, GemmProb *probs, int nprobs, GemmTypes t, bool wide, GemmProb *dev, int cap,
                  const char *label = "gemm");
    int64_t launches = 0;

private:
    void launch_once(hipStream_t st, GemmProb *probs, int nprobs, GemmTypes t, bool wide, GemmProb *dev, int cap,
                     const char *label);
    DBuf<GemmProb> dprobs_;
};

}  // namespace tsvd
