// Shared GEMM host-side declarations: the canonical column-major problem and
// the drivers implemented in gemm.hip / gemm_split.hip.
#pragma once

#include "gemm_kernel.hpp"

namespace hcb {

// Canonical column-major problem C[M x N] = alpha * A'[M x K] * B'[K x N]
// + beta * C, after the row-major operand swap.
struct Canon {
  int M, N, K;
  bool a_k, b_k;        // A' k-contiguous, B' k-contiguous
  bool conj_a, conj_b;  // complex conjugation of A', B'
  const void *A;
  int64_t lda, aOff, strideA;
  const void *const *Aptrs;
  const void *B;
  int64_t ldb, bOff, strideB;
  const void *const *Bptrs;
};

template <Elem E, typename OutT>
hcblasStatus_t gemm_real(hcblasHandle_t h, const Canon &c, double alpha,
                         double beta, void *C, int64_t ldc, int64_t cOff,
                         int64_t strideC, void *const *Cptrs, int64_t batch);

hcblasStatus_t gemm_split_bf16(hcblasHandle_t h, const Canon &c, float alpha,
                               float beta, float *C, int64_t ldc, int64_t cOff,
                               int64_t strideC, void *const *Cptrs,
                               int64_t batch);
hcblasStatus_t gemm_ozaki_int8(hcblasHandle_t h, const Canon &c, double alpha,
                               double beta, double *C, int64_t ldc,
                               int64_t cOff, int64_t strideC,
                               void *const *Cptrs, int64_t batch);
hcblasStatus_t gemm_ozaki_crt(hcblasHandle_t h, const Canon &c, double alpha,
                              double beta, double *C, int64_t ldc,
                              int64_t cOff, int64_t strideC,
                              void *const *Cptrs, int64_t batch);
// C = alpha * A' * B' (beta = 0) with A' FP32 and B' BF16 (a_f32) or A'
// BF16 and B' FP32, written as FP32 or (c_bf16) BF16; strided batch
// (gemm_mixed.hip).
hcblasStatus_t gemm_mixed_bf16(hcblasHandle_t h, const Canon &c, bool a_f32,
                               float alpha, void *C, bool c_bf16, int64_t ldc,
                               int64_t strideC, int64_t batch);
hcblasStatus_t gemm_ozaki_crt_f32(hcblasHandle_t h, const Canon &c,
                                  float alpha, float beta, float *C,
                                  int64_t ldc, int64_t cOff, int64_t strideC,
                                  void *const *Cptrs, int64_t batch);

}  // namespace hcb
