// Host launcher for the generic MFMA GEMM kernel: picks a macro-tile shape
// and the template instance matching the operand orientations.
#pragma once

#include "gemm_kernel.hpp"

namespace hcb {

// Launches C = alpha*A'*B' + beta*C for `batch` problems described by p
// (column-major storage; a_k / b_k give the operand orientations; vload
// allows 16-byte global loads for both operands).
template <Elem E, typename OutT>
hcblasStatus_t launch_gemm(hipStream_t stream, int num_cus, GemmParams p,
                           bool a_k, bool b_k, bool vload, int64_t batch);

// Split-K form for one problem (no batch, no pointer arrays, vector loads):
// S partial tiles of K step ks into W (S * M * N accumulators), then the
// reduction into C.
template <Elem E, typename OutT>
hcblasStatus_t launch_gemm_splitk(hipStream_t stream, int num_cus,
                                  GemmParams p, bool a_k, bool b_k, int S,
                                  int ks, typename ElemT<E>::Acc *W);

// Split count for a single problem, 1 when splitting K does not pay. It
// pays when the 64x64 tiles leave at least half the CUs idle and K is long
// enough that every split keeps a K loop of 1024 or more after the extra
// launch and the S * M * N partials: K >= 8192, aiming at four workgroups
// per CU, at most 8 splits (OverFeat's 3072 x 128 x 36864 runs 8-way;
// 1024^3, with one tile per CU, measured slower split 4-way: 0.045 ms
// against 0.038).
inline int splitk_count(int M, int N, int K, int num_cus, int bk, int &ks) {
  const int64_t tiles = ((int64_t)M + 63) / 64 * (((int64_t)N + 63) / 64);
  ks = K;
  if (K < 8192 || 2 * tiles > num_cus) return 1;
  int S = (int)((4 * (int64_t)num_cus) / tiles);
  S = S > 8 ? 8 : S;
  S = S > K / 1024 ? K / 1024 : S;
  if (S < 2) return 1;
  ks = (K + S - 1) / S;
  ks = (ks + bk - 1) / bk * bk;
  return (K + ks - 1) / ks;
}
}  // namespace hcb
