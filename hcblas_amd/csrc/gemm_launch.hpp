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

}  // namespace hcb
