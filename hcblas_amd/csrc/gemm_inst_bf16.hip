// Template instances of the MFMA GEMM for one element type.
#include "gemm_launch_impl.hpp"

namespace hcb {
template hcblasStatus_t launch_gemm<Elem::BF16, float>(hipStream_t, int,
                                                     GemmParams, bool, bool,
                                                     bool, int64_t);
template hcblasStatus_t launch_gemm_splitk<Elem::BF16, float>(
    hipStream_t, int, GemmParams, bool, bool, int, int,
    typename ElemT<Elem::BF16>::Acc *);
}  // namespace hcb
