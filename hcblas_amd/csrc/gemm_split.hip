// Split-precision GEMM drivers (BF16 split SGEMM, Ozaki INT8 DGEMM).
#include "gemm_launch.hpp"

namespace hcb {
struct Canon;
hcblasStatus_t gemm_split_bf16(hcblasHandle_t, const Canon &, float, float,
                               float *, int64_t, int64_t, int64_t,
                               void *const *, int64_t) {
  return HCBLAS_STATUS_INTERNAL_ERROR;
}
hcblasStatus_t gemm_ozaki_int8(hcblasHandle_t, const Canon &, double, double,
                               double *, int64_t, int64_t, int64_t,
                               void *const *, int64_t) {
  return HCBLAS_STATUS_INTERNAL_ERROR;
}
}  // namespace hcb
