// hipBLAS-compatible entry points over libhcblas (no GPU code here).
//
// Replaces the reference's hipBLAS-over-hcBLAS shim
// (lib/src/hcc_detail/hipblas.cpp:29-548): each hipblasX forwards to the
// library's hcblasX with the operation code translated (HIPBLAS_OP_N/T/C =
// 111/112/113 -> HCBLAS_OP_N/T/C = 0/1/2; any other value is INVALID_VALUE
// where the reference threw, hipblas.cpp:87-99) and the status translated
// name for name (hipblas.cpp:107-126). A hipblasHandle_t is the library's
// handle. The handle-less Set/Get helpers use a per-thread handle of the
// current device (the reference created and leaked one per call,
// hipblas.cpp:141-205).
//
// Built as its own shared object (libhcblas_hipblas.so, linked to
// libhcblas.so) so that a process that also loads ROCm's hipBLAS never has
// these names interposed on it.
#include "../../include/hcblas.h"
#include "../../include/hipblas_hcblas.h"

#include <hip/hip_runtime_api.h>

#include <unordered_map>

namespace {

hipblasStatus_t to_hip(hcblasStatus_t s) {
  switch (s) {
    case HCBLAS_STATUS_SUCCESS:
      return HIPBLAS_STATUS_SUCCESS;
    case HCBLAS_STATUS_NOT_INITIALIZED:
      return HIPBLAS_STATUS_NOT_INITIALIZED;
    case HCBLAS_STATUS_ALLOC_FAILED:
      return HIPBLAS_STATUS_ALLOC_FAILED;
    case HCBLAS_STATUS_INVALID_VALUE:
      return HIPBLAS_STATUS_INVALID_VALUE;
    case HCBLAS_STATUS_MAPPING_ERROR:
      return HIPBLAS_STATUS_MAPPING_ERROR;
    case HCBLAS_STATUS_EXECUTION_FAILED:
      return HIPBLAS_STATUS_EXECUTION_FAILED;
    case HCBLAS_STATUS_INTERNAL_ERROR:
      return HIPBLAS_STATUS_INTERNAL_ERROR;
  }
  return HIPBLAS_STATUS_UNKNOWN;
}

// hipBLAS operation code -> library code; false for an unknown code.
bool to_hc(hipblasOperation_t op, hcblasOperation_t *out) {
  switch (op) {
    case HIPBLAS_OP_N:
      *out = HCBLAS_OP_N;
      return true;
    case HIPBLAS_OP_T:
      *out = HCBLAS_OP_T;
      return true;
    case HIPBLAS_OP_C:
      *out = HCBLAS_OP_C;
      return true;
  }
  return false;
}

hcblasHandle_t H(hipblasHandle_t h) { return static_cast<hcblasHandle_t>(h); }

// Per-thread handle of the current device for the handle-less helpers.
hcblasHandle_t helper_handle() {
  thread_local std::unordered_map<int, hcblasHandle_t> handles;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  auto it = handles.find(dev);
  if (it != handles.end()) return it->second;
  hcblasHandle_t h = nullptr;
  if (hcblasCreate(&h) != HCBLAS_STATUS_SUCCESS) return nullptr;
  handles.emplace(dev, h);
  return h;
}

}  // namespace

#define OP2(a, b)                                              \
  hcblasOperation_t oa, ob;                                    \
  if (!to_hc(a, &oa) || !to_hc(b, &ob)) return HIPBLAS_STATUS_INVALID_VALUE;

extern "C" {

hipblasStatus_t hipblasCreate(hipblasHandle_t *handle) {
  if (handle == nullptr) return HIPBLAS_STATUS_INVALID_VALUE;
  hcblasHandle_t h = nullptr;
  const hcblasStatus_t st = hcblasCreate(&h);
  *handle = h;
  return to_hip(st);
}
hipblasStatus_t hipblasDestroy(hipblasHandle_t handle) {
  hcblasHandle_t h = H(handle);
  return to_hip(hcblasDestroy(&h));
}
hipblasStatus_t hipblasSetStream(hipblasHandle_t handle, hipStream_t streamId) {
  return to_hip(hcblasSetStream(H(handle), streamId));
}
hipblasStatus_t hipblasGetStream(hipblasHandle_t handle,
                                 hipStream_t *streamId) {
  return to_hip(hcblasGetStream(H(handle), streamId));
}

hipblasStatus_t hipblasSetVector(int n, int elemSize, const void *x, int incx,
                                 void *y, int incy) {
  return to_hip(hcblasSetVector(helper_handle(), n, elemSize, x, incx, y, incy));
}
hipblasStatus_t hipblasGetVector(int n, int elemSize, const void *x, int incx,
                                 void *y, int incy) {
  return to_hip(hcblasGetVector(helper_handle(), n, elemSize, x, incx, y, incy));
}
hipblasStatus_t hipblasSetMatrix(int rows, int cols, int elemSize,
                                 const void *A, int lda, void *B, int ldb) {
  return to_hip(
      hcblasSetMatrix(helper_handle(), rows, cols, elemSize, A, lda, B, ldb));
}
hipblasStatus_t hipblasGetMatrix(int rows, int cols, int elemSize,
                                 const void *A, int lda, void *B, int ldb) {
  return to_hip(
      hcblasGetMatrix(helper_handle(), rows, cols, elemSize, A, lda, B, ldb));
}

// ---- level 1 --------------------------------------------------------------
hipblasStatus_t hipblasSasum(hipblasHandle_t handle, int n, const float *x,
                             int incx, float *result) {
  return to_hip(hcblasSasum(H(handle), n, x, incx, result));
}
hipblasStatus_t hipblasDasum(hipblasHandle_t handle, int n, const double *x,
                             int incx, double *result) {
  return to_hip(hcblasDasum(H(handle), n, x, incx, result));
}
hipblasStatus_t hipblasSasumBatched(hipblasHandle_t handle, int n, float *x,
                                    int incx, float *result, int batchCount) {
  return to_hip(hcblasSasumBatched(H(handle), n, x, incx, result, batchCount));
}
hipblasStatus_t hipblasDasumBatched(hipblasHandle_t handle, int n, double *x,
                                    int incx, double *result, int batchCount) {
  return to_hip(hcblasDasumBatched(H(handle), n, x, incx, result, batchCount));
}
hipblasStatus_t hipblasSaxpy(hipblasHandle_t handle, int n, const float *alpha,
                             const float *x, int incx, float *y, int incy) {
  return to_hip(hcblasSaxpy(H(handle), n, alpha, x, incx, y, incy));
}
hipblasStatus_t hipblasDaxpy(hipblasHandle_t handle, int n, const double *alpha,
                             const double *x, int incx, double *y, int incy) {
  return to_hip(hcblasDaxpy(H(handle), n, alpha, x, incx, y, incy));
}
hipblasStatus_t hipblasSaxpyBatched(hipblasHandle_t handle, int n,
                                    const float *alpha, const float *x,
                                    int incx, float *y, int incy,
                                    int batchCount) {
  return to_hip(
      hcblasSaxpyBatched(H(handle), n, alpha, x, incx, y, incy, batchCount));
}
hipblasStatus_t hipblasScopy(hipblasHandle_t handle, int n, const float *x,
                             int incx, float *y, int incy) {
  return to_hip(hcblasScopy(H(handle), n, x, incx, y, incy));
}
hipblasStatus_t hipblasDcopy(hipblasHandle_t handle, int n, const double *x,
                             int incx, double *y, int incy) {
  return to_hip(hcblasDcopy(H(handle), n, x, incx, y, incy));
}
hipblasStatus_t hipblasScopyBatched(hipblasHandle_t handle, int n,
                                    const float *x, int incx, float *y,
                                    int incy, int batchCount) {
  return to_hip(
      hcblasScopyBatched(H(handle), n, x, incx, y, incy, batchCount));
}
hipblasStatus_t hipblasDcopyBatched(hipblasHandle_t handle, int n,
                                    const double *x, int incx, double *y,
                                    int incy, int batchCount) {
  return to_hip(
      hcblasDcopyBatched(H(handle), n, x, incx, y, incy, batchCount));
}
hipblasStatus_t hipblasSdot(hipblasHandle_t handle, int n, const float *x,
                            int incx, const float *y, int incy,
                            float *result) {
  return to_hip(hcblasSdot(H(handle), n, x, incx, y, incy, result));
}
hipblasStatus_t hipblasDdot(hipblasHandle_t handle, int n, const double *x,
                            int incx, const double *y, int incy,
                            double *result) {
  return to_hip(hcblasDdot(H(handle), n, x, incx, y, incy, result));
}
hipblasStatus_t hipblasSdotBatched(hipblasHandle_t handle, int n,
                                   const float *x, int incx, const float *y,
                                   int incy, float *result, int batchCount) {
  return to_hip(
      hcblasSdotBatched(H(handle), n, x, incx, y, incy, result, batchCount));
}
hipblasStatus_t hipblasDdotBatched(hipblasHandle_t handle, int n,
                                   const double *x, int incx, const double *y,
                                   int incy, double *result, int batchCount) {
  return to_hip(
      hcblasDdotBatched(H(handle), n, x, incx, y, incy, result, batchCount));
}
hipblasStatus_t hipblasSscal(hipblasHandle_t handle, int n, const float *alpha,
                             float *x, int incx) {
  return to_hip(hcblasSscal(H(handle), n, alpha, x, incx));
}
hipblasStatus_t hipblasDscal(hipblasHandle_t handle, int n, const double *alpha,
                             double *x, int incx) {
  return to_hip(hcblasDscal(H(handle), n, alpha, x, incx));
}
hipblasStatus_t hipblasCscal(hipblasHandle_t handle, int n,
                             const hipblasShimComplex *alpha,
                             hipblasShimComplex *x, int incx) {
  return to_hip(hcblasCscal(H(handle), n, (const hcComplex *)alpha,
                            (hcComplex *)x, incx));
}
hipblasStatus_t hipblasZscal(hipblasHandle_t handle, int n,
                             const hipblasShimDoubleComplex *alpha,
                             hipblasShimDoubleComplex *x, int incx) {
  return to_hip(hcblasZscal(H(handle), n, (const hcDoubleComplex *)alpha,
                            (hcDoubleComplex *)x, incx));
}
hipblasStatus_t hipblasCsscal(hipblasHandle_t handle, int n, const float *alpha,
                              hipblasShimComplex *x, int incx) {
  return to_hip(hcblasCsscal(H(handle), n, alpha, (hcComplex *)x, incx));
}
hipblasStatus_t hipblasZdscal(hipblasHandle_t handle, int n,
                              const double *alpha,
                              hipblasShimDoubleComplex *x, int incx) {
  return to_hip(
      hcblasZdscal(H(handle), n, alpha, (hcDoubleComplex *)x, incx));
}
hipblasStatus_t hipblasSscalBatched(hipblasHandle_t handle, int n,
                                    const float *alpha, float *x, int incx,
                                    int batchCount) {
  return to_hip(hcblasSscalBatched(H(handle), n, alpha, x, incx, batchCount));
}
hipblasStatus_t hipblasDscalBatched(hipblasHandle_t handle, int n,
                                    const double *alpha, double *x, int incx,
                                    int batchCount) {
  return to_hip(hcblasDscalBatched(H(handle), n, alpha, x, incx, batchCount));
}
hipblasStatus_t hipblasCscalBatched(hipblasHandle_t handle, int n,
                                    const hipblasShimComplex *alpha,
                                    hipblasShimComplex *x, int incx,
                                    int batchCount) {
  return to_hip(hcblasCscalBatched(H(handle), n, (const hcComplex *)alpha,
                                   (hcComplex *)x, incx, batchCount));
}
hipblasStatus_t hipblasZscalBatched(hipblasHandle_t handle, int n,
                                    const hipblasShimDoubleComplex *alpha,
                                    hipblasShimDoubleComplex *x, int incx,
                                    int batchCount) {
  return to_hip(hcblasZscalBatched(H(handle), n,
                                   (const hcDoubleComplex *)alpha,
                                   (hcDoubleComplex *)x, incx, batchCount));
}
hipblasStatus_t hipblasCsscalBatched(hipblasHandle_t handle, int n,
                                     const float *alpha, hipblasShimComplex *x,
                                     int incx, int batchCount) {
  return to_hip(hcblasCsscalBatched(H(handle), n, alpha, (hcComplex *)x, incx,
                                    batchCount));
}
hipblasStatus_t hipblasZdscalBatched(hipblasHandle_t handle, int n,
                                     const double *alpha,
                                     hipblasShimDoubleComplex *x, int incx,
                                     int batchCount) {
  return to_hip(hcblasZdscalBatched(H(handle), n, alpha,
                                    (hcDoubleComplex *)x, incx, batchCount));
}

// ---- level 2 --------------------------------------------------------------
hipblasStatus_t hipblasSgemv(hipblasHandle_t handle, hipblasOperation_t trans,
                             int m, int n, const float *alpha, const float *A,
                             int lda, const float *x, int incx,
                             const float *beta, float *y, int incy) {
  hcblasOperation_t t;
  if (!to_hc(trans, &t)) return HIPBLAS_STATUS_INVALID_VALUE;
  return to_hip(
      hcblasSgemv(H(handle), t, m, n, alpha, A, lda, x, incx, beta, y, incy));
}
hipblasStatus_t hipblasDgemv(hipblasHandle_t handle, hipblasOperation_t trans,
                             int m, int n, const double *alpha, const double *A,
                             int lda, const double *x, int incx,
                             const double *beta, double *y, int incy) {
  hcblasOperation_t t;
  if (!to_hc(trans, &t)) return HIPBLAS_STATUS_INVALID_VALUE;
  return to_hip(
      hcblasDgemv(H(handle), t, m, n, alpha, A, lda, x, incx, beta, y, incy));
}
hipblasStatus_t hipblasSgemvBatched(hipblasHandle_t handle,
                                    hipblasOperation_t trans, int m, int n,
                                    const float *alpha, float *A, int lda,
                                    float *x, int incx, const float *beta,
                                    float *y, int incy, int batchCount) {
  hcblasOperation_t t;
  if (!to_hc(trans, &t)) return HIPBLAS_STATUS_INVALID_VALUE;
  return to_hip(hcblasSgemvBatched(H(handle), t, m, n, alpha, A, lda, x, incx,
                                   beta, y, incy, batchCount));
}
hipblasStatus_t hipblasSger(hipblasHandle_t handle, int m, int n,
                            const float *alpha, const float *x, int incx,
                            const float *y, int incy, float *A, int lda) {
  return to_hip(hcblasSger(H(handle), m, n, alpha, x, incx, y, incy, A, lda));
}
hipblasStatus_t hipblasDger(hipblasHandle_t handle, int m, int n,
                            const double *alpha, const double *x, int incx,
                            const double *y, int incy, double *A, int lda) {
  return to_hip(hcblasDger(H(handle), m, n, alpha, x, incx, y, incy, A, lda));
}
hipblasStatus_t hipblasSgerBatched(hipblasHandle_t handle, int m, int n,
                                   const float *alpha, const float *x, int incx,
                                   const float *y, int incy, float *A, int lda,
                                   int batchCount) {
  return to_hip(hcblasSgerBatched(H(handle), m, n, alpha, x, incx, y, incy, A,
                                  lda, batchCount));
}
hipblasStatus_t hipblasDgerBatched(hipblasHandle_t handle, int m, int n,
                                   const double *alpha, const double *x,
                                   int incx, const double *y, int incy,
                                   double *A, int lda, int batchCount) {
  return to_hip(hcblasDgerBatched(H(handle), m, n, alpha, x, incx, y, incy, A,
                                  lda, batchCount));
}

// ---- level 3 --------------------------------------------------------------
hipblasStatus_t hipblasSgemm(hipblasHandle_t handle, hipblasOperation_t transa,
                             hipblasOperation_t transb, int m, int n, int k,
                             const float *alpha, const float *A, int lda,
                             const float *B, int ldb, const float *beta,
                             float *C, int ldc) {
  OP2(transa, transb)
  return to_hip(hcblasSgemm(H(handle), oa, ob, m, n, k, alpha, A, lda, B, ldb,
                            beta, C, ldc));
}
hipblasStatus_t hipblasDgemm(hipblasHandle_t handle, hipblasOperation_t transa,
                             hipblasOperation_t transb, int m, int n, int k,
                             const double *alpha, const double *A, int lda,
                             const double *B, int ldb, const double *beta,
                             double *C, int ldc) {
  OP2(transa, transb)
  return to_hip(hcblasDgemm(H(handle), oa, ob, m, n, k, alpha, A, lda, B, ldb,
                            beta, C, ldc));
}
hipblasStatus_t hipblasCgemm(hipblasHandle_t handle, hipblasOperation_t transa,
                             hipblasOperation_t transb, int m, int n, int k,
                             const hipblasShimComplex *alpha,
                             const hipblasShimComplex *A, int lda,
                             const hipblasShimComplex *B, int ldb,
                             const hipblasShimComplex *beta,
                             hipblasShimComplex *C, int ldc) {
  OP2(transa, transb)
  return to_hip(hcblasCgemm(H(handle), oa, ob, m, n, k,
                            (const hcComplex *)alpha, (const hcComplex *)A,
                            lda, (const hcComplex *)B, ldb,
                            (const hcComplex *)beta, (hcComplex *)C, ldc));
}
hipblasStatus_t hipblasZgemm(hipblasHandle_t handle, hipblasOperation_t transa,
                             hipblasOperation_t transb, int m, int n, int k,
                             const hipblasShimDoubleComplex *alpha,
                             const hipblasShimDoubleComplex *A, int lda,
                             const hipblasShimDoubleComplex *B, int ldb,
                             const hipblasShimDoubleComplex *beta,
                             hipblasShimDoubleComplex *C, int ldc) {
  OP2(transa, transb)
  return to_hip(hcblasZgemm(
      H(handle), oa, ob, m, n, k, (const hcDoubleComplex *)alpha,
      (const hcDoubleComplex *)A, lda, (const hcDoubleComplex *)B, ldb,
      (const hcDoubleComplex *)beta, (hcDoubleComplex *)C, ldc));
}
hipblasStatus_t hipblasHgemm(hipblasHandle_t handle, hipblasOperation_t transa,
                             hipblasOperation_t transb, int m, int n, int k,
                             const hipblasShimHalf *alpha,
                             const hipblasShimHalf *A, int lda,
                             const hipblasShimHalf *B, int ldb,
                             const hipblasShimHalf *beta, hipblasShimHalf *C,
                             int ldc) {
  OP2(transa, transb)
  return to_hip(hcblasHgemm(H(handle), oa, ob, m, n, k, alpha, A, lda, B, ldb,
                            beta, C, ldc));
}
hipblasStatus_t hipblasSgemmBatched(hipblasHandle_t handle,
                                    hipblasOperation_t transa,
                                    hipblasOperation_t transb, int m, int n,
                                    int k, const float *alpha, const float *A[],
                                    int lda, const float *B[], int ldb,
                                    const float *beta, float *C[], int ldc,
                                    int batchCount) {
  OP2(transa, transb)
  return to_hip(hcblasSgemmBatched(H(handle), oa, ob, m, n, k, alpha, A, lda,
                                   B, ldb, beta, C, ldc, batchCount));
}
hipblasStatus_t hipblasDgemmBatched(hipblasHandle_t handle,
                                    hipblasOperation_t transa,
                                    hipblasOperation_t transb, int m, int n,
                                    int k, const double *alpha,
                                    const double *A[], int lda,
                                    const double *B[], int ldb,
                                    const double *beta, double *C[], int ldc,
                                    int batchCount) {
  OP2(transa, transb)
  return to_hip(hcblasDgemmBatched(H(handle), oa, ob, m, n, k, alpha, A, lda,
                                   B, ldb, beta, C, ldc, batchCount));
}
hipblasStatus_t hipblasCgemmBatched(hipblasHandle_t handle,
                                    hipblasOperation_t transa,
                                    hipblasOperation_t transb, int m, int n,
                                    int k, const hipblasShimComplex *alpha,
                                    const hipblasShimComplex *A[], int lda,
                                    const hipblasShimComplex *B[], int ldb,
                                    const hipblasShimComplex *beta,
                                    hipblasShimComplex *C[], int ldc,
                                    int batchCount) {
  OP2(transa, transb)
  return to_hip(hcblasCgemmBatched(
      H(handle), oa, ob, m, n, k, (const hcComplex *)alpha,
      (const hcComplex *const *)A, lda, (const hcComplex *const *)B, ldb,
      (const hcComplex *)beta, (hcComplex *const *)C, ldc, batchCount));
}
hipblasStatus_t hipblasZgemmBatched(hipblasHandle_t handle,
                                    hipblasOperation_t transa,
                                    hipblasOperation_t transb, int m, int n,
                                    int k,
                                    const hipblasShimDoubleComplex *alpha,
                                    const hipblasShimDoubleComplex *A[],
                                    int lda,
                                    const hipblasShimDoubleComplex *B[],
                                    int ldb,
                                    const hipblasShimDoubleComplex *beta,
                                    hipblasShimDoubleComplex *C[], int ldc,
                                    int batchCount) {
  OP2(transa, transb)
  return to_hip(hcblasZgemmBatched(
      H(handle), oa, ob, m, n, k, (const hcDoubleComplex *)alpha,
      (const hcDoubleComplex *const *)A, lda,
      (const hcDoubleComplex *const *)B, ldb, (const hcDoubleComplex *)beta,
      (hcDoubleComplex *const *)C, ldc, batchCount));
}

}  // extern "C"
