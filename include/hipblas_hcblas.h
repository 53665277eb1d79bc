/*
 * hipblas_hcblas.h — hipBLAS-compatible C entry points over libhcblas
 * (library: hcblas_amd/lib/libhcblas_hipblas.so).
 *
 * Replaces the reference's hipBLAS shim over hcBLAS:
 *   lib/include/hcc_detail/hip_blas.h:47-259 (declarations)
 *   lib/src/hcc_detail/hipblas.cpp:29-548     (implementation)
 * with the same routine names and argument lists. The status and operation
 * enums take the values of ROCm's hipBLAS (hipblas-common.h: statuses 0..11,
 * HIPBLAS_OP_N/T/C = 111/112/113), so a caller compiled against hipBLAS's
 * headers links against this library unchanged for the routines below.
 * Status mapping from the library's codes follows hipblas.cpp:107-126 (equal
 * names map to each other).
 *
 * hipblasSetVector / GetVector / SetMatrix / GetMatrix take no handle (as in
 * the reference, hipblas.cpp:141-205); they use a per-thread context of the
 * current device, created on first use.
 */
#ifndef HCBLAS_MI355X_HIPBLAS_HCBLAS_H_
#define HCBLAS_MI355X_HIPBLAS_HCBLAS_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t *hipStream_t;

#ifndef HIPBLAS_COMMON_H
typedef enum {
  HIPBLAS_STATUS_SUCCESS = 0,
  HIPBLAS_STATUS_NOT_INITIALIZED = 1,
  HIPBLAS_STATUS_ALLOC_FAILED = 2,
  HIPBLAS_STATUS_INVALID_VALUE = 3,
  HIPBLAS_STATUS_MAPPING_ERROR = 4,
  HIPBLAS_STATUS_EXECUTION_FAILED = 5,
  HIPBLAS_STATUS_INTERNAL_ERROR = 6,
  HIPBLAS_STATUS_NOT_SUPPORTED = 7,
  HIPBLAS_STATUS_ARCH_MISMATCH = 8,
  HIPBLAS_STATUS_HANDLE_IS_NULLPTR = 9,
  HIPBLAS_STATUS_INVALID_ENUM = 10,
  HIPBLAS_STATUS_UNKNOWN = 11
} hipblasStatus_t;
#endif
#ifndef HIPBLAS_OPERATION_DECLARED
#define HIPBLAS_OPERATION_DECLARED
typedef enum {
  HIPBLAS_OP_N = 111,
  HIPBLAS_OP_T = 112,
  HIPBLAS_OP_C = 113
} hipblasOperation_t;
#endif

typedef void *hipblasHandle_t;
/* Interleaved complex values (layout of hip/hip_complex.h). */
typedef struct { float x, y; } hipblasShimComplex;
typedef struct { double x, y; } hipblasShimDoubleComplex;
typedef uint16_t hipblasShimHalf; /* IEEE binary16 bits */

hipblasStatus_t hipblasCreate(hipblasHandle_t *handle);
hipblasStatus_t hipblasDestroy(hipblasHandle_t handle);
hipblasStatus_t hipblasSetStream(hipblasHandle_t handle, hipStream_t streamId);
hipblasStatus_t hipblasGetStream(hipblasHandle_t handle, hipStream_t *streamId);

hipblasStatus_t hipblasSetVector(int n, int elemSize, const void *x, int incx,
                                 void *y, int incy);
hipblasStatus_t hipblasGetVector(int n, int elemSize, const void *x, int incx,
                                 void *y, int incy);
hipblasStatus_t hipblasSetMatrix(int rows, int cols, int elemSize,
                                 const void *A, int lda, void *B, int ldb);
hipblasStatus_t hipblasGetMatrix(int rows, int cols, int elemSize,
                                 const void *A, int lda, void *B, int ldb);

hipblasStatus_t hipblasSasum(hipblasHandle_t handle, int n, const float *x,
                             int incx, float *result);
hipblasStatus_t hipblasDasum(hipblasHandle_t handle, int n, const double *x,
                             int incx, double *result);
hipblasStatus_t hipblasSasumBatched(hipblasHandle_t handle, int n, float *x,
                                    int incx, float *result, int batchCount);
hipblasStatus_t hipblasDasumBatched(hipblasHandle_t handle, int n, double *x,
                                    int incx, double *result, int batchCount);
hipblasStatus_t hipblasSaxpy(hipblasHandle_t handle, int n, const float *alpha,
                             const float *x, int incx, float *y, int incy);
hipblasStatus_t hipblasDaxpy(hipblasHandle_t handle, int n, const double *alpha,
                             const double *x, int incx, double *y, int incy);
hipblasStatus_t hipblasSaxpyBatched(hipblasHandle_t handle, int n,
                                    const float *alpha, const float *x,
                                    int incx, float *y, int incy,
                                    int batchCount);
hipblasStatus_t hipblasScopy(hipblasHandle_t handle, int n, const float *x,
                             int incx, float *y, int incy);
hipblasStatus_t hipblasDcopy(hipblasHandle_t handle, int n, const double *x,
                             int incx, double *y, int incy);
hipblasStatus_t hipblasScopyBatched(hipblasHandle_t handle, int n,
                                    const float *x, int incx, float *y,
                                    int incy, int batchCount);
hipblasStatus_t hipblasDcopyBatched(hipblasHandle_t handle, int n,
                                    const double *x, int incx, double *y,
                                    int incy, int batchCount);
hipblasStatus_t hipblasSdot(hipblasHandle_t handle, int n, const float *x,
                            int incx, const float *y, int incy, float *result);
hipblasStatus_t hipblasDdot(hipblasHandle_t handle, int n, const double *x,
                            int incx, const double *y, int incy,
                            double *result);
hipblasStatus_t hipblasSdotBatched(hipblasHandle_t handle, int n,
                                   const float *x, int incx, const float *y,
                                   int incy, float *result, int batchCount);
hipblasStatus_t hipblasDdotBatched(hipblasHandle_t handle, int n,
                                   const double *x, int incx, const double *y,
                                   int incy, double *result, int batchCount);
hipblasStatus_t hipblasSscal(hipblasHandle_t handle, int n, const float *alpha,
                             float *x, int incx);
hipblasStatus_t hipblasDscal(hipblasHandle_t handle, int n, const double *alpha,
                             double *x, int incx);
hipblasStatus_t hipblasCscal(hipblasHandle_t handle, int n,
                             const hipblasShimComplex *alpha,
                             hipblasShimComplex *x, int incx);
hipblasStatus_t hipblasZscal(hipblasHandle_t handle, int n,
                             const hipblasShimDoubleComplex *alpha,
                             hipblasShimDoubleComplex *x, int incx);
hipblasStatus_t hipblasCsscal(hipblasHandle_t handle, int n, const float *alpha,
                              hipblasShimComplex *x, int incx);
hipblasStatus_t hipblasZdscal(hipblasHandle_t handle, int n,
                              const double *alpha,
                              hipblasShimDoubleComplex *x, int incx);
hipblasStatus_t hipblasSscalBatched(hipblasHandle_t handle, int n,
                                    const float *alpha, float *x, int incx,
                                    int batchCount);
hipblasStatus_t hipblasDscalBatched(hipblasHandle_t handle, int n,
                                    const double *alpha, double *x, int incx,
                                    int batchCount);
hipblasStatus_t hipblasCscalBatched(hipblasHandle_t handle, int n,
                                    const hipblasShimComplex *alpha,
                                    hipblasShimComplex *x, int incx,
                                    int batchCount);
hipblasStatus_t hipblasZscalBatched(hipblasHandle_t handle, int n,
                                    const hipblasShimDoubleComplex *alpha,
                                    hipblasShimDoubleComplex *x, int incx,
                                    int batchCount);
hipblasStatus_t hipblasCsscalBatched(hipblasHandle_t handle, int n,
                                     const float *alpha, hipblasShimComplex *x,
                                     int incx, int batchCount);
hipblasStatus_t hipblasZdscalBatched(hipblasHandle_t handle, int n,
                                     const double *alpha,
                                     hipblasShimDoubleComplex *x, int incx,
                                     int batchCount);

hipblasStatus_t hipblasSgemv(hipblasHandle_t handle, hipblasOperation_t trans,
                             int m, int n, const float *alpha, const float *A,
                             int lda, const float *x, int incx,
                             const float *beta, float *y, int incy);
hipblasStatus_t hipblasDgemv(hipblasHandle_t handle, hipblasOperation_t trans,
                             int m, int n, const double *alpha, const double *A,
                             int lda, const double *x, int incx,
                             const double *beta, double *y, int incy);
hipblasStatus_t hipblasSgemvBatched(hipblasHandle_t handle,
                                    hipblasOperation_t trans, int m, int n,
                                    const float *alpha, float *A, int lda,
                                    float *x, int incx, const float *beta,
                                    float *y, int incy, int batchCount);
hipblasStatus_t hipblasSger(hipblasHandle_t handle, int m, int n,
                            const float *alpha, const float *x, int incx,
                            const float *y, int incy, float *A, int lda);
hipblasStatus_t hipblasDger(hipblasHandle_t handle, int m, int n,
                            const double *alpha, const double *x, int incx,
                            const double *y, int incy, double *A, int lda);
hipblasStatus_t hipblasSgerBatched(hipblasHandle_t handle, int m, int n,
                                   const float *alpha, const float *x, int incx,
                                   const float *y, int incy, float *A, int lda,
                                   int batchCount);
hipblasStatus_t hipblasDgerBatched(hipblasHandle_t handle, int m, int n,
                                   const double *alpha, const double *x,
                                   int incx, const double *y, int incy,
                                   double *A, int lda, int batchCount);

hipblasStatus_t hipblasSgemm(hipblasHandle_t handle, hipblasOperation_t transa,
                             hipblasOperation_t transb, int m, int n, int k,
                             const float *alpha, const float *A, int lda,
                             const float *B, int ldb, const float *beta,
                             float *C, int ldc);
hipblasStatus_t hipblasDgemm(hipblasHandle_t handle, hipblasOperation_t transa,
                             hipblasOperation_t transb, int m, int n, int k,
                             const double *alpha, const double *A, int lda,
                             const double *B, int ldb, const double *beta,
                             double *C, int ldc);
hipblasStatus_t hipblasCgemm(hipblasHandle_t handle, hipblasOperation_t transa,
                             hipblasOperation_t transb, int m, int n, int k,
                             const hipblasShimComplex *alpha,
                             const hipblasShimComplex *A, int lda,
                             const hipblasShimComplex *B, int ldb,
                             const hipblasShimComplex *beta,
                             hipblasShimComplex *C, int ldc);
hipblasStatus_t hipblasZgemm(hipblasHandle_t handle, hipblasOperation_t transa,
                             hipblasOperation_t transb, int m, int n, int k,
                             const hipblasShimDoubleComplex *alpha,
                             const hipblasShimDoubleComplex *A, int lda,
                             const hipblasShimDoubleComplex *B, int ldb,
                             const hipblasShimDoubleComplex *beta,
                             hipblasShimDoubleComplex *C, int ldc);
hipblasStatus_t hipblasHgemm(hipblasHandle_t handle, hipblasOperation_t transa,
                             hipblasOperation_t transb, int m, int n, int k,
                             const hipblasShimHalf *alpha,
                             const hipblasShimHalf *A, int lda,
                             const hipblasShimHalf *B, int ldb,
                             const hipblasShimHalf *beta, hipblasShimHalf *C,
                             int ldc);
hipblasStatus_t hipblasSgemmBatched(hipblasHandle_t handle,
                                    hipblasOperation_t transa,
                                    hipblasOperation_t transb, int m, int n,
                                    int k, const float *alpha, const float *A[],
                                    int lda, const float *B[], int ldb,
                                    const float *beta, float *C[], int ldc,
                                    int batchCount);
hipblasStatus_t hipblasDgemmBatched(hipblasHandle_t handle,
                                    hipblasOperation_t transa,
                                    hipblasOperation_t transb, int m, int n,
                                    int k, const double *alpha,
                                    const double *A[], int lda,
                                    const double *B[], int ldb,
                                    const double *beta, double *C[], int ldc,
                                    int batchCount);
hipblasStatus_t hipblasCgemmBatched(hipblasHandle_t handle,
                                    hipblasOperation_t transa,
                                    hipblasOperation_t transb, int m, int n,
                                    int k, const hipblasShimComplex *alpha,
                                    const hipblasShimComplex *A[], int lda,
                                    const hipblasShimComplex *B[], int ldb,
                                    const hipblasShimComplex *beta,
                                    hipblasShimComplex *C[], int ldc,
                                    int batchCount);
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
                                    int batchCount);

#ifdef __cplusplus
}
#endif

#endif /* HCBLAS_MI355X_HIPBLAS_HCBLAS_H_ */
