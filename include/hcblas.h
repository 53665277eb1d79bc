/*
 * hcblas.h — C ABI of the MI355X-native hcBLAS drop-in (gfx950, one GPU).
 *
 * Every entry point here has plain C linkage, plain pointers and sizes, and no
 * C++ or torch types, so it can be bound from C, C++, ctypes or any FFI.
 * Each declaration cites the reference interface it replaces
 * (rocmarchive/hcBLAS, paths relative to the reference repository root):
 *   - lib/include/hcblas.h     : public C++ API (hcblasSgemm, ...)
 *   - lib/include/hcblaslib.h  : internal Hcblaslibrary overloads that carry the
 *                                storage order, per-matrix offsets and batch
 *                                offsets (hcblas_sgemm(...order..., aOffset...))
 *
 * Memory contract (same as the reference, lib/src/hcblas.cpp):
 *   - matrices / vectors are caller-owned DEVICE pointers;
 *   - alpha / beta are HOST pointers read at call time (hcblas.cpp:451,1321);
 *   - dot / asum write a HOST scalar before returning, so those calls block
 *     on the handle's stream (sdot_array.cpp:136-158);
 *   - every other routine is enqueued on the handle's stream and returns
 *     without synchronising.
 *
 * Differences from the reference, all in the direction of CBLAS semantics
 * (the reference's own test oracle):
 *   - empty problems (m, n or k == 0, n <= 0) return SUCCESS without work
 *     (reference: EXECUTION_FAILED, sgemm_array_wrapper.cpp:307-309); k == 0 or
 *     alpha == 0 computes C = beta*C, and beta == 0 overwrites C without
 *     reading it (NaN in C does not propagate);
 *   - NaN / Inf inputs propagate (reference zeroes them, sgemm_array_wrapper.cpp:171);
 *   - incx / incy / lda / ldb / ldc are honoured, negative increments follow BLAS
 *     (reference ignores them, sdot_array.cpp:56-58, sger_array.cpp:145);
 *   - HCBLAS_OP_C conjugates complex operands (reference transposes only,
 *     hcblas.cpp:1318);
 *   - HGEMM accumulates in fp32 (reference accumulates in half,
 *     hgemm_NN_kernels.cpp:1372);
 *   - row-major batched GEMM works (reference silently skips it,
 *     sgemm_array_wrapper.cpp:113-151);
 *   - C/Z scal compute the complex product correctly (cscal_array.cpp:49).
 *   Batched dot / asum keep the reference's contract of ONE scalar summed over
 *   the whole batch (sdot_array.cpp:267-277, test-blas.cpp:116-120).
 */
#ifndef HCBLAS_MI355X_HCBLAS_H_
#define HCBLAS_MI355X_HCBLAS_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* HIP stream handle; identical to the typedef in hip/hip_runtime_api.h, so
 * this header can be included with or without the HIP headers. */
typedef struct ihipStream_t *hipStream_t;

/* Opaque library context (reference: typedef struct Hcblaslibrary*
 * hcblasHandle_t, hcblas.h:47). */
typedef struct hcblasContext *hcblasHandle_t;

/* Status codes, same order and values as hcblas.h:56-65. */
typedef enum hcblasStatus_t {
  HCBLAS_STATUS_SUCCESS = 0,
  HCBLAS_STATUS_NOT_INITIALIZED = 1,
  HCBLAS_STATUS_ALLOC_FAILED = 2,
  HCBLAS_STATUS_INVALID_VALUE = 3,
  HCBLAS_STATUS_MAPPING_ERROR = 4,
  HCBLAS_STATUS_EXECUTION_FAILED = 5,
  HCBLAS_STATUS_INTERNAL_ERROR = 6
} hcblasStatus_t;

/* Operation on a matrix operand, same values as hcblas.h:74-78. */
typedef enum hcblasOperation_t {
  HCBLAS_OP_N = 0,
  HCBLAS_OP_T = 1,
  HCBLAS_OP_C = 2
} hcblasOperation_t;

/* The reference's internal transpose codes (hcblasTranspose, hcblaslib.h:51:
 * NoTrans = 'n', Trans = 't'). Every entry point that takes an
 * hcblasOperation_t also accepts these, as aliases of HCBLAS_OP_N and
 * HCBLAS_OP_T, so callers written against the Hcblaslibrary overloads pass
 * their codes unchanged. */
typedef enum hcblasTranspose_t {
  HCBLAS_NO_TRANS = 'n',
  HCBLAS_TRANS = 't'
} hcblasTranspose_t;

/* Storage order, same values as hcblasOrder in hcblaslib.h:47
 * (RowMajor = 0, ColMajor = 1). */
typedef enum hcblasOrder_t {
  HCBLAS_ROW_MAJOR = 0,
  HCBLAS_COL_MAJOR = 1
} hcblasOrder_t;

/* Interleaved complex values, layout of hcblas.h:80-97. */
typedef struct hcFloatComplex {
  float x, y;
} hcFloatComplex;
typedef hcFloatComplex hcComplex;
typedef struct hcDoubleComplex {
  double x, y;
} hcDoubleComplex;

/* IEEE-754 binary16 bit pattern (reference: hc::half, hcblas.h:687). */
typedef uint16_t hcblasHalf;

/* Element type of the entries that take untyped (void *) operands. */
typedef enum hcblasDatatype_t {
  HCBLAS_R_32F = 0,
  HCBLAS_R_64F = 1,
  HCBLAS_C_32F = 2,
  HCBLAS_C_64F = 3,
  HCBLAS_R_16F = 4,
  HCBLAS_R_16BF = 5
} hcblasDatatype_t;

/* GEMM arithmetic selected by the handle (no reference counterpart).
 *   DEFAULT          : plain FP32 / FP64 / FP16->FP32 MFMA kernels.
 *   SPLIT_BF16       : S/C GEMM run as three-way BF16 split products on BF16
 *                      MFMA with FP32 accumulation.
 *   OZAKI_INT8       : D/Z GEMM run as Ozaki exact INT8 slices on INT8 MFMA
 *                      (scheme I: s slices, s(s+1)/2 INT8 products).
 *   OZAKI_CRT        : D/Z GEMM run as Ozaki scheme II: one INT8 GEMM per
 *                      modulus of a set of pairwise coprime moduli <= 256,
 *                      exact Chinese-remainder reconstruction in the
 *                      epilogue. Takes precedence over OZAKI_INT8.
 *   OZAKI_CRT_F32    : S/C GEMM run as Ozaki scheme II on INT8 MFMA with
 *                      operands scaled by their row / column 2-norms
 *                      (8 moduli by default). Takes precedence over
 *                      SPLIT_BF16.
 *   Modes combine with bitwise OR. In the Ozaki modes a row of A or column
 *   of B holding an Inf or NaN makes every output it feeds NaN. */
typedef enum hcblasMathMode_t {
  HCBLAS_DEFAULT_MATH = 0,
  HCBLAS_SPLIT_BF16_MATH = 1,
  HCBLAS_OZAKI_INT8_MATH = 2,
  HCBLAS_OZAKI_CRT_MATH = 4,
  HCBLAS_OZAKI_CRT_F32_MATH = 8
} hcblasMathMode_t;

/* ------------------------------------------------------------------------ */
/* Handle and stream                                                         */
/* ------------------------------------------------------------------------ */

/* Creates a context on the current HIP device.
 * Replaces hcblasCreate(hcblasHandle_t*, hc::accelerator_view*) (hcblas.h:112,
 * hcblas.cpp:39-49). */
hcblasStatus_t hcblasCreate(hcblasHandle_t *handle);

/* Releases the context and nulls *handle; a second destroy returns
 * NOT_INITIALIZED (hcblas.h:125, hcblas.cpp:72-80). */
hcblasStatus_t hcblasDestroy(hcblasHandle_t *handle);

/* Selects the stream for all later calls; NULL is the default stream.
 * Replaces hcblasSetAcclView (hcblas.h:140, hcblas.cpp:95-103) with the
 * hipblasSetStream semantics of lib/src/hcc_detail/hipblas.cpp:51-62. */
hcblasStatus_t hcblasSetStream(hcblasHandle_t handle, hipStream_t stream);

/* Returns the stream set on the handle. Replaces hcblasGetAcclView
 * (hcblas.h:155, hcblas.cpp:116-125), which never wrote its out-parameter. */
hcblasStatus_t hcblasGetStream(hcblasHandle_t handle, hipStream_t *stream);

/* GEMM arithmetic mode (no reference counterpart). */
hcblasStatus_t hcblasSetMathMode(hcblasHandle_t handle, int mode);
hcblasStatus_t hcblasGetMathMode(hcblasHandle_t handle, int *mode);

/* Number of INT8 slices per operand for the Ozaki path, 3..12
 * (no reference counterpart). Default 8. */
hcblasStatus_t hcblasSetOzakiSlices(hcblasHandle_t handle, int slices);

/* Number of moduli for the OZAKI_CRT paths, 6..20, or 0 (the default).
 * FP64: 0 picks, per K, the fewest moduli whose operand precision matches
 * the plain FP64 kernel's error, also for inputs with a wide exponent
 * spread: b >= 59 - log2(K)/2 bits, i.e. 16 moduli (55 bits) at K = 8192.
 * FP32: 0 selects 8 moduli (row / column 2-norms scaled to 2^30, more
 * accurate than the FP32 kernel also for wide exponent spreads; 7 moduli
 * scale to 2^27, 6 to 2^23) (no reference counterpart). */
hcblasStatus_t hcblasSetOzakiModuli(hcblasHandle_t handle, int moduli);

/* ------------------------------------------------------------------------ */
/* Host <-> device helpers (hcblas.h:174-233, hcblas.cpp:142-312)            */
/* Element strides and leading dimensions are honoured; a copy that runs     */
/* past the end of the device allocation returns MAPPING_ERROR.              */
/* ------------------------------------------------------------------------ */
hcblasStatus_t hcblasSetVector(hcblasHandle_t handle, int n, int elemSize,
                               const void *x, int incx, void *y, int incy);
hcblasStatus_t hcblasGetVector(hcblasHandle_t handle, int n, int elemSize,
                               const void *x, int incx, void *y, int incy);
hcblasStatus_t hcblasSetMatrix(hcblasHandle_t handle, int rows, int cols,
                               int elemSize, const void *A, int lda, void *B,
                               int ldb);
hcblasStatus_t hcblasGetMatrix(hcblasHandle_t handle, int rows, int cols,
                               int elemSize, const void *A, int lda, void *B,
                               int ldb);

/* ------------------------------------------------------------------------ */
/* Level 1                                                                  */
/* Batched forms use a batch stride of n elements (hcblas.cpp:376).         */
/* ------------------------------------------------------------------------ */

/* sum |x_i| (hcblas.h:277-284) */
hcblasStatus_t hcblasSasum(hcblasHandle_t handle, int n, const float *x,
                           int incx, float *result);
hcblasStatus_t hcblasDasum(hcblasHandle_t handle, int n, const double *x,
                           int incx, double *result);
hcblasStatus_t hcblasSasumBatched(hcblasHandle_t handle, int n, const float *x,
                                  int incx, float *result, int batchCount);
hcblasStatus_t hcblasDasumBatched(hcblasHandle_t handle, int n,
                                  const double *x, int incx, double *result,
                                  int batchCount);

/* y = alpha*x + y, computed as fma(alpha, x, y) (hcblas.h:314-320) */
hcblasStatus_t hcblasSaxpy(hcblasHandle_t handle, int n, const float *alpha,
                           const float *x, int incx, float *y, int incy);
hcblasStatus_t hcblasDaxpy(hcblasHandle_t handle, int n, const double *alpha,
                           const double *x, int incx, double *y, int incy);
hcblasStatus_t hcblasSaxpyBatched(hcblasHandle_t handle, int n,
                                  const float *alpha, const float *x, int incx,
                                  float *y, int incy, int batchCount);
hcblasStatus_t hcblasDaxpyBatched(hcblasHandle_t handle, int n,
                                  const double *alpha, const double *x,
                                  int incx, double *y, int incy,
                                  int batchCount);

/* y = x (hcblas.h:347-355) */
hcblasStatus_t hcblasScopy(hcblasHandle_t handle, int n, const float *x,
                           int incx, float *y, int incy);
hcblasStatus_t hcblasDcopy(hcblasHandle_t handle, int n, const double *x,
                           int incx, double *y, int incy);
hcblasStatus_t hcblasScopyBatched(hcblasHandle_t handle, int n, const float *x,
                                  int incx, float *y, int incy,
                                  int batchCount);
hcblasStatus_t hcblasDcopyBatched(hcblasHandle_t handle, int n,
                                  const double *x, int incx, double *y,
                                  int incy, int batchCount);

/* sum x_i*y_i (hcblas.h:384-393) */
hcblasStatus_t hcblasSdot(hcblasHandle_t handle, int n, const float *x,
                          int incx, const float *y, int incy, float *result);
hcblasStatus_t hcblasDdot(hcblasHandle_t handle, int n, const double *x,
                          int incx, const double *y, int incy, double *result);
hcblasStatus_t hcblasSdotBatched(hcblasHandle_t handle, int n, const float *x,
                                 int incx, const float *y, int incy,
                                 float *result, int batchCount);
hcblasStatus_t hcblasDdotBatched(hcblasHandle_t handle, int n,
                                 const double *x, int incx, const double *y,
                                 int incy, double *result, int batchCount);

/* x = alpha*x (hcblas.h:420-450) */
hcblasStatus_t hcblasSscal(hcblasHandle_t handle, int n, const float *alpha,
                           float *x, int incx);
hcblasStatus_t hcblasDscal(hcblasHandle_t handle, int n, const double *alpha,
                           double *x, int incx);
hcblasStatus_t hcblasCscal(hcblasHandle_t handle, int n, const hcComplex *alpha,
                           hcComplex *x, int incx);
hcblasStatus_t hcblasZscal(hcblasHandle_t handle, int n,
                           const hcDoubleComplex *alpha, hcDoubleComplex *x,
                           int incx);
hcblasStatus_t hcblasCsscal(hcblasHandle_t handle, int n, const float *alpha,
                            hcComplex *x, int incx);
hcblasStatus_t hcblasZdscal(hcblasHandle_t handle, int n, const double *alpha,
                            hcDoubleComplex *x, int incx);
hcblasStatus_t hcblasSscalBatched(hcblasHandle_t handle, int n,
                                  const float *alpha, float *x, int incx,
                                  int batchCount);
hcblasStatus_t hcblasDscalBatched(hcblasHandle_t handle, int n,
                                  const double *alpha, double *x, int incx,
                                  int batchCount);
hcblasStatus_t hcblasCscalBatched(hcblasHandle_t handle, int n,
                                  const hcComplex *alpha, hcComplex *x,
                                  int incx, int batchCount);
hcblasStatus_t hcblasZscalBatched(hcblasHandle_t handle, int n,
                                  const hcDoubleComplex *alpha,
                                  hcDoubleComplex *x, int incx, int batchCount);
hcblasStatus_t hcblasCsscalBatched(hcblasHandle_t handle, int n,
                                   const float *alpha, hcComplex *x, int incx,
                                   int batchCount);
hcblasStatus_t hcblasZdscalBatched(hcblasHandle_t handle, int n,
                                   const double *alpha, hcDoubleComplex *x,
                                   int incx, int batchCount);

/* Level 1 with per-vector offsets and explicit batch offsets: C forms of
 * the Hcblaslibrary overloads, same argument order (the accelerator_view
 * argument becomes the handle, `const T &alpha` a host pointer, the dot /
 * asum output reference a host pointer). Vector i of a batch starts at
 * X + xOffset + i*X_batchOffset. Batched dot / asum return ONE scalar summed
 * over the batch, as the reference does (sdot_array.cpp:267-277). Output
 * batches that overlap race; input batches may overlap.
 *   axpy  hcblaslib.h:106-135     scal  hcblaslib.h:319-404
 *   copy  hcblaslib.h:406-436     dot   hcblaslib.h:438-464
 *   asum  hcblaslib.h:466-479 */
hcblasStatus_t hcblas_saxpy(hcblasHandle_t handle, int N, const float *alpha,
                            const float *X, int incX, float *Y, int incY,
                            int64_t xOffset, int64_t yOffset);
hcblasStatus_t hcblas_saxpy_batched(hcblasHandle_t handle, int N,
                                    const float *alpha, const float *X,
                                    int incX, int64_t X_batchOffset, float *Y,
                                    int incY, int64_t Y_batchOffset,
                                    int64_t xOffset, int64_t yOffset,
                                    int batchSize);
hcblasStatus_t hcblas_daxpy(hcblasHandle_t handle, int N, const double *alpha,
                            const double *X, int incX, double *Y, int incY,
                            int64_t xOffset, int64_t yOffset);
hcblasStatus_t hcblas_daxpy_batched(hcblasHandle_t handle, int N,
                                    const double *alpha, const double *X,
                                    int incX, int64_t X_batchOffset, double *Y,
                                    int incY, int64_t Y_batchOffset,
                                    int64_t xOffset, int64_t yOffset,
                                    int batchSize);
hcblasStatus_t hcblas_sscal(hcblasHandle_t handle, int N, const float *alpha,
                            float *X, int incX, int64_t xOffset);
hcblasStatus_t hcblas_sscal_batched(hcblasHandle_t handle, int N,
                                    const float *alpha, float *X, int incX,
                                    int64_t xOffset, int64_t X_batchOffset,
                                    int batchSize);
hcblasStatus_t hcblas_dscal(hcblasHandle_t handle, int N, const double *alpha,
                            double *X, int incX, int64_t xOffset);
hcblasStatus_t hcblas_dscal_batched(hcblasHandle_t handle, int N,
                                    const double *alpha, double *X, int incX,
                                    int64_t xOffset, int64_t X_batchOffset,
                                    int batchSize);
hcblasStatus_t hcblas_cscal(hcblasHandle_t handle, int N,
                            const hcComplex *alpha, hcComplex *X, int incX,
                            int64_t xOffset);
hcblasStatus_t hcblas_cscal_batched(hcblasHandle_t handle, int N,
                                    const hcComplex *alpha, hcComplex *X,
                                    int incX, int64_t xOffset,
                                    int64_t X_batchOffset, int batchSize);
hcblasStatus_t hcblas_zscal(hcblasHandle_t handle, int N,
                            const hcDoubleComplex *alpha, hcDoubleComplex *X,
                            int incX, int64_t xOffset);
hcblasStatus_t hcblas_zscal_batched(hcblasHandle_t handle, int N,
                                    const hcDoubleComplex *alpha,
                                    hcDoubleComplex *X, int incX,
                                    int64_t xOffset, int64_t X_batchOffset,
                                    int batchSize);
hcblasStatus_t hcblas_csscal(hcblasHandle_t handle, int N, const float *alpha,
                             hcComplex *X, int incX, int64_t xOffset);
hcblasStatus_t hcblas_csscal_batched(hcblasHandle_t handle, int N,
                                     const float *alpha, hcComplex *X,
                                     int incX, int64_t xOffset,
                                     int64_t X_batchOffset, int batchSize);
hcblasStatus_t hcblas_zdscal(hcblasHandle_t handle, int N, const double *alpha,
                             hcDoubleComplex *X, int incX, int64_t xOffset);
hcblasStatus_t hcblas_zdscal_batched(hcblasHandle_t handle, int N,
                                     const double *alpha, hcDoubleComplex *X,
                                     int incX, int64_t xOffset,
                                     int64_t X_batchOffset, int batchSize);
hcblasStatus_t hcblas_scopy(hcblasHandle_t handle, int N, const float *X,
                            int incX, int64_t xOffset, float *Y, int incY,
                            int64_t yOffset);
hcblasStatus_t hcblas_scopy_batched(hcblasHandle_t handle, int N,
                                    const float *X, int incX, int64_t xOffset,
                                    float *Y, int incY, int64_t yOffset,
                                    int64_t X_batchOffset,
                                    int64_t Y_batchOffset, int batchSize);
hcblasStatus_t hcblas_dcopy(hcblasHandle_t handle, int N, const double *X,
                            int incX, int64_t xOffset, double *Y, int incY,
                            int64_t yOffset);
hcblasStatus_t hcblas_dcopy_batched(hcblasHandle_t handle, int N,
                                    const double *X, int incX,
                                    int64_t xOffset, double *Y, int incY,
                                    int64_t yOffset, int64_t X_batchOffset,
                                    int64_t Y_batchOffset, int batchSize);
/* `result` is the reference's `dot` output argument. */
hcblasStatus_t hcblas_sdot(hcblasHandle_t handle, int N, const float *X,
                           int incX, int64_t xOffset, const float *Y, int incY,
                           int64_t yOffset, float *result);
hcblasStatus_t hcblas_sdot_batched(hcblasHandle_t handle, int N,
                                   const float *X, int incX, int64_t xOffset,
                                   const float *Y, int incY, int64_t yOffset,
                                   float *result, int64_t X_batchOffset,
                                   int64_t Y_batchOffset, int batchSize);
hcblasStatus_t hcblas_ddot(hcblasHandle_t handle, int N, const double *X,
                           int incX, int64_t xOffset, const double *Y,
                           int incY, int64_t yOffset, double *result);
hcblasStatus_t hcblas_ddot_batched(hcblasHandle_t handle, int N,
                                   const double *X, int incX, int64_t xOffset,
                                   const double *Y, int incY, int64_t yOffset,
                                   double *result, int64_t X_batchOffset,
                                   int64_t Y_batchOffset, int batchSize);
/* `result` is the reference's `Y` output argument (a host scalar). */
hcblasStatus_t hcblas_sasum(hcblasHandle_t handle, int N, const float *X,
                            int incX, int64_t xOffset, float *result);
hcblasStatus_t hcblas_sasum_batched(hcblasHandle_t handle, int N,
                                    const float *X, int incX, int64_t xOffset,
                                    float *result, int64_t X_batchOffset,
                                    int batchSize);
hcblasStatus_t hcblas_dasum(hcblasHandle_t handle, int N, const double *X,
                            int incX, int64_t xOffset, double *result);
hcblasStatus_t hcblas_dasum_batched(hcblasHandle_t handle, int N,
                                    const double *X, int incX,
                                    int64_t xOffset, double *result,
                                    int64_t X_batchOffset, int batchSize);

/* ------------------------------------------------------------------------ */
/* Level 2 (column-major, hcblas.h:516-586)                                 */
/* Batched strides: x = len(x), y = len(y), A = m*n (hcblas.cpp:1017-1019,  */
/* 1165-1167).                                                              */
/* ------------------------------------------------------------------------ */

/* y = alpha*op(A)*x + beta*y */
hcblasStatus_t hcblasSgemv(hcblasHandle_t handle, hcblasOperation_t trans,
                           int m, int n, const float *alpha, const float *A,
                           int lda, const float *x, int incx, const float *beta,
                           float *y, int incy);
hcblasStatus_t hcblasDgemv(hcblasHandle_t handle, hcblasOperation_t trans,
                           int m, int n, const double *alpha, const double *A,
                           int lda, const double *x, int incx,
                           const double *beta, double *y, int incy);
hcblasStatus_t hcblasSgemvBatched(hcblasHandle_t handle,
                                  hcblasOperation_t trans, int m, int n,
                                  const float *alpha, const float *A, int lda,
                                  const float *x, int incx, const float *beta,
                                  float *y, int incy, int batchCount);
hcblasStatus_t hcblasDgemvBatched(hcblasHandle_t handle,
                                  hcblasOperation_t trans, int m, int n,
                                  const double *alpha, const double *A,
                                  int lda, const double *x, int incx,
                                  const double *beta, double *y, int incy,
                                  int batchCount);

/* A = alpha*x*y^T + A, computed per element as fma(x_i, alpha*y_j, a_ij) */
hcblasStatus_t hcblasSger(hcblasHandle_t handle, int m, int n,
                          const float *alpha, const float *x, int incx,
                          const float *y, int incy, float *A, int lda);
hcblasStatus_t hcblasDger(hcblasHandle_t handle, int m, int n,
                          const double *alpha, const double *x, int incx,
                          const double *y, int incy, double *A, int lda);
hcblasStatus_t hcblasSgerBatched(hcblasHandle_t handle, int m, int n,
                                 const float *alpha, const float *x, int incx,
                                 const float *y, int incy, float *A, int lda,
                                 int batchCount);
hcblasStatus_t hcblasDgerBatched(hcblasHandle_t handle, int m, int n,
                                 const double *alpha, const double *x,
                                 int incx, const double *y, int incy,
                                 double *A, int lda, int batchCount);

/* Explicit order and offsets: C form of Hcblaslibrary::hcblas_sgemv /
 * hcblas_dgemv (hcblaslib.h:178-213) and hcblas_sger / hcblas_dger
 * (hcblaslib.h:140-176). */
hcblasStatus_t hcblas_sgemv(hcblasHandle_t handle, hcblasOrder_t order,
                            hcblasOperation_t type, int M, int N,
                            const float *alpha, const float *A,
                            int64_t aOffset, int64_t lda, const float *X,
                            int64_t xOffset, int64_t incX, const float *beta,
                            float *Y, int64_t yOffset, int64_t incY);
hcblasStatus_t hcblas_dgemv(hcblasHandle_t handle, hcblasOrder_t order,
                            hcblasOperation_t type, int M, int N,
                            const double *alpha, const double *A,
                            int64_t aOffset, int64_t lda, const double *X,
                            int64_t xOffset, int64_t incX, const double *beta,
                            double *Y, int64_t yOffset, int64_t incY);
hcblasStatus_t hcblas_sger(hcblasHandle_t handle, hcblasOrder_t order, int M,
                           int N, const float *alpha, const float *X,
                           int64_t xOffset, int64_t incX, const float *Y,
                           int64_t yOffset, int64_t incY, float *A,
                           int64_t aOffset, int64_t lda);
hcblasStatus_t hcblas_dger(hcblasHandle_t handle, hcblasOrder_t order, int M,
                           int N, const double *alpha, const double *X,
                           int64_t xOffset, int64_t incX, const double *Y,
                           int64_t yOffset, int64_t incY, double *A,
                           int64_t aOffset, int64_t lda);

/* Explicit order, offsets and batch offsets: C form of the batch overloads
 * Hcblaslibrary::hcblas_{s,d}gemv (hcblaslib.h:187-190, 206-212) and
 * hcblas_{s,d}ger (hcblaslib.h:148-156, 167-175), same argument order.
 * Problem i uses A + aOffset + i*A_batchOffset, X + xOffset +
 * i*X_batchOffset and Y + yOffset + i*Y_batchOffset. Output batches that
 * overlap race. */
hcblasStatus_t hcblas_sgemv_batched(
    hcblasHandle_t handle, hcblasOrder_t order, hcblasOperation_t type, int M,
    int N, const float *alpha, const float *A, int64_t aOffset,
    int64_t A_batchOffset, int64_t lda, const float *X, int64_t xOffset,
    int64_t X_batchOffset, int64_t incX, const float *beta, float *Y,
    int64_t yOffset, int64_t Y_batchOffset, int64_t incY, int batchSize);
hcblasStatus_t hcblas_dgemv_batched(
    hcblasHandle_t handle, hcblasOrder_t order, hcblasOperation_t type, int M,
    int N, const double *alpha, const double *A, int64_t aOffset,
    int64_t A_batchOffset, int64_t lda, const double *X, int64_t xOffset,
    int64_t X_batchOffset, int64_t incX, const double *beta, double *Y,
    int64_t yOffset, int64_t Y_batchOffset, int64_t incY, int batchSize);
hcblasStatus_t hcblas_sger_batched(
    hcblasHandle_t handle, hcblasOrder_t order, int M, int N,
    const float *alpha, const float *X, int64_t xOffset,
    int64_t X_batchOffset, int64_t incX, const float *Y, int64_t yOffset,
    int64_t Y_batchOffset, int64_t incY, float *A, int64_t aOffset,
    int64_t A_batchOffset, int64_t lda, int batchSize);
hcblasStatus_t hcblas_dger_batched(
    hcblasHandle_t handle, hcblasOrder_t order, int M, int N,
    const double *alpha, const double *X, int64_t xOffset,
    int64_t X_batchOffset, int64_t incX, const double *Y, int64_t yOffset,
    int64_t Y_batchOffset, int64_t incY, double *A, int64_t aOffset,
    int64_t A_batchOffset, int64_t lda, int batchSize);

/* Allocation-free GEMV (no counterpart in the reference). For callers that
 * cannot allocate device memory at call time: code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc. Same arguments and
 * results (bit for bit) as hcblas_sgemv (dtype HCBLAS_R_32F) or hcblas_dgemv
 * (HCBLAS_R_64F), but the split partials of the reduction live in a fixed
 * device array of HCBLAS_GEMV_NOALLOC_SCRATCH_BYTES on each GPU instead of
 * the handle's grow-on-demand workspace, and the call makes no allocation.
 * A problem whose partials do not fit returns HCBLAS_STATUS_INVALID_VALUE
 * (16 MiB holds every single GEMV up to 16 GiB of A, e.g. 16384 x 16384 in
 * both types and both transposes). Calls that overlap on one GPU (other
 * streams, other handles, other threads) share the array and race; the
 * array keeps the last call's partials, none of which needs clearing. */
#define HCBLAS_GEMV_NOALLOC_SCRATCH_BYTES ((size_t)16 << 20)
hcblasStatus_t hcblas_gemv_noalloc(hcblasHandle_t handle,
                                   hcblasDatatype_t dtype, hcblasOrder_t order,
                                   hcblasOperation_t type, int M, int N,
                                   const void *alpha, const void *A,
                                   int64_t aOffset, int64_t lda, const void *X,
                                   int64_t xOffset, int64_t incX,
                                   const void *beta, void *Y, int64_t yOffset,
                                   int64_t incY);

/* ------------------------------------------------------------------------ */
/* Level 3 (column-major C API, hcblas.h:662-790)                           */
/* Aarray / Barray / Carray are DEVICE arrays of device pointers             */
/* (test/unit-api/test-blas.cpp:2016-2018).                                 */
/* ------------------------------------------------------------------------ */
hcblasStatus_t hcblasSgemm(hcblasHandle_t handle, hcblasOperation_t transa,
                           hcblasOperation_t transb, int m, int n, int k,
                           const float *alpha, const float *A, int lda,
                           const float *B, int ldb, const float *beta, float *C,
                           int ldc);
hcblasStatus_t hcblasDgemm(hcblasHandle_t handle, hcblasOperation_t transa,
                           hcblasOperation_t transb, int m, int n, int k,
                           const double *alpha, const double *A, int lda,
                           const double *B, int ldb, const double *beta,
                           double *C, int ldc);
hcblasStatus_t hcblasCgemm(hcblasHandle_t handle, hcblasOperation_t transa,
                           hcblasOperation_t transb, int m, int n, int k,
                           const hcComplex *alpha, const hcComplex *A, int lda,
                           const hcComplex *B, int ldb, const hcComplex *beta,
                           hcComplex *C, int ldc);
hcblasStatus_t hcblasZgemm(hcblasHandle_t handle, hcblasOperation_t transa,
                           hcblasOperation_t transb, int m, int n, int k,
                           const hcDoubleComplex *alpha,
                           const hcDoubleComplex *A, int lda,
                           const hcDoubleComplex *B, int ldb,
                           const hcDoubleComplex *beta, hcDoubleComplex *C,
                           int ldc);
/* FP16 in/out, FP32 accumulation; alpha/beta are binary16 like the reference
 * (hcblas.h:685-689). */
hcblasStatus_t hcblasHgemm(hcblasHandle_t handle, hcblasOperation_t transa,
                           hcblasOperation_t transb, int m, int n, int k,
                           const hcblasHalf *alpha, const hcblasHalf *A,
                           int lda, const hcblasHalf *B, int ldb,
                           const hcblasHalf *beta, hcblasHalf *C, int ldc);

hcblasStatus_t hcblasSgemmBatched(hcblasHandle_t handle,
                                  hcblasOperation_t transa,
                                  hcblasOperation_t transb, int m, int n, int k,
                                  const float *alpha,
                                  const float *const Aarray[], int lda,
                                  const float *const Barray[], int ldb,
                                  const float *beta, float *const Carray[],
                                  int ldc, int batchCount);
hcblasStatus_t hcblasDgemmBatched(hcblasHandle_t handle,
                                  hcblasOperation_t transa,
                                  hcblasOperation_t transb, int m, int n, int k,
                                  const double *alpha,
                                  const double *const Aarray[], int lda,
                                  const double *const Barray[], int ldb,
                                  const double *beta, double *const Carray[],
                                  int ldc, int batchCount);
hcblasStatus_t hcblasCgemmBatched(hcblasHandle_t handle,
                                  hcblasOperation_t transa,
                                  hcblasOperation_t transb, int m, int n, int k,
                                  const hcComplex *alpha,
                                  const hcComplex *const Aarray[], int lda,
                                  const hcComplex *const Barray[], int ldb,
                                  const hcComplex *beta,
                                  hcComplex *const Carray[], int ldc,
                                  int batchCount);
hcblasStatus_t hcblasZgemmBatched(hcblasHandle_t handle,
                                  hcblasOperation_t transa,
                                  hcblasOperation_t transb, int m, int n, int k,
                                  const hcDoubleComplex *alpha,
                                  const hcDoubleComplex *const Aarray[],
                                  int lda,
                                  const hcDoubleComplex *const Barray[],
                                  int ldb, const hcDoubleComplex *beta,
                                  hcDoubleComplex *const Carray[], int ldc,
                                  int batchCount);
/* No reference counterpart (hgemm_array_wrapper.cpp:315+ is commented out). */
hcblasStatus_t hcblasHgemmBatched(hcblasHandle_t handle,
                                  hcblasOperation_t transa,
                                  hcblasOperation_t transb, int m, int n, int k,
                                  const hcblasHalf *alpha,
                                  const hcblasHalf *const Aarray[], int lda,
                                  const hcblasHalf *const Barray[], int ldb,
                                  const hcblasHalf *beta,
                                  hcblasHalf *const Carray[], int ldc,
                                  int batchCount);

/* Explicit order and per-matrix offsets: C form of
 * Hcblaslibrary::hcblas_{s,d,h,c,z}gemm (hcblaslib.h:215-242, 268-293).
 * Argument order follows the reference overloads; typeA / typeB take
 * HCBLAS_OP_N/T/C or the reference's 'n' / 't'. The return value is the
 * hcblasStatus_t of the C API: HCBLAS_STATUS_SUCCESS where the overload
 * returned HCBLAS_SUCCEEDS, a non-zero code (INVALID_VALUE for a bad
 * argument) where it returned HCBLAS_INVALID. */
hcblasStatus_t hcblas_sgemm(hcblasHandle_t handle, hcblasOrder_t order,
                            hcblasOperation_t typeA, hcblasOperation_t typeB,
                            int M, int N, int K, const float *alpha,
                            const float *A, int64_t lda, const float *B,
                            int64_t ldb, const float *beta, float *C,
                            int64_t ldc, int64_t aOffset, int64_t bOffset,
                            int64_t cOffset);
hcblasStatus_t hcblas_dgemm(hcblasHandle_t handle, hcblasOrder_t order,
                            hcblasOperation_t typeA, hcblasOperation_t typeB,
                            int M, int N, int K, const double *alpha,
                            const double *A, int64_t lda, const double *B,
                            int64_t ldb, const double *beta, double *C,
                            int64_t ldc, int64_t aOffset, int64_t bOffset,
                            int64_t cOffset);
/* alpha / beta are binary16, as the reference's `const hc::half &`
 * (hcblaslib.h:236-238); they are widened exactly to FP32 for the
 * FP32-accumulating kernel. */
hcblasStatus_t hcblas_hgemm(hcblasHandle_t handle, hcblasOrder_t order,
                            hcblasOperation_t typeA, hcblasOperation_t typeB,
                            int M, int N, int K, const hcblasHalf *alpha,
                            const hcblasHalf *A, int64_t lda,
                            const hcblasHalf *B, int64_t ldb,
                            const hcblasHalf *beta, hcblasHalf *C, int64_t ldc,
                            int64_t aOffset, int64_t bOffset, int64_t cOffset);
hcblasStatus_t hcblas_cgemm(hcblasHandle_t handle, hcblasOrder_t order,
                            hcblasOperation_t typeA, hcblasOperation_t typeB,
                            int M, int N, int K, const hcComplex *alpha,
                            const hcComplex *A, int64_t aOffset, int64_t lda,
                            const hcComplex *B, int64_t bOffset, int64_t ldb,
                            const hcComplex *beta, hcComplex *C,
                            int64_t cOffset, int64_t ldc);
hcblasStatus_t hcblas_zgemm(hcblasHandle_t handle, hcblasOrder_t order,
                            hcblasOperation_t typeA, hcblasOperation_t typeB,
                            int M, int N, int K, const hcDoubleComplex *alpha,
                            const hcDoubleComplex *A, int64_t aOffset,
                            int64_t lda, const hcDoubleComplex *B,
                            int64_t bOffset, int64_t ldb,
                            const hcDoubleComplex *beta, hcDoubleComplex *C,
                            int64_t cOffset, int64_t ldc);

/* Uniform strided batch with explicit order: matrix i of A starts at
 * A + aOffset + i*strideA. Replaces the strided row-major batch kernels that
 * the reference never reaches (sgemm_array_kernels_rMajor_batch.cpp,
 * sgemm_array_wrapper.cpp:113-151). dtype selects the element type of A, B, C
 * (alpha/beta point to that type; for HCBLAS_R_16F they point to float). */
hcblasStatus_t hcblas_gemm_strided_batched(
    hcblasHandle_t handle, hcblasDatatype_t dtype, hcblasOrder_t order,
    hcblasOperation_t typeA, hcblasOperation_t typeB, int M, int N, int K,
    const void *alpha, const void *A, int64_t lda, int64_t aOffset,
    int64_t strideA, const void *B, int64_t ldb, int64_t bOffset,
    int64_t strideB, const void *beta, void *C, int64_t ldc, int64_t cOffset,
    int64_t strideC, int batchCount);

/* Pointer-array batch with explicit order and per-matrix offsets: matrix i of
 * A is Aarray[i] + aOffset (hcblaslib.h:244-266, 294-315). */
hcblasStatus_t hcblas_gemm_batched(
    hcblasHandle_t handle, hcblasDatatype_t dtype, hcblasOrder_t order,
    hcblasOperation_t typeA, hcblasOperation_t typeB, int M, int N, int K,
    const void *alpha, const void *const Aarray[], int64_t lda, int64_t aOffset,
    const void *const Barray[], int64_t ldb, int64_t bOffset, const void *beta,
    void *const Carray[], int64_t ldc, int64_t cOffset, int batchCount);

/* ------------------------------------------------------------------------ */
/* Split-precision GEMM (no reference counterpart)                          */
/* Same arguments as hcblas_sgemm / hcblas_dgemm / hcblas_cgemm /           */
/* hcblas_zgemm; they run regardless of the handle's math mode.             */
/* ------------------------------------------------------------------------ */
hcblasStatus_t hcblas_sgemm_split_bf16(hcblasHandle_t handle,
                                       hcblasOrder_t order,
                                       hcblasOperation_t typeA,
                                       hcblasOperation_t typeB, int M, int N,
                                       int K, const float *alpha,
                                       const float *A, int64_t lda,
                                       const float *B, int64_t ldb,
                                       const float *beta, float *C,
                                       int64_t ldc, int64_t aOffset,
                                       int64_t bOffset, int64_t cOffset);
hcblasStatus_t hcblas_sgemm_ozaki_crt(hcblasHandle_t handle,
                                      hcblasOrder_t order,
                                      hcblasOperation_t typeA,
                                      hcblasOperation_t typeB, int M, int N,
                                      int K, const float *alpha,
                                      const float *A, int64_t lda,
                                      const float *B, int64_t ldb,
                                      const float *beta, float *C,
                                      int64_t ldc, int64_t aOffset,
                                      int64_t bOffset, int64_t cOffset);
hcblasStatus_t hcblas_dgemm_ozaki_int8(hcblasHandle_t handle,
                                       hcblasOrder_t order,
                                       hcblasOperation_t typeA,
                                       hcblasOperation_t typeB, int M, int N,
                                       int K, const double *alpha,
                                       const double *A, int64_t lda,
                                       const double *B, int64_t ldb,
                                       const double *beta, double *C,
                                       int64_t ldc, int64_t aOffset,
                                       int64_t bOffset, int64_t cOffset);
hcblasStatus_t hcblas_dgemm_ozaki_crt(hcblasHandle_t handle,
                                      hcblasOrder_t order,
                                      hcblasOperation_t typeA,
                                      hcblasOperation_t typeB, int M, int N,
                                      int K, const double *alpha,
                                      const double *A, int64_t lda,
                                      const double *B, int64_t ldb,
                                      const double *beta, double *C,
                                      int64_t ldc, int64_t aOffset,
                                      int64_t bOffset, int64_t cOffset);

/* ------------------------------------------------------------------------ */
/* Adjoints (no reference counterpart)                                       */
/* ------------------------------------------------------------------------ */

/* Reverse-mode adjoints of hcblas_gemm_strided_batched,
 *   C = alpha * op(A) * op(B) + beta * C_in,
 * for a cotangent dC of C (same order; A, B, dC at offset 0):
 *   dA    = alpha * dC * op(B)^T, stored like A (alpha * op(B) * dC^T when
 *           op(A) = A^T);
 *   dB    = alpha * op(A)^T * dC, stored like B (alpha * dC^T * op(A) when
 *           op(B) = B^T);
 *   dC_in = beta * dC.
 * dtype HCBLAS_R_32F / HCBLAS_R_64F: every matrix of that type, alpha / beta
 * point to float / double. HCBLAS_R_16BF: A, B, dA and dB are bfloat16 (dA
 * and dB accumulated in FP32 from the exact BF16 split of dC and rounded
 * once to nearest even), dC and dC_in are float (the forward's accumulation
 * type), alpha / beta are float. dA, dB or dCin may be NULL to skip that
 * adjoint. op = C is treated as T for these real types. The handle's math
 * mode applies to the FP32 / FP64 adjoint GEMMs as to any GEMM. */
hcblasStatus_t hcblas_gemm_strided_batched_backward(
    hcblasHandle_t handle, hcblasDatatype_t dtype, hcblasOrder_t order,
    hcblasOperation_t typeA, hcblasOperation_t typeB, int M, int N, int K,
    const void *alpha, const void *beta, const void *A, int64_t lda,
    int64_t strideA, const void *B, int64_t ldb, int64_t strideB,
    const void *dC, int64_t lddc, int64_t strideDC, void *dA, int64_t ldda,
    int64_t strideDA, void *dB, int64_t lddb, int64_t strideDB, void *dCin,
    int64_t lddcin, int64_t strideDCin, int batchCount);

/* Adjoint of hcblas_sger / hcblas_dger, A_out = alpha * x * y^T + A_in, for a
 * cotangent dA (M x N in the given order): dx = alpha * dA * y and
 * dy = alpha * dA^T * x, written with unit stride (either may be NULL);
 * dA_in = dA. */
hcblasStatus_t hcblas_sger_backward(hcblasHandle_t handle,
                                    hcblasOrder_t order, int M, int N,
                                    const float *alpha, const float *X,
                                    int64_t incX, const float *Y, int64_t incY,
                                    const float *dA, int64_t ldda, float *dX,
                                    float *dY);
hcblasStatus_t hcblas_dger_backward(hcblasHandle_t handle,
                                    hcblasOrder_t order, int M, int N,
                                    const double *alpha, const double *X,
                                    int64_t incX, const double *Y,
                                    int64_t incY, const double *dA,
                                    int64_t ldda, double *dX, double *dY);

/* ------------------------------------------------------------------------ */
/* Grouped GEMM (no reference counterpart; the reference's batched GEMM     */
/* takes one shape for the whole batch, sgemm_array_kernels_batch.cpp)      */
/* ------------------------------------------------------------------------ */

/* One problem of a group: C = alpha*op(A)*op(B) + beta*C with its own dtype
 * (HCBLAS_R_32F, or HCBLAS_R_16F = binary16 A/B/C with fp32 accumulation),
 * storage order, transposes (N or T), sizes and leading dimensions. */
typedef struct hcblasGemmProblem {
  int32_t dtype;
  int32_t order;
  int32_t transA;
  int32_t transB;
  int32_t M, N, K;
  int32_t reserved;
  float alpha, beta;
  const void *A;
  int64_t lda;
  const void *B;
  int64_t ldb;
  void *C;
  int64_t ldc;
} hcblasGemmProblem;

typedef struct hcblasGroupedPlan *hcblasGroupedPlan_t;

/* Validates a HOST problem table, copies it to the device in canonical form,
 * builds the tile work list and captures the persistent launch in a hipGraph.
 * The device pointers in the table must stay valid while the plan is used. */
hcblasStatus_t hcblasGroupedPlanCreate(hcblasHandle_t handle,
                                       const hcblasGemmProblem *problems,
                                       int count, hcblasGroupedPlan_t *plan);
/* Times every (tile shape, work order) candidate `reps` times on the plan's
 * own problems -- which executes them, so C is updated each time -- and
 * keeps the fastest. */
hcblasStatus_t hcblasGroupedPlanAutotune(hcblasHandle_t handle,
                                         hcblasGroupedPlan_t plan, int reps);
/* Runs every problem of the plan: one graph launch on the handle's stream. */
hcblasStatus_t hcblasGroupedGemm(hcblasHandle_t handle,
                                 hcblasGroupedPlan_t plan);
/* Tile edge, work order (0 longest first, 1 table order), number of tiles and
 * persistent workgroups of the plan's current schedule. */
hcblasStatus_t hcblasGroupedPlanInfo(hcblasGroupedPlan_t plan, int *tile,
                                     int *order, int *tiles, int *workgroups);
hcblasStatus_t hcblasGroupedPlanDestroy(hcblasGroupedPlan_t plan);

/* ------------------------------------------------------------------------ */
/* Library information                                                      */
/* ------------------------------------------------------------------------ */

/* Writes a NUL-terminated build description (arch, version). */
hcblasStatus_t hcblasGetVersionString(char *buf, int len);

#ifdef __cplusplus
}
#endif

#endif /* HCBLAS_MI355X_HCBLAS_H_ */
