/*
 * hcblaslib.hpp — header-only C++ mirror of the reference's internal
 * `Hcblaslibrary` interface (lib/include/hcblaslib.h:67-480), implemented on
 * the C ABI of hcblas.h. Host code only; compiles with g++ or hipcc.
 *
 * Code written against the reference's overloads (its lib-level unit tests,
 * test/unit/hcblas_*.cpp, call them directly) ports by changing the type of
 * the accelerator-view argument:
 *   - `hc::accelerator_view`     -> `hipStream_t` (the stream the call runs
 *                                   on; NULL is the default stream);
 *   - `Hcblaslibrary hc(&av)`    -> `Hcblaslibrary hc(&stream)`;
 *   - `hc.currentAcclView`       -> a `hipStream_t`;
 *   - `hc::short_vector::float_2 / double_2` -> `hcComplex / hcDoubleComplex`
 *     (same {x, y} layout); `hc::half` -> `hcblasHalf` (binary16 bits).
 * Names, argument order, the enums hcblasStatus {HCBLAS_SUCCEEDS = 0,
 * HCBLAS_INVALID = -1}, hcblasOrder {RowMajor, ColMajor} and
 * hcblasTranspose {NoTrans = 'n', Trans = 't'} (hcblaslib.h:44-51) and the
 * argument checks are the reference's: each overload returns HCBLAS_INVALID
 * for the inputs the reference rejects (NULL operands, N <= 0, non-positive
 * increments, a zero GEMM dimension; e.g. saxpy_array.cpp:186,
 * sgemm_array_wrapper.cpp:306), and for any failure of the call itself.
 * The arithmetic is the C ABI's (CBLAS semantics, FP32-accumulating HGEMM).
 * The reference's hcblasTranspose has no conjugate code, so HCBLAS_OP_C is
 * not reachable through this class.
 *
 * Memory: matrices and vectors are device pointers; the pointer-array
 * overloads take a device array of device pointers, and their
 * A/B/C_batchOffset is added to every matrix of the batch (the reference's
 * batch kernels index A[i][aOffset + A_batchOffset + ...],
 * sgemm_array_kernels_batch.cpp:56). Level-1 / level-2 batch overloads step
 * vector i of a batch by i * X_batchOffset. dot / asum write a host scalar
 * before returning.
 */
#ifndef HCBLAS_MI355X_HCBLASLIB_HPP_
#define HCBLAS_MI355X_HCBLASLIB_HPP_

#include <cstdint>

#include "hcblas.h"

/* Status of an Hcblaslibrary call (hcblaslib.h:44). */
enum hcblasStatus { HCBLAS_SUCCEEDS = 0, HCBLAS_INVALID = -1 };

/* Storage order (hcblaslib.h:47). */
enum hcblasOrder : unsigned short { RowMajor, ColMajor };

/* Operation on a matrix operand (hcblaslib.h:51). */
enum hcblasTranspose { NoTrans = 'n', Trans = 't' };

struct Hcblaslibrary {
 public:
  /* Creates a context on the current HIP device that runs on *av (the
   * default stream when av or *av is NULL). Replaces
   * Hcblaslibrary(hc::accelerator_view *) (hcblaslib.h:69-82). */
  explicit Hcblaslibrary(hipStream_t *av = nullptr)
      : currentAcclView(av ? *av : nullptr) {
    initialized = hcblasCreate(&handle) == HCBLAS_STATUS_SUCCESS;
    if (initialized)
      initialized =
          hcblasSetStream(handle, currentAcclView) == HCBLAS_STATUS_SUCCESS;
  }
  ~Hcblaslibrary() {
    if (handle) hcblasDestroy(&handle);
    initialized = false;
  }
  Hcblaslibrary(const Hcblaslibrary &) = delete;
  Hcblaslibrary &operator=(const Hcblaslibrary &) = delete;

  bool initialized = false;
  hipStream_t currentAcclView = nullptr;
  void *currentStream = nullptr;
  hcblasOrder Order = ColMajor;
  /* The C-ABI context behind this object. */
  hcblasHandle_t handle = nullptr;

  /* ---------------------------------------------------------- AXPY */
  hcblasStatus hcblas_saxpy(hipStream_t accl_view, const int N,
                            const float &alpha, const float *X, const int incX,
                            float *Y, const int incY, const int64_t xOffset,
                            const int64_t yOffset) {
    if (X == nullptr || Y == nullptr || N <= 0 || incX <= 0 || incY <= 0)
      return HCBLAS_INVALID;
    return done(on(accl_view) &&
                ::hcblas_saxpy(handle, N, &alpha, X, incX, Y, incY, xOffset,
                               yOffset) == 0);
  }
  hcblasStatus hcblas_saxpy(hipStream_t accl_view, const int N,
                            const float &alpha, const float *X, const int incX,
                            const int64_t X_batchOffset, float *Y,
                            const int incY, const int64_t Y_batchOffset,
                            const int64_t xOffset, const int64_t yOffset,
                            const int batchSize) {
    if (X == nullptr || Y == nullptr || N <= 0 || incX <= 0 || incY <= 0)
      return HCBLAS_INVALID;
    return done(on(accl_view) &&
                ::hcblas_saxpy_batched(handle, N, &alpha, X, incX,
                                       X_batchOffset, Y, incY, Y_batchOffset,
                                       xOffset, yOffset, batchSize) == 0);
  }
  hcblasStatus hcblas_daxpy(hipStream_t accl_view, const int N,
                            const double &alpha, const double *X,
                            const int incX, double *Y, const int incY,
                            const int64_t xOffset, const int64_t yOffset) {
    if (X == nullptr || Y == nullptr || N <= 0 || incX <= 0 || incY <= 0)
      return HCBLAS_INVALID;
    return done(on(accl_view) &&
                ::hcblas_daxpy(handle, N, &alpha, X, incX, Y, incY, xOffset,
                               yOffset) == 0);
  }
  hcblasStatus hcblas_daxpy(hipStream_t accl_view, const int N,
                            const double &alpha, const double *X,
                            const int incX, const int64_t X_batchOffset,
                            double *Y, const int incY,
                            const int64_t Y_batchOffset, const int64_t xOffset,
                            const int64_t yOffset, const int batchSize) {
    if (X == nullptr || Y == nullptr || N <= 0 || incX <= 0 || incY <= 0)
      return HCBLAS_INVALID;
    return done(on(accl_view) &&
                ::hcblas_daxpy_batched(handle, N, &alpha, X, incX,
                                       X_batchOffset, Y, incY, Y_batchOffset,
                                       xOffset, yOffset, batchSize) == 0);
  }

  /* ----------------------------------------------------------- GER */
  hcblasStatus hcblas_sger(hipStream_t accl_view, hcblasOrder order,
                           const int M, const int N, const float &alpha,
                           const float *X, const int64_t xOffset,
                           const int incX, const float *Y,
                           const int64_t yOffset, const int incY, float *A,
                           const int64_t aOffset, const int lda) {
    if (X == nullptr || Y == nullptr || A == nullptr || N <= 0 || M <= 0 ||
        incX == 0 || incY == 0)
      return HCBLAS_INVALID;
    return done(on(accl_view) &&
                ::hcblas_sger(handle, ord(order), M, N, &alpha, X, xOffset,
                              incX, Y, yOffset, incY, A, aOffset, lda) == 0);
  }
  hcblasStatus hcblas_sger(hipStream_t accl_view, hcblasOrder order,
                           const int M, const int N, const float &alpha,
                           const float *X, const int64_t xOffset,
                           const int64_t X_batchOffset, const int incX,
                           const float *Y, const int64_t yOffset,
                           const int64_t Y_batchOffset, const int incY,
                           float *A, const int64_t aOffset,
                           const int64_t A_batchOffset, const int lda,
                           const int batchSize) {
    if (X == nullptr || Y == nullptr || A == nullptr || N <= 0 || M <= 0 ||
        incX == 0 || incY == 0)
      return HCBLAS_INVALID;
    return done(on(accl_view) &&
                ::hcblas_sger_batched(handle, ord(order), M, N, &alpha, X,
                                      xOffset, X_batchOffset, incX, Y, yOffset,
                                      Y_batchOffset, incY, A, aOffset,
                                      A_batchOffset, lda, batchSize) == 0);
  }
  hcblasStatus hcblas_dger(hipStream_t accl_view, hcblasOrder order,
                           const int M, const int N, const double &alpha,
                           const double *X, const int64_t xOffset,
                           const int incX, const double *Y,
                           const int64_t yOffset, const int incY, double *A,
                           const int64_t aOffset, const int lda) {
    if (X == nullptr || Y == nullptr || A == nullptr || N <= 0 || M <= 0 ||
        incX == 0 || incY == 0)
      return HCBLAS_INVALID;
    return done(on(accl_view) &&
                ::hcblas_dger(handle, ord(order), M, N, &alpha, X, xOffset,
                              incX, Y, yOffset, incY, A, aOffset, lda) == 0);
  }
  hcblasStatus hcblas_dger(hipStream_t accl_view, hcblasOrder order,
                           const int M, const int N, const double &alpha,
                           const double *X, const int64_t xOffset,
                           const int64_t X_batchOffset, const int incX,
                           const double *Y, const int64_t yOffset,
                           const int64_t Y_batchOffset, const int incY,
                           double *A, const int64_t aOffset,
                           const int64_t A_batchOffset, const int lda,
                           const int batchSize) {
    if (X == nullptr || Y == nullptr || A == nullptr || N <= 0 || M <= 0 ||
        incX == 0 || incY == 0)
      return HCBLAS_INVALID;
    return done(on(accl_view) &&
                ::hcblas_dger_batched(handle, ord(order), M, N, &alpha, X,
                                      xOffset, X_batchOffset, incX, Y, yOffset,
                                      Y_batchOffset, incY, A, aOffset,
                                      A_batchOffset, lda, batchSize) == 0);
  }

  /* ---------------------------------------------------------- GEMV */
  hcblasStatus hcblas_sgemv(hipStream_t accl_view, hcblasOrder order,
                            hcblasTranspose type, const int M, const int N,
                            const float &alpha, float *A,
                            const int64_t aOffset, const int lda, float *X,
                            const int64_t xOffset, const int incX,
                            const float &beta, float *Y,
                            const int64_t yOffset, const int incY) {
    if (X == nullptr || Y == nullptr || A == nullptr || M <= 0 || N <= 0 ||
        incX <= 0 || incY <= 0)
      return HCBLAS_INVALID;
    return done(on(accl_view) &&
                ::hcblas_sgemv(handle, ord(order), op(type), M, N, &alpha, A,
                               aOffset, lda, X, xOffset, incX, &beta, Y,
                               yOffset, incY) == 0);
  }
  hcblasStatus hcblas_sgemv(hipStream_t accl_view, hcblasOrder order,
                            hcblasTranspose type, const int M, const int N,
                            const float &alpha, float *A,
                            const int64_t aOffset, const int64_t A_batchOffset,
                            const int lda, float *X, const int64_t xOffset,
                            const int64_t X_batchOffset, const int incX,
                            const float &beta, float *Y,
                            const int64_t yOffset, const int64_t Y_batchOffset,
                            const int incY, const int batchSize) {
    if (X == nullptr || Y == nullptr || A == nullptr || M <= 0 || N <= 0 ||
        incX <= 0 || incY <= 0)
      return HCBLAS_INVALID;
    return done(on(accl_view) &&
                ::hcblas_sgemv_batched(handle, ord(order), op(type), M, N,
                                       &alpha, A, aOffset, A_batchOffset, lda,
                                       X, xOffset, X_batchOffset, incX, &beta,
                                       Y, yOffset, Y_batchOffset, incY,
                                       batchSize) == 0);
  }
  hcblasStatus hcblas_dgemv(hipStream_t accl_view, hcblasOrder order,
                            hcblasTranspose type, const int M, const int N,
                            const double &alpha, double *A,
                            const int64_t aOffset, const int lda, double *X,
                            const int64_t xOffset, const int incX,
                            const double &beta, double *Y,
                            const int64_t yOffset, const int incY) {
    if (X == nullptr || Y == nullptr || A == nullptr || M <= 0 || N <= 0 ||
        incX <= 0 || incY <= 0)
      return HCBLAS_INVALID;
    return done(on(accl_view) &&
                ::hcblas_dgemv(handle, ord(order), op(type), M, N, &alpha, A,
                               aOffset, lda, X, xOffset, incX, &beta, Y,
                               yOffset, incY) == 0);
  }
  hcblasStatus hcblas_dgemv(hipStream_t accl_view, hcblasOrder order,
                            hcblasTranspose type, const int M, const int N,
                            const double &alpha, double *A,
                            const int64_t aOffset, const int64_t A_batchOffset,
                            const int lda, double *X, const int64_t xOffset,
                            const int64_t X_batchOffset, const int incX,
                            const double &beta, double *Y,
                            const int64_t yOffset, const int64_t Y_batchOffset,
                            const int incY, const int batchSize) {
    if (X == nullptr || Y == nullptr || A == nullptr || M <= 0 || N <= 0 ||
        incX <= 0 || incY <= 0)
      return HCBLAS_INVALID;
    return done(on(accl_view) &&
                ::hcblas_dgemv_batched(handle, ord(order), op(type), M, N,
                                       &alpha, A, aOffset, A_batchOffset, lda,
                                       X, xOffset, X_batchOffset, incX, &beta,
                                       Y, yOffset, Y_batchOffset, incY,
                                       batchSize) == 0);
  }

  /* ---------------------------------------------------------- GEMM */
  hcblasStatus hcblas_sgemm(hipStream_t accl_view, hcblasOrder order,
                            hcblasTranspose typeA, hcblasTranspose typeB,
                            const int M, const int N, const int K,
                            const float &alpha, float *A, const int64_t lda,
                            float *B, const int64_t ldb, const float &beta,
                            float *C, const int64_t ldc, const int64_t aOffset,
                            const int64_t bOffset, const int64_t cOffset) {
    if (A == nullptr || B == nullptr || C == nullptr || !M || !N || !K)
      return HCBLAS_INVALID;
    return done(on(accl_view) &&
                ::hcblas_sgemm(handle, ord(order), op(typeA), op(typeB), M, N,
                               K, &alpha, A, lda, B, ldb, &beta, C, ldc,
                               aOffset, bOffset, cOffset) == 0);
  }
  hcblasStatus hcblas_dgemm(hipStream_t accl_view, hcblasOrder order,
                            hcblasTranspose typeA, hcblasTranspose typeB,
                            const int M, const int N, const int K,
                            const double &alpha, double *A, const int64_t lda,
                            double *B, const int64_t ldb, const double &beta,
                            double *C, const int64_t ldc,
                            const int64_t aOffset, const int64_t bOffset,
                            const int64_t cOffset) {
    if (A == nullptr || B == nullptr || C == nullptr || !M || !N || !K)
      return HCBLAS_INVALID;
    return done(on(accl_view) &&
                ::hcblas_dgemm(handle, ord(order), op(typeA), op(typeB), M, N,
                               K, &alpha, A, lda, B, ldb, &beta, C, ldc,
                               aOffset, bOffset, cOffset) == 0);
  }
  hcblasStatus hcblas_hgemm(hipStream_t accl_view, hcblasOrder order,
                            hcblasTranspose typeA, hcblasTranspose typeB,
                            const int M, const int N, const int K,
                            const hcblasHalf &alpha, hcblasHalf *A,
                            const int64_t lda, hcblasHalf *B,
                            const int64_t ldb, const hcblasHalf &beta,
                            hcblasHalf *C, const int64_t ldc,
                            const int64_t aOffset, const int64_t bOffset,
                            const int64_t cOffset) {
    if (A == nullptr || B == nullptr || C == nullptr || !M || !N || !K)
      return HCBLAS_INVALID;
    return done(on(accl_view) &&
                ::hcblas_hgemm(handle, ord(order), op(typeA), op(typeB), M, N,
                               K, &alpha, A, lda, B, ldb, &beta, C, ldc,
                               aOffset, bOffset, cOffset) == 0);
  }
  hcblasStatus hcblas_sgemm(hipStream_t accl_view, hcblasOrder order,
                            hcblasTranspose typeA, hcblasTranspose typeB,
                            const int M, const int N, const int K,
                            const float &alpha, float *A[], const int64_t lda,
                            const int64_t A_batchOffset, float *B[],
                            const int64_t ldb, const int64_t B_batchOffset,
                            const float &beta, float *C[], const int64_t ldc,
                            const int64_t C_batchOffset, const int64_t aOffset,
                            const int64_t bOffset, const int64_t cOffset,
                            const int batchSize) {
    if (A == nullptr || B == nullptr || C == nullptr || !M || !N || !K)
      return HCBLAS_INVALID;
    return batched(accl_view, HCBLAS_R_32F, order, typeA, typeB, M, N, K,
                   &alpha, A, lda, aOffset + A_batchOffset, B, ldb,
                   bOffset + B_batchOffset, &beta, C, ldc,
                   cOffset + C_batchOffset, batchSize);
  }
  hcblasStatus hcblas_dgemm(hipStream_t accl_view, hcblasOrder order,
                            hcblasTranspose typeA, hcblasTranspose typeB,
                            const int M, const int N, const int K,
                            const double &alpha, double *A[],
                            const int64_t lda, const int64_t A_batchOffset,
                            double *B[], const int64_t ldb,
                            const int64_t B_batchOffset, const double &beta,
                            double *C[], const int64_t ldc,
                            const int64_t C_batchOffset, const int64_t aOffset,
                            const int64_t bOffset, const int64_t cOffset,
                            const int batchSize) {
    if (A == nullptr || B == nullptr || C == nullptr || !M || !N || !K)
      return HCBLAS_INVALID;
    return batched(accl_view, HCBLAS_R_64F, order, typeA, typeB, M, N, K,
                   &alpha, A, lda, aOffset + A_batchOffset, B, ldb,
                   bOffset + B_batchOffset, &beta, C, ldc,
                   cOffset + C_batchOffset, batchSize);
  }
  hcblasStatus hcblas_cgemm(hipStream_t accl_view, hcblasOrder order,
                            hcblasTranspose typeA, hcblasTranspose typeB,
                            const int M, const int N, const int K,
                            const hcComplex &alpha, hcComplex *A,
                            const int64_t aOffset, const int64_t lda,
                            hcComplex *B, const int64_t bOffset,
                            const int64_t ldb, const hcComplex &beta,
                            hcComplex *C, const int64_t cOffset,
                            const int64_t ldc) {
    if (A == nullptr || B == nullptr || C == nullptr || !M || !N || !K)
      return HCBLAS_INVALID;
    return done(on(accl_view) &&
                ::hcblas_cgemm(handle, ord(order), op(typeA), op(typeB), M, N,
                               K, &alpha, A, aOffset, lda, B, bOffset, ldb,
                               &beta, C, cOffset, ldc) == 0);
  }
  hcblasStatus hcblas_zgemm(hipStream_t accl_view, hcblasOrder order,
                            hcblasTranspose typeA, hcblasTranspose typeB,
                            const int M, const int N, const int K,
                            const hcDoubleComplex &alpha, hcDoubleComplex *A,
                            const int64_t aOffset, const int64_t lda,
                            hcDoubleComplex *B, const int64_t bOffset,
                            const int64_t ldb, const hcDoubleComplex &beta,
                            hcDoubleComplex *C, const int64_t cOffset,
                            const int64_t ldc) {
    if (A == nullptr || B == nullptr || C == nullptr || !M || !N || !K)
      return HCBLAS_INVALID;
    return done(on(accl_view) &&
                ::hcblas_zgemm(handle, ord(order), op(typeA), op(typeB), M, N,
                               K, &alpha, A, aOffset, lda, B, bOffset, ldb,
                               &beta, C, cOffset, ldc) == 0);
  }
  hcblasStatus hcblas_cgemm(hipStream_t accl_view, hcblasOrder order,
                            hcblasTranspose typeA, hcblasTranspose typeB,
                            const int M, const int N, const int K,
                            const hcComplex &alpha, hcComplex *A[],
                            const int64_t aOffset, const int64_t A_batchOffset,
                            const int64_t lda, hcComplex *B[],
                            const int64_t bOffset, const int64_t B_batchOffset,
                            const int64_t ldb, const hcComplex &beta,
                            hcComplex *C[], const int64_t cOffset,
                            const int64_t C_batchOffset, const int64_t ldc,
                            const int batchSize) {
    if (A == nullptr || B == nullptr || C == nullptr || !M || !N || !K)
      return HCBLAS_INVALID;
    return batched(accl_view, HCBLAS_C_32F, order, typeA, typeB, M, N, K,
                   &alpha, A, lda, aOffset + A_batchOffset, B, ldb,
                   bOffset + B_batchOffset, &beta, C, ldc,
                   cOffset + C_batchOffset, batchSize);
  }
  hcblasStatus hcblas_zgemm(hipStream_t accl_view, hcblasOrder order,
                            hcblasTranspose typeA, hcblasTranspose typeB,
                            const int M, const int N, const int K,
                            const hcDoubleComplex &alpha, hcDoubleComplex *A[],
                            const int64_t aOffset, const int64_t A_batchOffset,
                            const int64_t lda, hcDoubleComplex *B[],
                            const int64_t bOffset, const int64_t B_batchOffset,
                            const int64_t ldb, const hcDoubleComplex &beta,
                            hcDoubleComplex *C[], const int64_t cOffset,
                            const int64_t C_batchOffset, const int64_t ldc,
                            const int batchSize) {
    if (A == nullptr || B == nullptr || C == nullptr || !M || !N || !K)
      return HCBLAS_INVALID;
    return batched(accl_view, HCBLAS_C_64F, order, typeA, typeB, M, N, K,
                   &alpha, A, lda, aOffset + A_batchOffset, B, ldb,
                   bOffset + B_batchOffset, &beta, C, ldc,
                   cOffset + C_batchOffset, batchSize);
  }

  /* ---------------------------------------------------------- SCAL */
#define HCBLASLIB_SCAL(NAME, AT, XT)                                          \
  hcblasStatus NAME(hipStream_t accl_view, const int N, const AT &alpha,      \
                    XT *X, const int incX, const int64_t xOffset) {           \
    if (X == nullptr || N <= 0 || incX <= 0) return HCBLAS_INVALID;           \
    return done(on(accl_view) &&                                              \
                ::NAME(handle, N, &alpha, X, incX, xOffset) == 0);            \
  }                                                                           \
  hcblasStatus NAME(hipStream_t accl_view, const int N, const AT &alpha,      \
                    XT *X, const int incX, const int64_t xOffset,             \
                    const int64_t X_batchOffset, const int batchSize) {       \
    if (X == nullptr || N <= 0 || incX <= 0) return HCBLAS_INVALID;           \
    return done(on(accl_view) &&                                              \
                ::NAME##_batched(handle, N, &alpha, X, incX, xOffset,         \
                                 X_batchOffset, batchSize) == 0);             \
  }
  HCBLASLIB_SCAL(hcblas_sscal, float, float)
  HCBLASLIB_SCAL(hcblas_dscal, double, double)
  HCBLASLIB_SCAL(hcblas_cscal, hcComplex, hcComplex)
  HCBLASLIB_SCAL(hcblas_zscal, hcDoubleComplex, hcDoubleComplex)
  HCBLASLIB_SCAL(hcblas_csscal, float, hcComplex)
  HCBLASLIB_SCAL(hcblas_zdscal, double, hcDoubleComplex)
#undef HCBLASLIB_SCAL

  /* ---------------------------------------------------------- COPY */
#define HCBLASLIB_COPY(NAME, T)                                               \
  hcblasStatus NAME(hipStream_t accl_view, const int N, const T *X,           \
                    const int incX, const int64_t xOffset, T *Y,              \
                    const int incY, const int64_t yOffset) {                  \
    if (X == nullptr || Y == nullptr || N <= 0 || incX <= 0 || incY <= 0)     \
      return HCBLAS_INVALID;                                                  \
    return done(on(accl_view) && ::NAME(handle, N, X, incX, xOffset, Y, incY, \
                                        yOffset) == 0);                       \
  }                                                                           \
  hcblasStatus NAME(hipStream_t accl_view, const int N, const T *X,           \
                    const int incX, const int64_t xOffset, T *Y,              \
                    const int incY, const int64_t yOffset,                    \
                    const int64_t X_batchOffset, const int64_t Y_batchOffset, \
                    const int batchSize) {                                    \
    if (X == nullptr || Y == nullptr || N <= 0 || incX <= 0 || incY <= 0)     \
      return HCBLAS_INVALID;                                                  \
    return done(on(accl_view) &&                                              \
                ::NAME##_batched(handle, N, X, incX, xOffset, Y, incY,        \
                                 yOffset, X_batchOffset, Y_batchOffset,       \
                                 batchSize) == 0);                            \
  }
  HCBLASLIB_COPY(hcblas_scopy, float)
  HCBLASLIB_COPY(hcblas_dcopy, double)
#undef HCBLASLIB_COPY

  /* ----------------------------------------------------------- DOT */
#define HCBLASLIB_DOT(NAME, T)                                                \
  hcblasStatus NAME(hipStream_t accl_view, const int N, const T *X,           \
                    const int incX, const int64_t xOffset, const T *Y,        \
                    const int incY, const int64_t yOffset, T &dot) {          \
    if (X == nullptr || Y == nullptr || N <= 0 || incX <= 0 || incY <= 0)     \
      return HCBLAS_INVALID;                                                  \
    return done(on(accl_view) && ::NAME(handle, N, X, incX, xOffset, Y, incY, \
                                        yOffset, &dot) == 0);                 \
  }                                                                           \
  hcblasStatus NAME(hipStream_t accl_view, const int N, const T *X,           \
                    const int incX, const int64_t xOffset, const T *Y,        \
                    const int incY, const int64_t yOffset, T &dot,            \
                    const int64_t X_batchOffset, const int64_t Y_batchOffset, \
                    const int batchSize) {                                    \
    if (X == nullptr || Y == nullptr || N <= 0 || incX <= 0 || incY <= 0)     \
      return HCBLAS_INVALID;                                                  \
    return done(on(accl_view) &&                                              \
                ::NAME##_batched(handle, N, X, incX, xOffset, Y, incY,        \
                                 yOffset, &dot, X_batchOffset, Y_batchOffset, \
                                 batchSize) == 0);                            \
  }
  HCBLASLIB_DOT(hcblas_sdot, float)
  HCBLASLIB_DOT(hcblas_ddot, double)
#undef HCBLASLIB_DOT

  /* ---------------------------------------------------------- ASUM */
#define HCBLASLIB_ASUM(NAME, T)                                               \
  hcblasStatus NAME(hipStream_t accl_view, const int N, T *X, const int incX, \
                    const int64_t xOffset, T *Y) {                            \
    if (X == nullptr || N <= 0 || incX <= 0) return HCBLAS_INVALID;           \
    return done(on(accl_view) &&                                              \
                ::NAME(handle, N, X, incX, xOffset, Y) == 0);                 \
  }                                                                           \
  hcblasStatus NAME(hipStream_t accl_view, const int N, T *X, const int incX, \
                    const int64_t xOffset, T *Y, const int64_t X_batchOffset, \
                    const int batchSize) {                                    \
    if (X == nullptr || N <= 0 || incX <= 0) return HCBLAS_INVALID;           \
    return done(on(accl_view) &&                                              \
                ::NAME##_batched(handle, N, X, incX, xOffset, Y,              \
                                 X_batchOffset, batchSize) == 0);             \
  }
  HCBLASLIB_ASUM(hcblas_sasum, float)
  HCBLASLIB_ASUM(hcblas_dasum, double)
#undef HCBLASLIB_ASUM

 private:
  /* Runs the next call on `s` (the reference's per-call accelerator view). */
  bool on(hipStream_t s) {
    if (!initialized) return false;
    currentAcclView = s;
    return hcblasSetStream(handle, s) == HCBLAS_STATUS_SUCCESS;
  }
  static hcblasStatus done(bool ok) { return ok ? HCBLAS_SUCCEEDS : HCBLAS_INVALID; }
  static hcblasOrder_t ord(hcblasOrder o) {
    return o == RowMajor ? HCBLAS_ROW_MAJOR : HCBLAS_COL_MAJOR;
  }
  /* 'n' / 't' are accepted by the C ABI as they are. */
  static hcblasOperation_t op(hcblasTranspose t) {
    return static_cast<hcblasOperation_t>(static_cast<int>(t));
  }
  template <typename P, typename S>
  hcblasStatus batched(hipStream_t accl_view, hcblasDatatype_t dt,
                       hcblasOrder order, hcblasTranspose typeA,
                       hcblasTranspose typeB, int M, int N, int K,
                       const S *alpha, P *A[], int64_t lda, int64_t aOff,
                       P *B[], int64_t ldb, int64_t bOff, const S *beta,
                       P *C[], int64_t ldc, int64_t cOff, int batchSize) {
    return done(on(accl_view) &&
                ::hcblas_gemm_batched(
                    handle, dt, ord(order), op(typeA), op(typeB), M, N, K,
                    alpha, reinterpret_cast<const void *const *>(A), lda, aOff,
                    reinterpret_cast<const void *const *>(B), ldb, bOff, beta,
                    reinterpret_cast<void *const *>(C), ldc, cOff,
                    batchSize) == 0);
  }
};

#endif  // HCBLAS_MI355X_HCBLASLIB_HPP_
