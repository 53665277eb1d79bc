/*
 * oracle/refblas.c -- CPU restatement of the BLAS arithmetic on the hcBLAS
 * hot path. TEST INFRASTRUCTURE ONLY: it is the checker for the gfx950
 * library (tests, the smoke check and the benchmark's CPU baseline), never
 * part of the product path.
 *
 * The reference (rocmarchive/hcBLAS) checks every routine against host CBLAS
 * (test/unit/CMakeLists.txt:27 links -lblas; e.g. test/unit/hcblas_sgemm.cpp,
 * test/unit-api/test-blas.cpp). This file restates CBLAS's semantics:
 *   - loop orders follow netlib reference BLAS (the -lblas of the reference's
 *     install docs): sgemm/sgemv/sger/saxpy/sdot/sasum/sscal/scopy and the
 *     d/c/z variants;
 *   - row-major calls are mapped to column-major by the CBLAS argument swap;
 *   - contraction: compiled with -ffp-contract=off; the two routines whose
 *     results must be bit-exact use an explicit fma in the form the bundled
 *     OpenBLAS produces (verified bitwise in this container):
 *       axpy  y = fma(alpha, x, y)
 *       ger   a = fma(x_i, alpha*y_j, a)   (netlib temp = alpha*y(j))
 *   - scal multiplies (alpha == 0 keeps NaN, as netlib and OpenBLAS do).
 * Plus accuracy references: FP64 evaluation of FP32 problems (products of
 * floats are exact in double), and double-double (TwoProd via fma + TwoSum)
 * on sampled entries for FP64 problems.
 *
 * Reference file:line followed per routine is given at each function.
 * Parity pinning (golden vectors, closed forms, exact rationals) lives in
 * tests/test_oracle.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ROW_MAJOR 0 /* hcblaslib.h:47 RowMajor = 0 */
#define COL_MAJOR 1
#define OP_N 0 /* hcblas.h:74-78 */
#define OP_T 1
#define OP_C 2

/* ------------------------------------------------------------------------ */
/* Counter-based generator shared by every test: value i of stream (seed)  */
/* depends only on (seed, i), so any slice can be regenerated.             */
/* ------------------------------------------------------------------------ */
static uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static uint64_t rng_at(uint64_t seed, uint64_t i) {
  return splitmix64(splitmix64(seed) ^ (i * 0xD1B54A32D192ED03ull));
}

/* Uniform[-1,1) with a 24-bit grid: exactly representable in fp32. */
void orc_fill_uniform_f32(float *out, int64_t n, uint64_t seed,
                          uint64_t start) {
  for (int64_t i = 0; i < n; ++i) {
    int64_t u = (int64_t)(rng_at(seed, start + (uint64_t)i) >> 40);
    out[i] = (float)((double)(u - (1 << 23)) / (double)(1 << 23));
  }
}
/* Uniform[-1,1) with a 53-bit grid: exactly representable in fp64. */
void orc_fill_uniform_f64(double *out, int64_t n, uint64_t seed,
                          uint64_t start) {
  for (int64_t i = 0; i < n; ++i) {
    int64_t u = (int64_t)(rng_at(seed, start + (uint64_t)i) >> 11);
    out[i] = (double)(u - (1ll << 52)) / (double)(1ll << 52);
  }
}
/* sign * 2^u with u ~ U[-emax, emax] continuous (wide exponent spread). */
void orc_fill_wide_f64(double *out, int64_t n, uint64_t seed, uint64_t start,
                       double emax) {
  for (int64_t i = 0; i < n; ++i) {
    uint64_t r = rng_at(seed, start + (uint64_t)i);
    double u = (double)(r >> 11) / (double)(1ull << 53); /* [0,1) */
    double e = -emax + 2.0 * emax * u;
    double s = (r & 1) ? -1.0 : 1.0;
    out[i] = s * exp2(e);
  }
}
void orc_fill_wide_f32(float *out, int64_t n, uint64_t seed, uint64_t start,
                       double emax) {
  for (int64_t i = 0; i < n; ++i) {
    double d;
    orc_fill_wide_f64(&d, 1, seed, start + (uint64_t)i, emax);
    out[i] = (float)d;
  }
}

static int64_t bidx(int64_t i, int64_t n, int64_t inc) {
  return inc > 0 ? i * inc : (n - 1 - i) * (-inc);
}

/* ------------------------------------------------------------------------ */
/* Level 1 (netlib saxpy.f / sscal.f / scopy.f / sdot.f / sasum.f)          */
/* reference kernels: saxpy_array.cpp:29-97, sscal_array.cpp:29-47,        */
/* scopy_array.cpp:29-45, sdot_array.cpp:30-159, sasum_array.cpp:30-143     */
/* ------------------------------------------------------------------------ */
void orc_saxpy(int64_t n, float alpha, const float *x, int64_t incx, float *y,
               int64_t incy) {
  if (n <= 0 || alpha == 0.0f) return; /* netlib quick return */
  for (int64_t i = 0; i < n; ++i) {
    float *py = &y[bidx(i, n, incy)];
    *py = fmaf(alpha, x[bidx(i, n, incx)], *py);
  }
}
void orc_daxpy(int64_t n, double alpha, const double *x, int64_t incx,
               double *y, int64_t incy) {
  if (n <= 0 || alpha == 0.0) return;
  for (int64_t i = 0; i < n; ++i) {
    double *py = &y[bidx(i, n, incy)];
    *py = fma(alpha, x[bidx(i, n, incx)], *py);
  }
}
void orc_sscal(int64_t n, float alpha, float *x, int64_t incx) {
  if (n <= 0 || incx <= 0) return;
  for (int64_t i = 0; i < n; ++i) x[i * incx] = alpha * x[i * incx];
}
void orc_dscal(int64_t n, double alpha, double *x, int64_t incx) {
  if (n <= 0 || incx <= 0) return;
  for (int64_t i = 0; i < n; ++i) x[i * incx] = alpha * x[i * incx];
}
/* Complex scal: separately rounded products (cscal_array.cpp:28-52 computes
 * the imaginary part wrongly; this follows the definition). */
void orc_cscal(int64_t n, float ar, float ai, float *x, int64_t incx) {
  if (n <= 0 || incx <= 0) return;
  for (int64_t i = 0; i < n; ++i) {
    float *p = &x[2 * i * incx];
    float xr = p[0], xi = p[1];
    float a = ar * xr, b = ai * xi, c = ar * xi, d = ai * xr;
    p[0] = a - b;
    p[1] = c + d;
  }
}
void orc_zscal(int64_t n, double ar, double ai, double *x, int64_t incx) {
  if (n <= 0 || incx <= 0) return;
  for (int64_t i = 0; i < n; ++i) {
    double *p = &x[2 * i * incx];
    double xr = p[0], xi = p[1];
    double a = ar * xr, b = ai * xi, c = ar * xi, d = ai * xr;
    p[0] = a - b;
    p[1] = c + d;
  }
}
void orc_csscal(int64_t n, float a, float *x, int64_t incx) {
  if (n <= 0 || incx <= 0) return;
  for (int64_t i = 0; i < n; ++i) {
    x[2 * i * incx] = a * x[2 * i * incx];
    x[2 * i * incx + 1] = a * x[2 * i * incx + 1];
  }
}
void orc_zdscal(int64_t n, double a, double *x, int64_t incx) {
  if (n <= 0 || incx <= 0) return;
  for (int64_t i = 0; i < n; ++i) {
    x[2 * i * incx] = a * x[2 * i * incx];
    x[2 * i * incx + 1] = a * x[2 * i * incx + 1];
  }
}
void orc_scopy(int64_t n, const float *x, int64_t incx, float *y,
               int64_t incy) {
  for (int64_t i = 0; i < n; ++i) y[bidx(i, n, incy)] = x[bidx(i, n, incx)];
}
void orc_dcopy(int64_t n, const double *x, int64_t incx, double *y,
               int64_t incy) {
  for (int64_t i = 0; i < n; ++i) y[bidx(i, n, incy)] = x[bidx(i, n, incx)];
}
/* netlib sdot: sequential fp32 sum of rounded products. */
float orc_sdot(int64_t n, const float *x, int64_t incx, const float *y,
               int64_t incy) {
  float s = 0.0f;
  for (int64_t i = 0; i < n; ++i) {
    float p = x[bidx(i, n, incx)] * y[bidx(i, n, incy)];
    s = s + p;
  }
  return s;
}
double orc_ddot(int64_t n, const double *x, int64_t incx, const double *y,
                int64_t incy) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    double p = x[bidx(i, n, incx)] * y[bidx(i, n, incy)];
    s = s + p;
  }
  return s;
}
float orc_sasum(int64_t n, const float *x, int64_t incx) {
  float s = 0.0f;
  if (n <= 0 || incx <= 0) return 0.0f;
  for (int64_t i = 0; i < n; ++i) s = s + fabsf(x[i * incx]);
  return s;
}
double orc_dasum(int64_t n, const double *x, int64_t incx) {
  double s = 0.0;
  if (n <= 0 || incx <= 0) return 0.0;
  for (int64_t i = 0; i < n; ++i) s = s + fabs(x[i * incx]);
  return s;
}

/* Double-double helpers (Dekker/Knuth): TwoSum and TwoProd via fma. */
typedef struct {
  double hi, lo;
} dd_t;
static dd_t dd_add_d(dd_t a, double b) {
  double s = a.hi + b;
  double bb = s - a.hi;
  double err = (a.hi - (s - bb)) + (b - bb);
  double lo = a.lo + err;
  double hi = s + lo;
  dd_t r = {hi, lo - (hi - s)};
  return r;
}
static dd_t dd_add(dd_t a, dd_t b) {
  dd_t r = dd_add_d(a, b.hi);
  return dd_add_d(r, b.lo);
}
static dd_t two_prod(double a, double b) {
  double p = a * b;
  dd_t r = {p, fma(a, b, -p)};
  return r;
}
static dd_t dd_mul_d(dd_t a, double b) {
  dd_t p = two_prod(a.hi, b);
  p.lo += a.lo * b;
  double hi = p.hi + p.lo;
  dd_t r = {hi, p.lo - (hi - p.hi)};
  return r;
}

/* Accuracy references for dot/asum: FP64 accumulation of exact fp32
 * products; double-double for fp64 inputs. */
double orc_sdot_f64(int64_t n, const float *x, int64_t incx, const float *y,
                    int64_t incy) {
  dd_t s = {0, 0};
  for (int64_t i = 0; i < n; ++i)
    s = dd_add_d(s, (double)x[bidx(i, n, incx)] * (double)y[bidx(i, n, incy)]);
  return s.hi + s.lo;
}
void orc_ddot_dd(int64_t n, const double *x, int64_t incx, const double *y,
                 int64_t incy, double *hi, double *lo) {
  dd_t s = {0, 0};
  for (int64_t i = 0; i < n; ++i)
    s = dd_add(s, two_prod(x[bidx(i, n, incx)], y[bidx(i, n, incy)]));
  *hi = s.hi;
  *lo = s.lo;
}
void orc_dasum_dd(int64_t n, const double *x, int64_t incx, double *hi,
                  double *lo) {
  dd_t s = {0, 0};
  if (incx <= 0) n = 0; /* netlib dasum.f: non-positive increment -> 0 */
  for (int64_t i = 0; i < n; ++i) s = dd_add_d(s, fabs(x[i * incx]));
  *hi = s.hi;
  *lo = s.lo;
}

/* ------------------------------------------------------------------------ */
/* Level 2 (netlib sgemv.f, sger.f); reference sgemv_array.cpp:826-880,     */
/* sger_array.cpp:128-186.                                                  */
/* ------------------------------------------------------------------------ */
#define GEMV_BODY(T)                                                          \
  int64_t m = M, n = N;                                                       \
  int tr = trans != OP_N;                                                     \
  if (order == ROW_MAJOR) { /* CBLAS: row-major A = col-major A^T */          \
    m = N;                                                                    \
    n = M;                                                                    \
    tr = !tr;                                                                 \
  }                                                                           \
  int64_t leny = tr ? n : m, lenx = tr ? m : n;                               \
  if (M <= 0 || N <= 0) {                                                     \
    if (leny <= 0) return;                                                    \
  }                                                                           \
  if (beta != (T)1) {                                                         \
    for (int64_t i = 0; i < leny; ++i) {                                      \
      T *p = &y[bidx(i, leny, incy)];                                         \
      *p = (beta == (T)0) ? (T)0 : beta * *p;                                 \
    }                                                                         \
  }                                                                           \
  if (alpha == (T)0 || lenx == 0) return;                                     \
  if (!tr) {                                                                  \
    for (int64_t j = 0; j < n; ++j) {                                         \
      T temp = alpha * x[bidx(j, lenx, incx)];                                \
      for (int64_t i = 0; i < m; ++i) {                                       \
        T prod = temp * A[i + j * lda];                                       \
        T *p = &y[bidx(i, leny, incy)];                                       \
        *p = *p + prod;                                                       \
      }                                                                       \
    }                                                                         \
  } else {                                                                    \
    for (int64_t j = 0; j < n; ++j) {                                         \
      T temp = 0;                                                             \
      for (int64_t i = 0; i < m; ++i) {                                       \
        T prod = A[i + j * lda] * x[bidx(i, lenx, incx)];                     \
        temp = temp + prod;                                                   \
      }                                                                       \
      T *p = &y[bidx(j, leny, incy)];                                         \
      T at = alpha * temp;                                                    \
      *p = *p + at;                                                           \
    }                                                                         \
  }

void orc_sgemv(int order, int trans, int64_t M, int64_t N, float alpha,
               const float *A, int64_t lda, const float *x, int64_t incx,
               float beta, float *y, int64_t incy) {
  GEMV_BODY(float)
}
void orc_dgemv(int order, int trans, int64_t M, int64_t N, double alpha,
               const double *A, int64_t lda, const double *x, int64_t incx,
               double beta, double *y, int64_t incy) {
  GEMV_BODY(double)
}

/* FP64 / double-double accuracy references for gemv: y_out (double) =
 * alpha*op(A)*x + beta*y computed in double-double and rounded once. */
#define GEMV_DD_BODY(T)                                                        \
  int64_t m = M, n = N;                                                        \
  int tr = trans != OP_N;                                                      \
  if (order == ROW_MAJOR) {                                                    \
    m = N;                                                                     \
    n = M;                                                                     \
    tr = !tr;                                                                  \
  }                                                                            \
  int64_t leny = tr ? n : m, lenx = tr ? m : n;                                \
  for (int64_t o = 0; o < leny; ++o) {                                         \
    dd_t s = {0, 0};                                                           \
    for (int64_t r = 0; r < lenx; ++r) {                                       \
      double a = tr ? (double)A[r + o * lda] : (double)A[o + r * lda];         \
      s = dd_add(s, two_prod(a, (double)x[bidx(r, lenx, incx)]));              \
    }                                                                          \
    s = dd_mul_d(s, (double)alpha);                                            \
    if (beta != (T)0)                                                          \
      s = dd_add(s, two_prod((double)beta, (double)y[bidx(o, leny, incy)]));   \
    out[o] = s.hi + s.lo;                                                      \
  }

void orc_sgemv_ref(int order, int trans, int64_t M, int64_t N, float alpha,
                   const float *A, int64_t lda, const float *x, int64_t incx,
                   float beta, const float *y, int64_t incy, double *out) {
  GEMV_DD_BODY(float)
}
void orc_dgemv_ref(int order, int trans, int64_t M, int64_t N, double alpha,
                   const double *A, int64_t lda, const double *x, int64_t incx,
                   double beta, const double *y, int64_t incy, double *out) {
  GEMV_DD_BODY(double)
}

#define GER_BODY(T, FMA)                                                      \
  const T *v = x, *w = y;                                                     \
  int64_t m = M, n = N, incv = incx, incw = incy;                             \
  if (order == ROW_MAJOR) { /* CBLAS: swap (m,n) and (x,y) */                 \
    m = N;                                                                    \
    n = M;                                                                    \
    v = y;                                                                    \
    w = x;                                                                    \
    incv = incy;                                                              \
    incw = incx;                                                              \
  }                                                                           \
  if (m <= 0 || n <= 0 || alpha == (T)0) return;                              \
  for (int64_t j = 0; j < n; ++j) {                                           \
    T temp = alpha * w[bidx(j, n, incw)];                                     \
    for (int64_t i = 0; i < m; ++i) {                                         \
      T *p = &A[i + j * lda];                                                 \
      *p = FMA(v[bidx(i, m, incv)], temp, *p);                                \
    }                                                                         \
  }

void orc_sger(int order, int64_t M, int64_t N, float alpha, const float *x,
              int64_t incx, const float *y, int64_t incy, float *A,
              int64_t lda) {
  GER_BODY(float, fmaf)
}
void orc_dger(int order, int64_t M, int64_t N, double alpha, const double *x,
              int64_t incx, const double *y, int64_t incy, double *A,
              int64_t lda) {
  GER_BODY(double, fma)
}

/* ------------------------------------------------------------------------ */
/* Level 3 (netlib sgemm.f / cgemm.f); reference sgemm_array_wrapper.cpp:    */
/* 27-364, cgemm_array_wrapper.cpp:212-371.                                  */
/* Row-major is the CBLAS swap: C^T = op(B)^T op(A)^T.                      */
/* ------------------------------------------------------------------------ */
#define SWAP_ROW_MAJOR(PT)                                              \
  if (order == ROW_MAJOR) {                                             \
    int t = ta;                                                         \
    ta = tb;                                                            \
    tb = t;                                                             \
    int64_t tmn = M;                                                    \
    M = N;                                                              \
    N = tmn;                                                            \
    const PT *tp = A;                                                   \
    A = B;                                                              \
    B = tp;                                                             \
    int64_t tl = lda;                                                   \
    lda = ldb;                                                          \
    ldb = tl;                                                           \
  }

#define GEMM_BODY(T)                                                          \
  SWAP_ROW_MAJOR(T)                                                           \
  if (M == 0 || N == 0) return;                                               \
  int nota = ta == OP_N, notb = tb == OP_N;                                   \
  for (int64_t j = 0; j < N; ++j) {                                           \
    if (nota) {                                                               \
      if (beta == (T)0) {                                                     \
        for (int64_t i = 0; i < M; ++i) C[i + j * ldc] = 0;                   \
      } else if (beta != (T)1) {                                              \
        for (int64_t i = 0; i < M; ++i) C[i + j * ldc] *= beta;               \
      }                                                                       \
      if (alpha == (T)0) continue;                                            \
      for (int64_t l = 0; l < K; ++l) {                                       \
        T bl = notb ? B[l + j * ldb] : B[j + l * ldb];                        \
        T temp = alpha * bl;                                                  \
        for (int64_t i = 0; i < M; ++i) {                                     \
          T prod = temp * A[i + l * lda];                                     \
          C[i + j * ldc] = C[i + j * ldc] + prod;                             \
        }                                                                     \
      }                                                                       \
    } else {                                                                  \
      for (int64_t i = 0; i < M; ++i) {                                       \
        T temp = 0;                                                           \
        if (alpha != (T)0)                                                    \
          for (int64_t l = 0; l < K; ++l) {                                   \
            T bl = notb ? B[l + j * ldb] : B[j + l * ldb];                    \
            T prod = A[l + i * lda] * bl;                                     \
            temp = temp + prod;                                               \
          }                                                                   \
        T r = alpha * temp;                                                   \
        if (beta != (T)0) {                                                   \
          T bc = beta * C[i + j * ldc];                                       \
          r = r + bc;                                                         \
        }                                                                     \
        C[i + j * ldc] = r;                                                   \
      }                                                                       \
    }                                                                         \
  }

void orc_sgemm(int order, int ta, int tb, int64_t M, int64_t N, int64_t K,
               float alpha, const float *A, int64_t lda, const float *B,
               int64_t ldb, float beta, float *C, int64_t ldc) {
  GEMM_BODY(float)
}
void orc_dgemm(int order, int ta, int tb, int64_t M, int64_t N, int64_t K,
               double alpha, const double *A, int64_t lda, const double *B,
               int64_t ldb, double beta, double *C, int64_t ldc) {
  GEMM_BODY(double)
}

/* FP64 reference of an fp32 GEMM: every product is exact in double; sums
 * carry ~K*2^-53 relative error, far below fp32 resolution. Output is a
 * column-major M x N double array (ld = M) in the caller's order:
 * out[i + j*M] for col-major, out[i*N + j] for row-major. */
void orc_sgemm_ref64(int order, int ta, int tb, int64_t M, int64_t N,
                     int64_t K, float alpha, const float *A, int64_t lda,
                     const float *B, int64_t ldb, float beta, const float *C,
                     int64_t ldc, double *out) {
  int64_t M0 = M, N0 = N;
  SWAP_ROW_MAJOR(float)
  (void)M0;
  (void)N0;
  int nota = ta == OP_N, notb = tb == OP_N;
  for (int64_t j = 0; j < N; ++j)
    for (int64_t i = 0; i < M; ++i) {
      double s = 0.0;
      for (int64_t l = 0; l < K; ++l) {
        double a = nota ? A[i + l * lda] : A[l + i * lda];
        double b = notb ? B[l + j * ldb] : B[j + l * ldb];
        s += a * b;
      }
      double r = (double)alpha * s;
      if (beta != 0.0f) r += (double)beta * (double)C[i + j * ldc];
      out[i + j * M] = r; /* = row-major out[j*M + i] when swapped */
    }
}

/* Double-double reference of sampled entries of an fp64 GEMM (column-major
 * entry (i, j) of the caller's C for col-major; row-major entry (i, j) for
 * row-major). */
void orc_dgemm_dd_sample(int order, int ta, int tb, int64_t M, int64_t N,
                         int64_t K, double alpha, const double *A, int64_t lda,
                         const double *B, int64_t ldb, double beta,
                         const double *C, int64_t ldc, int64_t nsamp,
                         const int64_t *ri, const int64_t *cj, double *out) {
  SWAP_ROW_MAJOR(double)
  int nota = ta == OP_N, notb = tb == OP_N;
  for (int64_t s = 0; s < nsamp; ++s) {
    /* In swapped (column-major C^T) terms, caller's (i, j) is (j, i). */
    int64_t i = order == ROW_MAJOR ? cj[s] : ri[s];
    int64_t j = order == ROW_MAJOR ? ri[s] : cj[s];
    dd_t acc = {0, 0};
    for (int64_t l = 0; l < K; ++l) {
      double a = nota ? A[i + l * lda] : A[l + i * lda];
      double b = notb ? B[l + j * ldb] : B[j + l * ldb];
      acc = dd_add(acc, two_prod(a, b));
    }
    acc = dd_mul_d(acc, alpha);
    if (beta != 0.0) acc = dd_add(acc, two_prod(beta, C[i + j * ldc]));
    out[s] = acc.hi + acc.lo;
  }
}

/* Complex GEMM (netlib cgemm.f semantics incl. conjugate transpose). Data are
 * interleaved (re, im); lda etc. count complex elements. */
#define CGEMM_BODY(T)                                                          \
  SWAP_ROW_MAJOR(T)                                                            \
  if (M == 0 || N == 0) return;                                                \
  int alpha0 = alpha_r == 0 && alpha_i == 0;                                   \
  for (int64_t j = 0; j < N; ++j)                                              \
    for (int64_t i = 0; i < M; ++i) {                                          \
      T sr = 0, si = 0;                                                        \
      if (!alpha0)                                                             \
        for (int64_t l = 0; l < K; ++l) {                                      \
          const T *pa = ta == OP_N ? &A[2 * (i + l * lda)]                     \
                                   : &A[2 * (l + i * lda)];                    \
          const T *pb = tb == OP_N ? &B[2 * (l + j * ldb)]                     \
                                   : &B[2 * (j + l * ldb)];                    \
          T ar = pa[0], ai = ta == OP_C ? -pa[1] : pa[1];                      \
          T br = pb[0], bi = tb == OP_C ? -pb[1] : pb[1];                      \
          T p1 = ar * br, p2 = ai * bi, p3 = ar * bi, p4 = ai * br;            \
          T pr = p1 - p2, pi = p3 + p4;                                        \
          sr = sr + pr;                                                        \
          si = si + pi;                                                        \
        }                                                                      \
      T *pc = &C[2 * (i + j * ldc)];                                           \
      T q1 = alpha_r * sr, q2 = alpha_i * si, q3 = alpha_r * si,               \
        q4 = alpha_i * sr;                                                     \
      T rr = q1 - q2, ri2 = q3 + q4;                                           \
      if (!(beta_r == 0 && beta_i == 0)) {                                     \
        T cr = pc[0], ci = pc[1];                                              \
        T c1 = beta_r * cr, c2 = beta_i * ci, c3 = beta_r * ci,                \
          c4 = beta_i * cr;                                                    \
        rr = rr + (c1 - c2);                                                   \
        ri2 = ri2 + (c3 + c4);                                                 \
      }                                                                        \
      pc[0] = rr;                                                              \
      pc[1] = ri2;                                                             \
    }

void orc_cgemm(int order, int ta, int tb, int64_t M, int64_t N, int64_t K,
               float alpha_r, float alpha_i, const float *A, int64_t lda,
               const float *B, int64_t ldb, float beta_r, float beta_i,
               float *C, int64_t ldc) {
  CGEMM_BODY(float)
}
void orc_zgemm(int order, int ta, int tb, int64_t M, int64_t N, int64_t K,
               double alpha_r, double alpha_i, const double *A, int64_t lda,
               const double *B, int64_t ldb, double beta_r, double beta_i,
               double *C, int64_t ldc) {
  CGEMM_BODY(double)
}

/* Double-double reference of sampled entries of a complex GEMM (fp32 or fp64
 * inputs promoted to double). out has 2*nsamp doubles (re, im). */
#define ZDD_BODY(T)                                                           \
  SWAP_ROW_MAJOR(T)                                                           \
  for (int64_t s = 0; s < nsamp; ++s) {                                       \
    int64_t i = order == ROW_MAJOR ? cj[s] : ri[s];                           \
    int64_t j = order == ROW_MAJOR ? ri[s] : cj[s];                           \
    dd_t sr = {0, 0}, si = {0, 0};                                            \
    for (int64_t l = 0; l < K; ++l) {                                         \
      const T *pa =                                                           \
          ta == OP_N ? &A[2 * (i + l * lda)] : &A[2 * (l + i * lda)];         \
      const T *pb =                                                           \
          tb == OP_N ? &B[2 * (l + j * ldb)] : &B[2 * (j + l * ldb)];         \
      double ar = pa[0], ai = ta == OP_C ? -(double)pa[1] : (double)pa[1];    \
      double br = pb[0], bi = tb == OP_C ? -(double)pb[1] : (double)pb[1];    \
      sr = dd_add(sr, two_prod(ar, br));                                      \
      sr = dd_add(sr, two_prod(-ai, bi));                                     \
      si = dd_add(si, two_prod(ar, bi));                                      \
      si = dd_add(si, two_prod(ai, br));                                      \
    }                                                                         \
    dd_t rr = dd_add(dd_mul_d(sr, alpha_r), dd_mul_d(si, -alpha_i));          \
    dd_t rim = dd_add(dd_mul_d(si, alpha_r), dd_mul_d(sr, alpha_i));          \
    if (!(beta_r == 0 && beta_i == 0)) {                                      \
      const T *pc = &C[2 * (i + j * ldc)];                                    \
      rr = dd_add(rr, two_prod(beta_r, pc[0]));                               \
      rr = dd_add(rr, two_prod(-beta_i, pc[1]));                              \
      rim = dd_add(rim, two_prod(beta_r, pc[1]));                             \
      rim = dd_add(rim, two_prod(beta_i, pc[0]));                             \
    }                                                                         \
    out[2 * s] = rr.hi + rr.lo;                                               \
    out[2 * s + 1] = rim.hi + rim.lo;                                         \
  }

void orc_zgemm_dd_sample(int order, int ta, int tb, int64_t M, int64_t N,
                         int64_t K, double alpha_r, double alpha_i,
                         const double *A, int64_t lda, const double *B,
                         int64_t ldb, double beta_r, double beta_i,
                         const double *C, int64_t ldc, int64_t nsamp,
                         const int64_t *ri, const int64_t *cj, double *out) {
  ZDD_BODY(double)
}
void orc_cgemm_dd_sample(int order, int ta, int tb, int64_t M, int64_t N,
                         int64_t K, double alpha_r, double alpha_i,
                         const float *A, int64_t lda, const float *B,
                         int64_t ldb, double beta_r, double beta_i,
                         const float *C, int64_t ldc, int64_t nsamp,
                         const int64_t *ri, const int64_t *cj, double *out) {
  ZDD_BODY(float)
}
/* Same for real fp32 GEMM (sampled), used at sizes where the full FP64
 * reference is too slow. */
void orc_sgemm_dd_sample(int order, int ta, int tb, int64_t M, int64_t N,
                         int64_t K, double alpha, const float *A, int64_t lda,
                         const float *B, int64_t ldb, double beta,
                         const float *C, int64_t ldc, int64_t nsamp,
                         const int64_t *ri, const int64_t *cj, double *out) {
  SWAP_ROW_MAJOR(float)
  int nota = ta == OP_N, notb = tb == OP_N;
  for (int64_t s = 0; s < nsamp; ++s) {
    int64_t i = order == ROW_MAJOR ? cj[s] : ri[s];
    int64_t j = order == ROW_MAJOR ? ri[s] : cj[s];
    dd_t acc = {0, 0};
    for (int64_t l = 0; l < K; ++l) {
      double a = nota ? A[i + l * lda] : A[l + i * lda];
      double b = notb ? B[l + j * ldb] : B[j + l * ldb];
      acc = dd_add_d(acc, a * b); /* exact product of floats */
    }
    acc = dd_mul_d(acc, alpha);
    if (beta != 0.0) acc = dd_add(acc, two_prod(beta, (double)C[i + j * ldc]));
    out[s] = acc.hi + acc.lo;
  }
}
