// The reference's lib-level unit tests (test/unit/hcblas_*.cpp), ported to
// the C++ mirror include/hcblaslib.hpp: the same Hcblaslibrary overloads,
// argument order, rand_r(seed = 100) integer recipes and status checks, with
// the CPU oracle (oracle/lib/liboracle.so, netlib order) in place of CBLAS.
// Integer-valued data makes every result exact, so results are compared with
// ==, as the reference's EXPECT_EQ does. Cases beyond the reference (marked
// "extension") exercise non-zero offsets and batch offsets other than the
// reference's N.
//
// Build: make -C tests/cpp (hipcc host code, links libhcblas.so and
// liboracle.so). Run on a GPU: tests/cpp/bin/test_hcblaslib; exit status 0
// when every check passes. Prints one line per test.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/hcblaslib.hpp"

extern "C" {
void orc_saxpy(int64_t, float, const float *, int64_t, float *, int64_t);
void orc_daxpy(int64_t, double, const double *, int64_t, double *, int64_t);
void orc_sscal(int64_t, float, float *, int64_t);
void orc_dscal(int64_t, double, double *, int64_t);
void orc_cscal(int64_t, float, float, float *, int64_t);
void orc_zscal(int64_t, double, double, double *, int64_t);
void orc_csscal(int64_t, float, float *, int64_t);
void orc_zdscal(int64_t, double, double *, int64_t);
void orc_scopy(int64_t, const float *, int64_t, float *, int64_t);
void orc_dcopy(int64_t, const double *, int64_t, double *, int64_t);
float orc_sdot(int64_t, const float *, int64_t, const float *, int64_t);
double orc_ddot(int64_t, const double *, int64_t, const double *, int64_t);
float orc_sasum(int64_t, const float *, int64_t);
double orc_dasum(int64_t, const double *, int64_t);
void orc_sgemv(int, int, int64_t, int64_t, float, const float *, int64_t,
               const float *, int64_t, float, float *, int64_t);
void orc_dgemv(int, int, int64_t, int64_t, double, const double *, int64_t,
               const double *, int64_t, double, double *, int64_t);
void orc_sger(int, int64_t, int64_t, float, const float *, int64_t,
              const float *, int64_t, float *, int64_t);
void orc_dger(int, int64_t, int64_t, double, const double *, int64_t,
              const double *, int64_t, double *, int64_t);
void orc_sgemm(int, int, int, int64_t, int64_t, int64_t, float, const float *,
               int64_t, const float *, int64_t, float, float *, int64_t);
void orc_dgemm(int, int, int, int64_t, int64_t, int64_t, double,
               const double *, int64_t, const double *, int64_t, double,
               double *, int64_t);
void orc_cgemm(int, int, int, int64_t, int64_t, int64_t, float, float,
               const float *, int64_t, const float *, int64_t, float, float,
               float *, int64_t);
void orc_zgemm(int, int, int, int64_t, int64_t, int64_t, double, double,
               const double *, int64_t, const double *, int64_t, double,
               double, double *, int64_t);
}

namespace {

unsigned int global_seed = 100;
int g_checks = 0, g_fail = 0, g_test_fail = 0;

#define EXPECT_EQ(a, b)                                                     \
  do {                                                                      \
    ++g_checks;                                                             \
    if (!((a) == (b))) {                                                    \
      ++g_fail;                                                             \
      ++g_test_fail;                                                        \
      if (g_test_fail <= 5)                                                 \
        std::printf("  FAIL %s:%d: %s == %s\n", __FILE__, __LINE__, #a, #b); \
    }                                                                       \
  } while (0)

void hip_ok(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    std::printf("HIP error %s in %s\n", hipGetErrorString(e), what);
    std::exit(2);
  }
}

// Device buffer of n elements of T.
template <typename T>
struct Dev {
  T *p = nullptr;
  size_t n = 0;
  explicit Dev(size_t n_) : n(n_) {
    hip_ok(hipMalloc(&p, sizeof(T) * (n ? n : 1)), "hipMalloc");
  }
  Dev(const std::vector<T> &h) : Dev(h.size()) { put(h); }
  ~Dev() { (void)hipFree(p); }
  void put(const std::vector<T> &h) {
    hip_ok(hipMemcpy(p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice),
           "H2D");
  }
  std::vector<T> get() const {
    std::vector<T> h(n);
    hip_ok(hipDeviceSynchronize(), "sync");
    hip_ok(hipMemcpy(h.data(), p, sizeof(T) * n, hipMemcpyDeviceToHost),
           "D2H");
    return h;
  }
};

template <typename T>
std::vector<T> rnd(size_t n, int mod) {
  std::vector<T> v(n);
  for (auto &e : v) e = T(rand_r(&global_seed) % mod);
  return v;
}

template <typename T>
bool same(const std::vector<T> &a, const std::vector<T> &b) {
  return a.size() == b.size() &&
         std::memcmp(a.data(), b.data(), sizeof(T) * a.size()) == 0;
}

void begin(const char *name) {
  g_test_fail = 0;
  std::printf("[ RUN  ] %s\n", name);
}
void end(const char *name) {
  std::printf("[ %s ] %s\n", g_test_fail ? "FAIL" : " OK ", name);
}

// ---------------------------------------------------------------- axpy
template <typename T, typename F, typename FB, typename O>
void axpy_tests(const char *name, F axpy, FB axpy_b, O oracle) {
  begin(name);
  Hcblaslibrary hc;
  hipStream_t accl_view = hc.currentAcclView;
  // return_correct_*_Implementation_type_1 (hcblas_saxpy.cpp:34-101)
  int N = 279, incX = 1, incY = 1;
  int64_t xOffset = 0, yOffset = 0;
  T alpha = 1;
  std::vector<T> X = rnd<T>(N, 10), Y = rnd<T>(N, 15);
  Dev<T> devX(X), devY(Y);
  EXPECT_EQ((hc.*axpy)(accl_view, N, alpha, devX.p, incX, devY.p, incY,
                       xOffset, yOffset),
            HCBLAS_SUCCEEDS);
  EXPECT_EQ((hc.*axpy)(accl_view, N, alpha, nullptr, incX, devY.p, incY,
                       xOffset, yOffset),
            HCBLAS_INVALID);
  EXPECT_EQ((hc.*axpy)(accl_view, N, alpha, devX.p, incX, nullptr, incY,
                       xOffset, yOffset),
            HCBLAS_INVALID);
  EXPECT_EQ((hc.*axpy)(accl_view, N, T(0), devX.p, incX, devY.p, incY,
                       xOffset, yOffset),
            HCBLAS_SUCCEEDS);
  EXPECT_EQ((hc.*axpy)(accl_view, 0, T(0), devX.p, incX, devY.p, incY,
                       xOffset, yOffset),
            HCBLAS_INVALID);
  EXPECT_EQ((hc.*axpy)(accl_view, N, alpha, devX.p, 0, devY.p, incY, xOffset,
                       yOffset),
            HCBLAS_INVALID);
  EXPECT_EQ((hc.*axpy)(accl_view, N, alpha, devX.p, incX, devY.p, 0, xOffset,
                       yOffset),
            HCBLAS_INVALID);
  // func_correct_*_Implementation_type_1 (hcblas_saxpy.cpp:103-146)
  for (T a : {T(1), T(3)}) {
    devY.put(Y);
    EXPECT_EQ((hc.*axpy)(accl_view, N, a, devX.p, incX, devY.p, incY, 0, 0),
              HCBLAS_SUCCEEDS);
    std::vector<T> ref = Y;
    oracle(N, a, X.data(), incX, ref.data(), incY);
    EXPECT_EQ(same(devY.get(), ref), true);
  }
  // extension: offsets and increments
  {
    const int n = 100, ix = 2, iy = 3, xo = 7, yo = 5;
    std::vector<T> x = rnd<T>(xo + n * ix, 10), y = rnd<T>(yo + n * iy, 15);
    Dev<T> dx(x), dy(y);
    EXPECT_EQ((hc.*axpy)(accl_view, n, T(3), dx.p, ix, dy.p, iy, xo, yo),
              HCBLAS_SUCCEEDS);
    std::vector<T> ref = y;
    oracle(n, T(3), x.data() + xo, ix, ref.data() + yo, iy);
    EXPECT_EQ(same(dy.get(), ref), true);
  }
  // func_correct_*_Implementation_type_2 (hcblas_saxpy.cpp:226-280):
  // batchSize 128, X_batchOffset = Y_batchOffset = N
  const int batchSize = 128;
  std::vector<T> Xb = rnd<T>((size_t)N * batchSize, 10);
  std::vector<T> Yb = rnd<T>((size_t)N * batchSize, 15);
  Dev<T> dXb(Xb), dYb(Yb);
  EXPECT_EQ((hc.*axpy_b)(accl_view, N, alpha, dXb.p, incX, N, dYb.p, incY, N,
                         xOffset, yOffset, batchSize),
            HCBLAS_SUCCEEDS);
  std::vector<T> refb = Yb;
  for (int b = 0; b < batchSize; ++b)
    oracle(N, alpha, Xb.data() + (size_t)b * N, incX,
           refb.data() + (size_t)b * N, incY);
  EXPECT_EQ(same(dYb.get(), refb), true);
  EXPECT_EQ((hc.*axpy_b)(accl_view, N, alpha, nullptr, incX, N, dYb.p, incY,
                         N, xOffset, yOffset, batchSize),
            HCBLAS_INVALID);
  // extension: batch offsets other than N, with offsets
  {
    const int n = 61, bs = 9, xbo = 70, ybo = 64, xo = 3, yo = 1;
    std::vector<T> x = rnd<T>(xo + xbo * bs, 10), y = rnd<T>(yo + ybo * bs, 15);
    Dev<T> dx(x), dy(y);
    EXPECT_EQ((hc.*axpy_b)(accl_view, n, T(3), dx.p, 1, xbo, dy.p, 1, ybo, xo,
                           yo, bs),
              HCBLAS_SUCCEEDS);
    std::vector<T> ref = y;
    for (int b = 0; b < bs; ++b)
      oracle(n, T(3), x.data() + xo + b * xbo, 1, ref.data() + yo + b * ybo, 1);
    EXPECT_EQ(same(dy.get(), ref), true);
  }
  end(name);
}

// ---------------------------------------------------------------- scal
template <typename T, typename R, typename A, typename F, typename FB,
          typename O>
void scal_tests(const char *name, A alpha, F scal, FB scal_b, O oracle) {
  begin(name);
  Hcblaslibrary hc;
  hipStream_t accl_view = hc.currentAcclView;
  const int N = 119;
  // integer-valued components of the storage type R (T is R or {R, R})
  const std::vector<R> raw = rnd<R>(N * sizeof(T) / sizeof(R), 10);
  std::vector<T> X(N);
  std::memcpy(X.data(), raw.data(), sizeof(T) * N);
  Dev<T> devX(X);
  // return_correct (hcblas_sscal.cpp): NULL, N <= 0, incX <= 0
  EXPECT_EQ((hc.*scal)(accl_view, N, alpha, devX.p, 1, 0), HCBLAS_SUCCEEDS);
  EXPECT_EQ((hc.*scal)(accl_view, N, alpha, nullptr, 1, 0), HCBLAS_INVALID);
  EXPECT_EQ((hc.*scal)(accl_view, 0, alpha, devX.p, 1, 0), HCBLAS_INVALID);
  EXPECT_EQ((hc.*scal)(accl_view, N, alpha, devX.p, 0, 0), HCBLAS_INVALID);
  // func_correct type 1, with an offset and an increment (extension)
  for (int inc : {1, 2}) {
    const int n = (N - 5) / inc, off = 5;
    devX.put(X);
    EXPECT_EQ((hc.*scal)(accl_view, n, alpha, devX.p, inc, off),
              HCBLAS_SUCCEEDS);
    std::vector<T> ref = X;
    oracle(n, alpha, ref.data() + off, inc);
    EXPECT_EQ(same(devX.get(), ref), true);
  }
  // func_correct type 2: batchSize 32, X_batchOffset = n (reference), and a
  // batch offset larger than n (extension)
  for (int xbo : {40, 47}) {
    const int n = 40, bs = 32, off = 2;
    std::vector<T> xb((size_t)off + (size_t)xbo * bs);
    for (size_t i = 0; i < xb.size(); ++i) xb[i] = X[i % X.size()];
    Dev<T> dxb(xb);
    EXPECT_EQ((hc.*scal_b)(accl_view, n, alpha, dxb.p, 1, off, xbo, bs),
              HCBLAS_SUCCEEDS);
    std::vector<T> ref = xb;
    for (int b = 0; b < bs; ++b) oracle(n, alpha, ref.data() + off + b * xbo, 1);
    EXPECT_EQ(same(dxb.get(), ref), true);
  }
  end(name);
}

// ---------------------------------------------------------------- copy
template <typename T, typename F, typename FB, typename O>
void copy_tests(const char *name, F copy, FB copy_b, O oracle) {
  begin(name);
  Hcblaslibrary hc;
  hipStream_t accl_view = hc.currentAcclView;
  const int N = 279;
  std::vector<T> X = rnd<T>(N, 10), Y = rnd<T>(N, 15);
  Dev<T> devX(X), devY(Y);
  EXPECT_EQ((hc.*copy)(accl_view, N, devX.p, 1, 0, devY.p, 1, 0),
            HCBLAS_SUCCEEDS);
  EXPECT_EQ(same(devY.get(), X), true);
  EXPECT_EQ((hc.*copy)(accl_view, N, nullptr, 1, 0, devY.p, 1, 0),
            HCBLAS_INVALID);
  EXPECT_EQ((hc.*copy)(accl_view, N, devX.p, 1, 0, nullptr, 1, 0),
            HCBLAS_INVALID);
  EXPECT_EQ((hc.*copy)(accl_view, 0, devX.p, 1, 0, devY.p, 1, 0),
            HCBLAS_INVALID);
  EXPECT_EQ((hc.*copy)(accl_view, N, devX.p, 0, 0, devY.p, 1, 0),
            HCBLAS_INVALID);
  // extension: offsets and increments
  {
    const int n = 90, ix = 3, iy = 2, xo = 4, yo = 9;
    std::vector<T> x = rnd<T>(xo + n * ix, 10), y = rnd<T>(yo + n * iy, 15);
    Dev<T> dx(x), dy(y);
    EXPECT_EQ((hc.*copy)(accl_view, n, dx.p, ix, xo, dy.p, iy, yo),
              HCBLAS_SUCCEEDS);
    std::vector<T> ref = y;
    oracle(n, x.data() + xo, ix, ref.data() + yo, iy);
    EXPECT_EQ(same(dy.get(), ref), true);
  }
  // type 2: batchSize 128 with batch offsets N; then other batch offsets
  for (int extra : {0, 5}) {
    const int bs = 128, xbo = N + extra, ybo = N + 2 * extra;
    std::vector<T> xb = rnd<T>((size_t)xbo * bs, 10);
    std::vector<T> yb = rnd<T>((size_t)ybo * bs, 15);
    Dev<T> dxb(xb), dyb(yb);
    EXPECT_EQ((hc.*copy_b)(accl_view, N, dxb.p, 1, 0, dyb.p, 1, 0, xbo, ybo,
                           bs),
              HCBLAS_SUCCEEDS);
    std::vector<T> ref = yb;
    for (int b = 0; b < bs; ++b)
      oracle(N, xb.data() + (size_t)b * xbo, 1, ref.data() + (size_t)b * ybo,
             1);
    EXPECT_EQ(same(dyb.get(), ref), true);
  }
  end(name);
}

// ----------------------------------------------------------------- dot
template <typename T, typename F, typename FB, typename O>
void dot_tests(const char *name, F dot, FB dot_b, O oracle) {
  begin(name);
  Hcblaslibrary hc;
  hipStream_t accl_view = hc.currentAcclView;
  const int N = 279;
  std::vector<T> X = rnd<T>(N, 10), Y = rnd<T>(N, 15);
  Dev<T> devX(X), devY(Y);
  T r = -1;
  // return_correct (hcblas_sdot.cpp:34-78)
  EXPECT_EQ((hc.*dot)(accl_view, N, devX.p, 1, 0, devY.p, 1, 0, r),
            HCBLAS_SUCCEEDS);
  EXPECT_EQ(r, oracle(N, X.data(), 1, Y.data(), 1));
  EXPECT_EQ((hc.*dot)(accl_view, N, nullptr, 1, 0, devY.p, 1, 0, r),
            HCBLAS_INVALID);
  EXPECT_EQ((hc.*dot)(accl_view, N, devX.p, 1, 0, nullptr, 1, 0, r),
            HCBLAS_INVALID);
  EXPECT_EQ((hc.*dot)(accl_view, 0, devX.p, 1, 0, devY.p, 1, 0, r),
            HCBLAS_INVALID);
  EXPECT_EQ((hc.*dot)(accl_view, N, devX.p, 0, 0, devY.p, 1, 0, r),
            HCBLAS_INVALID);
  EXPECT_EQ((hc.*dot)(accl_view, N, devX.p, 1, 0, devY.p, 0, 0, r),
            HCBLAS_INVALID);
  // extension: offsets and increments
  {
    const int n = 77, ix = 2, iy = 3, xo = 11, yo = 6;
    std::vector<T> x = rnd<T>(xo + n * ix, 10), y = rnd<T>(yo + n * iy, 15);
    Dev<T> dx(x), dy(y);
    EXPECT_EQ((hc.*dot)(accl_view, n, dx.p, ix, xo, dy.p, iy, yo, r),
              HCBLAS_SUCCEEDS);
    EXPECT_EQ(r, oracle(n, x.data() + xo, ix, y.data() + yo, iy));
  }
  // type 2 (hcblas_sdot.cpp:267-330): one scalar summed over the batch
  for (int extra : {0, 3}) {
    const int bs = 128, xbo = N + extra, ybo = N + extra;
    std::vector<T> xb = rnd<T>((size_t)xbo * bs, 10);
    std::vector<T> yb = rnd<T>((size_t)ybo * bs, 15);
    Dev<T> dxb(xb), dyb(yb);
    EXPECT_EQ((hc.*dot_b)(accl_view, N, dxb.p, 1, 0, dyb.p, 1, 0, r, xbo, ybo,
                          bs),
              HCBLAS_SUCCEEDS);
    T ref = 0;
    for (int b = 0; b < bs; ++b)
      ref += oracle(N, xb.data() + (size_t)b * xbo, 1,
                    yb.data() + (size_t)b * ybo, 1);
    EXPECT_EQ(r, ref);
  }
  end(name);
}

// ---------------------------------------------------------------- asum
template <typename T, typename F, typename FB, typename O>
void asum_tests(const char *name, F asum, FB asum_b, O oracle) {
  begin(name);
  Hcblaslibrary hc;
  hipStream_t accl_view = hc.currentAcclView;
  const int N = 279;
  std::vector<T> X = rnd<T>(N, 10);
  for (int i = 0; i < N; i += 3) X[i] = -X[i];
  Dev<T> devX(X);
  T r = -1;
  EXPECT_EQ((hc.*asum)(accl_view, N, devX.p, 1, 0, &r), HCBLAS_SUCCEEDS);
  EXPECT_EQ(r, oracle(N, X.data(), 1));
  EXPECT_EQ((hc.*asum)(accl_view, N, nullptr, 1, 0, &r), HCBLAS_INVALID);
  EXPECT_EQ((hc.*asum)(accl_view, 0, devX.p, 1, 0, &r), HCBLAS_INVALID);
  EXPECT_EQ((hc.*asum)(accl_view, N, devX.p, 0, 0, &r), HCBLAS_INVALID);
  // extension: offset and increment
  EXPECT_EQ((hc.*asum)(accl_view, 60, devX.p, 4, 13, &r), HCBLAS_SUCCEEDS);
  EXPECT_EQ(r, oracle(60, X.data() + 13, 4));
  // type 2: Σ over the batch (sasum_array.cpp:278-290)
  for (int extra : {0, 7}) {
    const int bs = 128, xbo = N + extra;
    std::vector<T> xb = rnd<T>((size_t)xbo * bs, 10);
    for (size_t i = 0; i < xb.size(); i += 5) xb[i] = -xb[i];
    Dev<T> dxb(xb);
    EXPECT_EQ((hc.*asum_b)(accl_view, N, dxb.p, 1, 0, &r, xbo, bs),
              HCBLAS_SUCCEEDS);
    T ref = 0;
    for (int b = 0; b < bs; ++b) ref += oracle(N, xb.data() + (size_t)b * xbo, 1);
    EXPECT_EQ(r, ref);
  }
  end(name);
}

// ---------------------------------------------------------------- gemv
template <typename T, typename F, typename FB, typename O>
void gemv_tests(const char *name, F gemv, FB gemv_b, O oracle) {
  begin(name);
  Hcblaslibrary hc;
  hipStream_t accl_view = hc.currentAcclView;
  // hcblas_sgemv.cpp:197-356: M = 179, N = 19, alpha = 1, beta = 1
  const int M = 179, N = 19;
  const T alpha = 1, beta = 1;
  for (hcblasOrder order : {ColMajor, RowMajor}) {
    for (hcblasTranspose type : {NoTrans, Trans}) {
      const int lenx = type == NoTrans ? N : M, leny = type == NoTrans ? M : N;
      const int lda = (order == ColMajor ? M : N) + 3;  // extension: padded
      const int aoff = 5, xoff = 2, yoff = 4;
      std::vector<T> A = rnd<T>(aoff + (size_t)lda * (order == ColMajor ? N : M),
                                25);
      std::vector<T> x = rnd<T>(xoff + lenx, 10), y = rnd<T>(yoff + leny, 15);
      Dev<T> dA(A), dx(x), dy(y);
      EXPECT_EQ((hc.*gemv)(accl_view, order, type, M, N, alpha, dA.p, aoff,
                           lda, dx.p, xoff, 1, beta, dy.p, yoff, 1),
                HCBLAS_SUCCEEDS);
      std::vector<T> ref = y;
      oracle(order == RowMajor ? 0 : 1, type == NoTrans ? 0 : 1, M, N, alpha,
             A.data() + aoff, lda, x.data() + xoff, 1, beta,
             ref.data() + yoff, 1);
      EXPECT_EQ(same(dy.get(), ref), true);
      // return_correct (hcblas_sgemv.cpp:34-195)
      EXPECT_EQ((hc.*gemv)(accl_view, order, type, M, N, alpha, nullptr, aoff,
                           lda, dx.p, xoff, 1, beta, dy.p, yoff, 1),
                HCBLAS_INVALID);
      EXPECT_EQ((hc.*gemv)(accl_view, order, type, 0, N, alpha, dA.p, aoff,
                           lda, dx.p, xoff, 1, beta, dy.p, yoff, 1),
                HCBLAS_INVALID);
      EXPECT_EQ((hc.*gemv)(accl_view, order, type, M, N, alpha, dA.p, aoff,
                           lda, dx.p, xoff, 0, beta, dy.p, yoff, 1),
                HCBLAS_INVALID);
      // type 2 (hcblas_sgemv.cpp:358+): batchSize 128, A_batchOffset = M*N,
      // X_batchOffset = lenx, Y_batchOffset = leny; then larger offsets
      for (int extra : {0, 4}) {
        const int bs = 128, ldb = order == ColMajor ? M : N;
        const int64_t Abo = (int64_t)M * N + extra, xbo = lenx + extra,
                      ybo = leny + extra;
        std::vector<T> Ab = rnd<T>((size_t)Abo * bs, 25);
        std::vector<T> xb = rnd<T>((size_t)xbo * bs, 10);
        std::vector<T> yb = rnd<T>((size_t)ybo * bs, 15);
        Dev<T> dAb(Ab), dxb(xb), dyb(yb);
        EXPECT_EQ((hc.*gemv_b)(accl_view, order, type, M, N, alpha, dAb.p, 0,
                               Abo, ldb, dxb.p, 0, xbo, 1, beta, dyb.p, 0, ybo,
                               1, bs),
                  HCBLAS_SUCCEEDS);
        std::vector<T> refb = yb;
        for (int b = 0; b < bs; ++b)
          oracle(order == RowMajor ? 0 : 1, type == NoTrans ? 0 : 1, M, N,
                 alpha, Ab.data() + b * Abo, ldb, xb.data() + b * xbo, 1, beta,
                 refb.data() + b * ybo, 1);
        EXPECT_EQ(same(dyb.get(), refb), true);
      }
    }
  }
  end(name);
}

// ----------------------------------------------------------------- ger
template <typename T, typename F, typename FB, typename O>
void ger_tests(const char *name, F ger, FB ger_b, O oracle) {
  begin(name);
  Hcblaslibrary hc;
  hipStream_t accl_view = hc.currentAcclView;
  // hcblas_sger.cpp: M = 179, N = 19
  const int M = 179, N = 19;
  for (hcblasOrder order : {ColMajor, RowMajor}) {
    for (T alpha : {T(1), T(3)}) {
      const int lda = (order == ColMajor ? M : N) + 2;
      const int aoff = 3, xoff = 1, yoff = 6;
      std::vector<T> A = rnd<T>(aoff + (size_t)lda * (order == ColMajor ? N : M),
                                25);
      std::vector<T> x = rnd<T>(xoff + M, 10), y = rnd<T>(yoff + N, 15);
      Dev<T> dA(A), dx(x), dy(y);
      EXPECT_EQ((hc.*ger)(accl_view, order, M, N, alpha, dx.p, xoff, 1, dy.p,
                          yoff, 1, dA.p, aoff, lda),
                HCBLAS_SUCCEEDS);
      std::vector<T> ref = A;
      oracle(order == RowMajor ? 0 : 1, M, N, alpha, x.data() + xoff, 1,
             y.data() + yoff, 1, ref.data() + aoff, lda);
      EXPECT_EQ(same(dA.get(), ref), true);
      EXPECT_EQ((hc.*ger)(accl_view, order, M, N, alpha, nullptr, xoff, 1,
                          dy.p, yoff, 1, dA.p, aoff, lda),
                HCBLAS_INVALID);
      EXPECT_EQ((hc.*ger)(accl_view, order, M, 0, alpha, dx.p, xoff, 1, dy.p,
                          yoff, 1, dA.p, aoff, lda),
                HCBLAS_INVALID);
      EXPECT_EQ((hc.*ger)(accl_view, order, M, N, alpha, dx.p, xoff, 0, dy.p,
                          yoff, 1, dA.p, aoff, lda),
                HCBLAS_INVALID);
      // type 2: batchSize 128, A_batchOffset = M*N, X = M, Y = N (reference
      // strides), then larger batch offsets
      for (int extra : {0, 4}) {
        const int bs = 128, ldb = order == ColMajor ? M : N;
        const int64_t Abo = (int64_t)M * N + extra, xbo = M + extra,
                      ybo = N + extra;
        std::vector<T> Ab = rnd<T>((size_t)Abo * bs, 25);
        std::vector<T> xb = rnd<T>((size_t)xbo * bs, 10);
        std::vector<T> yb = rnd<T>((size_t)ybo * bs, 15);
        Dev<T> dAb(Ab), dxb(xb), dyb(yb);
        EXPECT_EQ((hc.*ger_b)(accl_view, order, M, N, alpha, dxb.p, 0, xbo, 1,
                              dyb.p, 0, ybo, 1, dAb.p, 0, Abo, ldb, bs),
                  HCBLAS_SUCCEEDS);
        std::vector<T> refb = Ab;
        for (int b = 0; b < bs; ++b)
          oracle(order == RowMajor ? 0 : 1, M, N, alpha, xb.data() + b * xbo,
                 1, yb.data() + b * ybo, 1, refb.data() + b * Abo, ldb);
        EXPECT_EQ(same(dAb.get(), refb), true);
      }
    }
  }
  end(name);
}

// ---------------------------------------------------------------- gemm
template <typename T, typename F, typename FB, typename O>
void gemm_tests(const char *name, F gemm, FB gemm_b, O oracle) {
  begin(name);
  Hcblaslibrary hc;
  hipStream_t accl_view = hc.currentAcclView;
  // hcblas_sgemm.cpp:34-960: M = 189, N = 9, K = 19, alpha = beta = 1,
  // A % 100, B % 15, C = 4 (type 2), every order x trans case
  const int M = 189, N = 9, K = 19;
  const T alpha = 1, beta = 1;
  for (hcblasOrder order : {ColMajor, RowMajor}) {
    for (hcblasTranspose ta : {NoTrans, Trans}) {
      for (hcblasTranspose tb : {NoTrans, Trans}) {
        const bool col = order == ColMajor;
        const int ar = col ? (ta == NoTrans ? M : K) : (ta == NoTrans ? K : M);
        const int ac = col ? (ta == NoTrans ? K : M) : (ta == NoTrans ? M : K);
        const int br = col ? (tb == NoTrans ? K : N) : (tb == NoTrans ? N : K);
        const int bc = col ? (tb == NoTrans ? N : K) : (tb == NoTrans ? K : N);
        const int lda = ar + 1, ldb = br + 2, ldc = (col ? M : N) + 3;
        const int aoff = 3, boff = 5, coff = 7;
        std::vector<T> A = rnd<T>(aoff + (size_t)lda * ac, 100);
        std::vector<T> B = rnd<T>(boff + (size_t)ldb * bc, 15);
        std::vector<T> C = rnd<T>(coff + (size_t)ldc * (col ? N : M), 10);
        Dev<T> dA(A), dB(B), dC(C);
        EXPECT_EQ((hc.*gemm)(accl_view, order, ta, tb, M, N, K, alpha, dA.p,
                             lda, dB.p, ldb, beta, dC.p, ldc, aoff, boff,
                             coff),
                  HCBLAS_SUCCEEDS);
        std::vector<T> ref = C;
        oracle(col ? 1 : 0, ta == NoTrans ? 0 : 1, tb == NoTrans ? 0 : 1, M,
               N, K, alpha, A.data() + aoff, lda, B.data() + boff, ldb, beta,
               ref.data() + coff, ldc);
        EXPECT_EQ(same(dC.get(), ref), true);
        if (ta == NoTrans && tb == NoTrans) {
          EXPECT_EQ((hc.*gemm)(accl_view, order, ta, tb, M, N, K, alpha,
                               nullptr, lda, dB.p, ldb, beta, dC.p, ldc, aoff,
                               boff, coff),
                    HCBLAS_INVALID);
          EXPECT_EQ((hc.*gemm)(accl_view, order, ta, tb, M, N, 0, alpha, dA.p,
                               lda, dB.p, ldb, beta, dC.p, ldc, aoff, boff,
                               coff),
                    HCBLAS_INVALID);
        }
      }
    }
  }
  // type 2 (hcblas_sgemm.cpp:1612+): batchSize 64 pointer arrays, C = 4
  {
    const int bs = 64;
    std::vector<T> A = rnd<T>((size_t)M * K * bs, 100);
    std::vector<T> B = rnd<T>((size_t)K * N * bs, 15);
    std::vector<T> C((size_t)M * N * bs, T(4));
    Dev<T> dA(A), dB(B), dC(C);
    std::vector<T *> pa(bs), pb(bs), pc(bs);
    for (int b = 0; b < bs; ++b) {
      pa[b] = dA.p + (size_t)b * M * K;
      pb[b] = dB.p + (size_t)b * K * N;
      pc[b] = dC.p + (size_t)b * M * N;
    }
    Dev<T *> da(pa), db(pb), dc(pc);
    EXPECT_EQ((hc.*gemm_b)(accl_view, ColMajor, NoTrans, NoTrans, M, N, K,
                           alpha, da.p, M, 0, db.p, K, 0, beta, dc.p, M, 0, 0,
                           0, 0, bs),
              HCBLAS_SUCCEEDS);
    std::vector<T> ref = C;
    for (int b = 0; b < bs; ++b)
      oracle(1, 0, 0, M, N, K, alpha, A.data() + (size_t)b * M * K, M,
             B.data() + (size_t)b * K * N, K, beta,
             ref.data() + (size_t)b * M * N, M);
    EXPECT_EQ(same(dC.get(), ref), true);
  }
  end(name);
}

template <typename T, typename C, typename F, typename O>
void complex_gemm_tests(const char *name, F gemm, O oracle) {
  begin(name);
  Hcblaslibrary hc;
  hipStream_t accl_view = hc.currentAcclView;
  // hcblas_cgemm.cpp: M = 189, N = 9, K = 19, alpha = beta = (1, 1)
  const int M = 189, N = 9, K = 19;
  const C alpha{1, 1}, beta{1, 1};
  for (hcblasOrder order : {ColMajor, RowMajor}) {
    for (hcblasTranspose ta : {NoTrans, Trans}) {
      for (hcblasTranspose tb : {NoTrans, Trans}) {
        const bool col = order == ColMajor;
        const int ar = col ? (ta == NoTrans ? M : K) : (ta == NoTrans ? K : M);
        const int ac = col ? (ta == NoTrans ? K : M) : (ta == NoTrans ? M : K);
        const int br = col ? (tb == NoTrans ? K : N) : (tb == NoTrans ? N : K);
        const int bc = col ? (tb == NoTrans ? N : K) : (tb == NoTrans ? K : N);
        const int lda = ar + 1, ldb = br, ldc = (col ? M : N) + 2;
        const int aoff = 1, boff = 2, coff = 3;
        std::vector<T> A = rnd<T>(2 * (aoff + (size_t)lda * ac), 10);
        std::vector<T> B = rnd<T>(2 * (boff + (size_t)ldb * bc), 10);
        std::vector<T> Cv = rnd<T>(2 * (coff + (size_t)ldc * (col ? N : M)), 10);
        Dev<T> dA(A), dB(B), dC(Cv);
        EXPECT_EQ((hc.*gemm)(accl_view, order, ta, tb, M, N, K, alpha,
                             reinterpret_cast<C *>(dA.p), aoff, lda,
                             reinterpret_cast<C *>(dB.p), boff, ldb, beta,
                             reinterpret_cast<C *>(dC.p), coff, ldc),
                  HCBLAS_SUCCEEDS);
        std::vector<T> ref = Cv;
        oracle(col ? 1 : 0, ta == NoTrans ? 0 : 1, tb == NoTrans ? 0 : 1, M,
               N, K, T(1), T(1), A.data() + 2 * aoff, lda,
               B.data() + 2 * boff, ldb, T(1), T(1), ref.data() + 2 * coff,
               ldc);
        EXPECT_EQ(same(dC.get(), ref), true);
      }
    }
  }
  end(name);
}

void hgemm_tests() {
  const char *name = "hcblas_hgemm";
  begin(name);
  Hcblaslibrary hc;
  hipStream_t accl_view = hc.currentAcclView;
  // Small integers: every product sum is below 2048, exact in binary16.
  const int M = 67, N = 9, K = 19;
  auto h = [](float f) {
    _Float16 v = (_Float16)f;
    hcblasHalf b;
    std::memcpy(&b, &v, 2);
    return b;
  };
  auto f = [](hcblasHalf b) {
    _Float16 v;
    std::memcpy(&v, &b, 2);
    return (float)v;
  };
  for (hcblasOrder order : {ColMajor, RowMajor}) {
    const bool col = order == ColMajor;
    std::vector<float> Af = rnd<float>((size_t)M * K, 4);
    std::vector<float> Bf = rnd<float>((size_t)K * N, 4);
    std::vector<float> Cf = rnd<float>((size_t)M * N, 4);
    std::vector<hcblasHalf> A(Af.size()), B(Bf.size()), C(Cf.size());
    for (size_t i = 0; i < A.size(); ++i) A[i] = h(Af[i]);
    for (size_t i = 0; i < B.size(); ++i) B[i] = h(Bf[i]);
    for (size_t i = 0; i < C.size(); ++i) C[i] = h(Cf[i]);
    Dev<hcblasHalf> dA(A), dB(B), dC(C);
    const hcblasHalf one = h(1.f), two = h(2.f);
    EXPECT_EQ(hc.hcblas_hgemm(accl_view, order, NoTrans, NoTrans, M, N, K, one,
                              dA.p, col ? M : K, dB.p, col ? K : N, two, dC.p,
                              col ? M : N, 0, 0, 0),
              HCBLAS_SUCCEEDS);
    std::vector<float> ref = Cf;
    orc_sgemm(col ? 1 : 0, 0, 0, M, N, K, 1.f, Af.data(), col ? M : K,
              Bf.data(), col ? K : N, 2.f, ref.data(), col ? M : N);
    std::vector<hcblasHalf> got = dC.get();
    bool ok = true;
    for (size_t i = 0; i < got.size(); ++i) ok = ok && f(got[i]) == ref[i];
    EXPECT_EQ(ok, true);
  }
  end(name);
}

}  // namespace

int main() {
  int dev = 0;
  if (hipGetDeviceCount(&dev) != hipSuccess || dev < 1) {
    std::printf("no GPU\n");
    return 2;
  }
  axpy_tests<float>("hcblas_saxpy",
                    static_cast<hcblasStatus (Hcblaslibrary::*)(
                        hipStream_t, int, const float &, const float *, int,
                        float *, int, int64_t, int64_t)>(
                        &Hcblaslibrary::hcblas_saxpy),
                    static_cast<hcblasStatus (Hcblaslibrary::*)(
                        hipStream_t, int, const float &, const float *, int,
                        int64_t, float *, int, int64_t, int64_t, int64_t,
                        int)>(&Hcblaslibrary::hcblas_saxpy),
                    orc_saxpy);
  axpy_tests<double>("hcblas_daxpy",
                     static_cast<hcblasStatus (Hcblaslibrary::*)(
                         hipStream_t, int, const double &, const double *,
                         int, double *, int, int64_t, int64_t)>(
                         &Hcblaslibrary::hcblas_daxpy),
                     static_cast<hcblasStatus (Hcblaslibrary::*)(
                         hipStream_t, int, const double &, const double *,
                         int, int64_t, double *, int, int64_t, int64_t,
                         int64_t, int)>(&Hcblaslibrary::hcblas_daxpy),
                     orc_daxpy);

#define SCAL_PTRS(NAME, AT, XT)                                              \
  static_cast<hcblasStatus (Hcblaslibrary::*)(hipStream_t, int, const AT &, \
                                              XT *, int, int64_t)>(          \
      &Hcblaslibrary::NAME),                                                 \
      static_cast<hcblasStatus (Hcblaslibrary::*)(                           \
          hipStream_t, int, const AT &, XT *, int, int64_t, int64_t, int)>(  \
          &Hcblaslibrary::NAME)
  scal_tests<float, float>("hcblas_sscal", 3.f, SCAL_PTRS(hcblas_sscal, float, float),
                    orc_sscal);
  scal_tests<double, double>("hcblas_dscal", 3.0,
                     SCAL_PTRS(hcblas_dscal, double, double), orc_dscal);
  scal_tests<hcComplex, float>(
      "hcblas_cscal", hcComplex{2.f, -1.f},
      SCAL_PTRS(hcblas_cscal, hcComplex, hcComplex),
      [](int64_t n, hcComplex a, hcComplex *x, int64_t inc) {
        orc_cscal(n, a.x, a.y, reinterpret_cast<float *>(x), inc);
      });
  scal_tests<hcDoubleComplex, double>(
      "hcblas_zscal", hcDoubleComplex{2.0, -1.0},
      SCAL_PTRS(hcblas_zscal, hcDoubleComplex, hcDoubleComplex),
      [](int64_t n, hcDoubleComplex a, hcDoubleComplex *x, int64_t inc) {
        orc_zscal(n, a.x, a.y, reinterpret_cast<double *>(x), inc);
      });
  scal_tests<hcComplex, float>("hcblas_csscal", 3.f,
                        SCAL_PTRS(hcblas_csscal, float, hcComplex),
                        [](int64_t n, float a, hcComplex *x, int64_t inc) {
                          orc_csscal(n, a, reinterpret_cast<float *>(x), inc);
                        });
  scal_tests<hcDoubleComplex, double>(
      "hcblas_zdscal", 3.0, SCAL_PTRS(hcblas_zdscal, double, hcDoubleComplex),
      [](int64_t n, double a, hcDoubleComplex *x, int64_t inc) {
        orc_zdscal(n, a, reinterpret_cast<double *>(x), inc);
      });
#undef SCAL_PTRS

  copy_tests<float>("hcblas_scopy",
                    static_cast<hcblasStatus (Hcblaslibrary::*)(
                        hipStream_t, int, const float *, int, int64_t, float *,
                        int, int64_t)>(&Hcblaslibrary::hcblas_scopy),
                    static_cast<hcblasStatus (Hcblaslibrary::*)(
                        hipStream_t, int, const float *, int, int64_t, float *,
                        int, int64_t, int64_t, int64_t, int)>(
                        &Hcblaslibrary::hcblas_scopy),
                    orc_scopy);
  copy_tests<double>("hcblas_dcopy",
                     static_cast<hcblasStatus (Hcblaslibrary::*)(
                         hipStream_t, int, const double *, int, int64_t,
                         double *, int, int64_t)>(&Hcblaslibrary::hcblas_dcopy),
                     static_cast<hcblasStatus (Hcblaslibrary::*)(
                         hipStream_t, int, const double *, int, int64_t,
                         double *, int, int64_t, int64_t, int64_t, int)>(
                         &Hcblaslibrary::hcblas_dcopy),
                     orc_dcopy);

  dot_tests<float>("hcblas_sdot",
                   static_cast<hcblasStatus (Hcblaslibrary::*)(
                       hipStream_t, int, const float *, int, int64_t,
                       const float *, int, int64_t, float &)>(
                       &Hcblaslibrary::hcblas_sdot),
                   static_cast<hcblasStatus (Hcblaslibrary::*)(
                       hipStream_t, int, const float *, int, int64_t,
                       const float *, int, int64_t, float &, int64_t, int64_t,
                       int)>(&Hcblaslibrary::hcblas_sdot),
                   orc_sdot);
  dot_tests<double>("hcblas_ddot",
                    static_cast<hcblasStatus (Hcblaslibrary::*)(
                        hipStream_t, int, const double *, int, int64_t,
                        const double *, int, int64_t, double &)>(
                        &Hcblaslibrary::hcblas_ddot),
                    static_cast<hcblasStatus (Hcblaslibrary::*)(
                        hipStream_t, int, const double *, int, int64_t,
                        const double *, int, int64_t, double &, int64_t,
                        int64_t, int)>(&Hcblaslibrary::hcblas_ddot),
                    orc_ddot);

  asum_tests<float>("hcblas_sasum",
                    static_cast<hcblasStatus (Hcblaslibrary::*)(
                        hipStream_t, int, float *, int, int64_t, float *)>(
                        &Hcblaslibrary::hcblas_sasum),
                    static_cast<hcblasStatus (Hcblaslibrary::*)(
                        hipStream_t, int, float *, int, int64_t, float *,
                        int64_t, int)>(&Hcblaslibrary::hcblas_sasum),
                    orc_sasum);
  asum_tests<double>("hcblas_dasum",
                     static_cast<hcblasStatus (Hcblaslibrary::*)(
                         hipStream_t, int, double *, int, int64_t, double *)>(
                         &Hcblaslibrary::hcblas_dasum),
                     static_cast<hcblasStatus (Hcblaslibrary::*)(
                         hipStream_t, int, double *, int, int64_t, double *,
                         int64_t, int)>(&Hcblaslibrary::hcblas_dasum),
                     orc_dasum);

  gemv_tests<float>(
      "hcblas_sgemv",
      static_cast<hcblasStatus (Hcblaslibrary::*)(
          hipStream_t, hcblasOrder, hcblasTranspose, int, int, const float &,
          float *, int64_t, int, float *, int64_t, int, const float &, float *,
          int64_t, int)>(&Hcblaslibrary::hcblas_sgemv),
      static_cast<hcblasStatus (Hcblaslibrary::*)(
          hipStream_t, hcblasOrder, hcblasTranspose, int, int, const float &,
          float *, int64_t, int64_t, int, float *, int64_t, int64_t, int,
          const float &, float *, int64_t, int64_t, int, int)>(
          &Hcblaslibrary::hcblas_sgemv),
      orc_sgemv);
  gemv_tests<double>(
      "hcblas_dgemv",
      static_cast<hcblasStatus (Hcblaslibrary::*)(
          hipStream_t, hcblasOrder, hcblasTranspose, int, int, const double &,
          double *, int64_t, int, double *, int64_t, int, const double &,
          double *, int64_t, int)>(&Hcblaslibrary::hcblas_dgemv),
      static_cast<hcblasStatus (Hcblaslibrary::*)(
          hipStream_t, hcblasOrder, hcblasTranspose, int, int, const double &,
          double *, int64_t, int64_t, int, double *, int64_t, int64_t, int,
          const double &, double *, int64_t, int64_t, int, int)>(
          &Hcblaslibrary::hcblas_dgemv),
      orc_dgemv);

  ger_tests<float>(
      "hcblas_sger",
      static_cast<hcblasStatus (Hcblaslibrary::*)(
          hipStream_t, hcblasOrder, int, int, const float &, const float *,
          int64_t, int, const float *, int64_t, int, float *, int64_t, int)>(
          &Hcblaslibrary::hcblas_sger),
      static_cast<hcblasStatus (Hcblaslibrary::*)(
          hipStream_t, hcblasOrder, int, int, const float &, const float *,
          int64_t, int64_t, int, const float *, int64_t, int64_t, int, float *,
          int64_t, int64_t, int, int)>(&Hcblaslibrary::hcblas_sger),
      orc_sger);
  ger_tests<double>(
      "hcblas_dger",
      static_cast<hcblasStatus (Hcblaslibrary::*)(
          hipStream_t, hcblasOrder, int, int, const double &, const double *,
          int64_t, int, const double *, int64_t, int, double *, int64_t, int)>(
          &Hcblaslibrary::hcblas_dger),
      static_cast<hcblasStatus (Hcblaslibrary::*)(
          hipStream_t, hcblasOrder, int, int, const double &, const double *,
          int64_t, int64_t, int, const double *, int64_t, int64_t, int,
          double *, int64_t, int64_t, int, int)>(&Hcblaslibrary::hcblas_dger),
      orc_dger);

  gemm_tests<float>(
      "hcblas_sgemm",
      static_cast<hcblasStatus (Hcblaslibrary::*)(
          hipStream_t, hcblasOrder, hcblasTranspose, hcblasTranspose, int, int,
          int, const float &, float *, int64_t, float *, int64_t,
          const float &, float *, int64_t, int64_t, int64_t, int64_t)>(
          &Hcblaslibrary::hcblas_sgemm),
      static_cast<hcblasStatus (Hcblaslibrary::*)(
          hipStream_t, hcblasOrder, hcblasTranspose, hcblasTranspose, int, int,
          int, const float &, float *[], int64_t, int64_t, float *[], int64_t,
          int64_t, const float &, float *[], int64_t, int64_t, int64_t,
          int64_t, int64_t, int)>(&Hcblaslibrary::hcblas_sgemm),
      orc_sgemm);
  gemm_tests<double>(
      "hcblas_dgemm",
      static_cast<hcblasStatus (Hcblaslibrary::*)(
          hipStream_t, hcblasOrder, hcblasTranspose, hcblasTranspose, int, int,
          int, const double &, double *, int64_t, double *, int64_t,
          const double &, double *, int64_t, int64_t, int64_t, int64_t)>(
          &Hcblaslibrary::hcblas_dgemm),
      static_cast<hcblasStatus (Hcblaslibrary::*)(
          hipStream_t, hcblasOrder, hcblasTranspose, hcblasTranspose, int, int,
          int, const double &, double *[], int64_t, int64_t, double *[],
          int64_t, int64_t, const double &, double *[], int64_t, int64_t,
          int64_t, int64_t, int64_t, int)>(&Hcblaslibrary::hcblas_dgemm),
      orc_dgemm);
  complex_gemm_tests<float, hcComplex>(
      "hcblas_cgemm",
      static_cast<hcblasStatus (Hcblaslibrary::*)(
          hipStream_t, hcblasOrder, hcblasTranspose, hcblasTranspose, int, int,
          int, const hcComplex &, hcComplex *, int64_t, int64_t, hcComplex *,
          int64_t, int64_t, const hcComplex &, hcComplex *, int64_t, int64_t)>(
          &Hcblaslibrary::hcblas_cgemm),
      orc_cgemm);
  complex_gemm_tests<double, hcDoubleComplex>(
      "hcblas_zgemm",
      static_cast<hcblasStatus (Hcblaslibrary::*)(
          hipStream_t, hcblasOrder, hcblasTranspose, hcblasTranspose, int, int,
          int, const hcDoubleComplex &, hcDoubleComplex *, int64_t, int64_t,
          hcDoubleComplex *, int64_t, int64_t, const hcDoubleComplex &,
          hcDoubleComplex *, int64_t, int64_t)>(&Hcblaslibrary::hcblas_zgemm),
      orc_zgemm);
  hgemm_tests();

  std::printf("%d checks, %d failed\n", g_checks, g_fail);
  return g_fail ? 1 : 0;
}
