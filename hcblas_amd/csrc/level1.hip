// Level-1 BLAS on gfx950: copy, scal (S/D/C/Z/Cs/Zd), axpy, dot, asum.
//
// Reference kernels replaced (rocmarchive/hcBLAS):
//   scopy_HC   lib/src/blas/scopy/scopy_array.cpp:29-45   (8-thread groups)
//   sscal_HC   lib/src/blas/sscal/sscal_array.cpp:29-47   (8-thread groups)
//   axpy_HC    lib/src/blas/saxpy/saxpy_array.cpp:29-97
//   sdot_HC    lib/src/blas/sdot/sdot_array.cpp:30-159    (host-side final sum)
//   sasum_HC   lib/src/blas/sasum/sasum_array.cpp:30-143
//
// Design: every routine is HBM-bound. Unit-stride, 16-byte-aligned operands
// take a dwordx4 path (four 16-byte loads in flight per thread, grid-stride
// over a grid sized to the CU count); everything else takes a scalar path that
// honours BLAS increments. dot/asum are one launch: per-workgroup partials in
// FP64, then the last workgroup to take a ticket sums the partials in a fixed
// order, so the result is deterministic and no host round trip is needed for
// the reduction itself.
#include "common.hpp"

namespace hcb {
namespace {

constexpr int kThreads = 256;
constexpr int kUnroll = 4;

template <typename T>
struct V16;  // 16-byte vector of T
template <>
struct V16<float> {
  using type = float4;
  static constexpr int n = 4;
};
template <>
struct V16<double> {
  using type = double2;
  static constexpr int n = 2;
};

__device__ __forceinline__ float &lane(float4 &v, int i) {
  return reinterpret_cast<float *>(&v)[i];
}
__device__ __forceinline__ double &lane(double2 &v, int i) {
  return reinterpret_cast<double *>(&v)[i];
}

// BLAS element index for an increment (negative increments walk backwards).
__device__ __forceinline__ int64_t blas_index(int64_t i, int64_t n,
                                              int64_t inc) {
  return inc > 0 ? i * inc : (n - 1 - i) * (-inc);
}

// ---------------------------------------------------------------------------
// Elementwise ops. Each op defines how one element of y (and x) is updated.
// ---------------------------------------------------------------------------
template <typename T>
struct CopyOp {
  __device__ __forceinline__ T operator()(T x, T) const { return x; }
  static constexpr bool reads_y = false;
};
template <typename T>
struct ScalOp {  // x = alpha*x (single rounding, NaN propagates)
  T alpha;
  __device__ __forceinline__ T operator()(T, T y) const { return y * alpha; }
  static constexpr bool reads_y = true;
};
template <typename T>
struct AxpyOp {  // y = fma(alpha, x, y)
  T alpha;
  __device__ __forceinline__ T operator()(T x, T y) const {
    return __builtin_fma(alpha, x, y);
  }
  static constexpr bool reads_y = true;
};
template <>
struct AxpyOp<float> {
  float alpha;
  __device__ __forceinline__ float operator()(float x, float y) const {
    return __builtin_fmaf(alpha, x, y);
  }
  static constexpr bool reads_y = true;
};

// Vector path: x, y 16-byte aligned, unit stride, nvec whole vectors,
// grid-stride with kUnroll independent vectors per thread per trip.
template <typename T, typename Op>
__global__ __launch_bounds__(kThreads) void elementwise_vec_kernel(
    int64_t nvec, const typename V16<T>::type *__restrict__ x,
    typename V16<T>::type *__restrict__ y, Op op) {
  using V = typename V16<T>::type;
  constexpr int VN = V16<T>::n;
  const int64_t stride = (int64_t)gridDim.x * kThreads;
  int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  for (; i + (kUnroll - 1) * stride < nvec; i += kUnroll * stride) {
    V xv[kUnroll], yv[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      xv[u] = x[i + u * stride];
      if (Op::reads_y) yv[u] = y[i + u * stride];
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      V r;
#pragma unroll
      for (int e = 0; e < VN; ++e) lane(r, e) = op(lane(xv[u], e), lane(yv[u], e));
      y[i + u * stride] = r;
    }
  }
  for (; i < nvec; i += stride) {
    V xv = x[i];
    V yv;
    if (Op::reads_y) yv = y[i];
    V r;
#pragma unroll
    for (int e = 0; e < VN; ++e) lane(r, e) = op(lane(xv, e), lane(yv, e));
    y[i] = r;
  }
}

// Scalar path: any increments. Element (b, i) of the batch lives at
// base + b*bstride + blas_index(i, n, inc).
template <typename T, typename Op>
__global__ __launch_bounds__(kThreads) void elementwise_strided_kernel(
    int64_t n, int64_t batch, const T *__restrict__ x, int64_t incx,
    int64_t bsx, T *__restrict__ y, int64_t incy, int64_t bsy, Op op) {
  const int64_t total = n * batch;
  const int64_t stride = (int64_t)gridDim.x * kThreads;
  for (int64_t f = (int64_t)blockIdx.x * kThreads + threadIdx.x; f < total;
       f += stride) {
    int64_t b = f / n, i = f - b * n;
    int64_t ix = b * bsx + blas_index(i, n, incx);
    int64_t iy = b * bsy + blas_index(i, n, incy);
    T yv = Op::reads_y ? y[iy] : T(0);
    y[iy] = op(x[ix], yv);
  }
}

// Complex scal. Products are rounded separately (no contraction), so the
// result is bit-identical to the CPU restatement compiled with
// -ffp-contract=off: re = ar*xr - ai*xi, im = ar*xi + ai*xr.
template <typename T>
struct CScalOp {
  T ar, ai;
};

template <typename T>
__device__ __forceinline__ T mul_rn(T a, T b);
template <>
__device__ __forceinline__ float mul_rn(float a, float b) {
  return __fmul_rn(a, b);
}
template <>
__device__ __forceinline__ double mul_rn(double a, double b) {
  return __dmul_rn(a, b);
}
template <typename T>
__device__ __forceinline__ T add_rn(T a, T b);
template <>
__device__ __forceinline__ float add_rn(float a, float b) {
  return __fadd_rn(a, b);
}
template <>
__device__ __forceinline__ double add_rn(double a, double b) {
  return __dadd_rn(a, b);
}
template <typename T>
__global__ __launch_bounds__(kThreads) void cscal_kernel(int64_t n,
                                                         int64_t batch, T *x,
                                                         int64_t incx,
                                                         int64_t bsx, T ar,
                                                         T ai, bool real_alpha) {
  const int64_t total = n * batch;
  const int64_t stride = (int64_t)gridDim.x * kThreads;
  for (int64_t f = (int64_t)blockIdx.x * kThreads + threadIdx.x; f < total;
       f += stride) {
    int64_t b = f / n, i = f - b * n;
    int64_t ix = 2 * (b * bsx + i * incx);
    T xr = x[ix], xi = x[ix + 1];
    if (real_alpha) {
      x[ix] = mul_rn(ar, xr);
      x[ix + 1] = mul_rn(ar, xi);
    } else {
      x[ix] = add_rn(mul_rn(ar, xr), -mul_rn(ai, xi));
      x[ix + 1] = add_rn(mul_rn(ar, xi), mul_rn(ai, xr));
    }
  }
}

// ---------------------------------------------------------------------------
// Reductions
// ---------------------------------------------------------------------------
template <typename T>
struct DotRed {
  __device__ __forceinline__ T operator()(T x, T y) const { return x * y; }
  static constexpr bool two = true;
};
template <typename T>
struct AsumRed {
  __device__ __forceinline__ T operator()(T x, T) const {
    return __builtin_fabs(x);
  }
  static constexpr bool two = false;
};
template <>
struct AsumRed<float> {
  __device__ __forceinline__ float operator()(float x, float) const {
    return __builtin_fabsf(x);
  }
  static constexpr bool two = false;
};

__device__ __forceinline__ double wave_sum(double v) {
  // 64-lane butterfly; lowered to DPP / ds_swizzle by the compiler.
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block sum of one double per thread, in a fixed order.
__device__ __forceinline__ double block_sum(double v, double *lds) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) lds[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < kThreads / 64; ++i) r += lds[i];
  }
  return r;  // valid on thread 0 only
}

// Second stage shared by both reduction kernels: publish this workgroup's
// partial; the last workgroup to arrive sums every partial in index order and
// writes the final value as T. The ticket returns to zero for the next call.
template <typename T>
__device__ void finish_reduction(double partial, double *partials,
                                 unsigned int *ticket, T *result,
                                 unsigned int nblocks, double *lds) {
  __shared__ bool last;
  const unsigned int bid = blockIdx.x;
  if (threadIdx.x == 0) {
    __hip_atomic_store(&partials[bid], partial, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    unsigned int t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL,
                                            __HIP_MEMORY_SCOPE_AGENT);
    last = (t == nblocks - 1);
  }
  __syncthreads();
  if (!last) return;
  double s = 0.0;
  for (unsigned int i = threadIdx.x; i < nblocks; i += kThreads)
    s += __hip_atomic_load(&partials[i], __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  double tot = block_sum(s, lds);
  if (threadIdx.x == 0) {
    *result = (T)tot;
    __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Vector-path reduction over nvec whole 16-byte vectors plus a scalar tail
// (tail handled by block 0). Accumulates in T per lane of the vector with
// kUnroll independent chains, then FP64 across lanes and workgroups.
template <typename T, typename Red>
__global__ __launch_bounds__(kThreads) void reduce_vec_kernel(
    int64_t n, const T *__restrict__ x, const T *__restrict__ y,
    double *partials, unsigned int *ticket, T *result, Red red) {
  using V = typename V16<T>::type;
  constexpr int VN = V16<T>::n;
  __shared__ double lds[kThreads / 64];
  const V *xv = reinterpret_cast<const V *>(x);
  const V *yv = reinterpret_cast<const V *>(y);
  const int64_t nvec = n / VN;
  const int64_t stride = (int64_t)gridDim.x * kThreads;
  T acc[kUnroll][VN];
#pragma unroll
  for (int u = 0; u < kUnroll; ++u)
#pragma unroll
    for (int e = 0; e < VN; ++e) acc[u][e] = T(0);
  int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  for (; i + (kUnroll - 1) * stride < nvec; i += kUnroll * stride) {
    V a[kUnroll], b[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      a[u] = xv[i + u * stride];
      if (Red::two) b[u] = yv[i + u * stride];
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u)
#pragma unroll
      for (int e = 0; e < VN; ++e)
        acc[u][e] += red(lane(a[u], e), lane(b[u], e));
  }
  for (; i < nvec; i += stride) {
    V a = xv[i], b;
    if (Red::two) b = yv[i];
#pragma unroll
    for (int e = 0; e < VN; ++e) acc[0][e] += red(lane(a, e), lane(b, e));
  }
  if (blockIdx.x == 0) {
    int64_t t = nvec * VN + threadIdx.x;
    if (t < n) acc[1][0] += red(x[t], Red::two ? y[t] : T(0));
  }
  double s = 0.0;
#pragma unroll
  for (int u = 0; u < kUnroll; ++u)
#pragma unroll
    for (int e = 0; e < VN; ++e) s += (double)acc[u][e];
  double part = block_sum(s, lds);
  finish_reduction<T>(part, partials, ticket, result, gridDim.x, lds);
}

// Strided / batched reduction: element (b, i) at b*bs + blas_index(i, n, inc).
template <typename T, typename Red>
__global__ __launch_bounds__(kThreads) void reduce_strided_kernel(
    int64_t n, int64_t batch, const T *__restrict__ x, int64_t incx,
    int64_t bsx, const T *__restrict__ y, int64_t incy, int64_t bsy,
    double *partials, unsigned int *ticket, T *result, Red red) {
  __shared__ double lds[kThreads / 64];
  const int64_t total = n * batch;
  const int64_t stride = (int64_t)gridDim.x * kThreads;
  T acc = T(0);
  for (int64_t f = (int64_t)blockIdx.x * kThreads + threadIdx.x; f < total;
       f += stride) {
    int64_t b = f / n, i = f - b * n;
    T xv = x[b * bsx + blas_index(i, n, incx)];
    T yv = Red::two ? y[b * bsy + blas_index(i, n, incy)] : T(0);
    acc += red(xv, yv);
  }
  double part = block_sum((double)acc, lds);
  finish_reduction<T>(part, partials, ticket, result, gridDim.x, lds);
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
int stream_grid(hcblasHandle_t h, int64_t work_items) {
  // Enough workgroups to cover every SIMD several times over, never more
  // than the work needs.
  int64_t want = cdiv(work_items, (int64_t)kThreads * kUnroll);
  int64_t cap = (int64_t)h->num_cus * 8;
  if (want < 1) want = 1;
  return (int)(want < cap ? want : cap);
}

template <typename T, typename Op>
hcblasStatus_t run_elementwise(hcblasHandle_t h, int64_t n, int64_t batch,
                               const T *x, int64_t incx, T *y, int64_t incy,
                               Op op) {
  constexpr int VN = V16<T>::n;
  if (n <= 0 || batch <= 0) return HCBLAS_STATUS_SUCCESS;
  // A unit-stride batch with stride n is one contiguous vector of n*batch.
  const bool contiguous = incx == 1 && incy == 1;
  const int64_t flat = n * batch;
  if (contiguous && aligned16(x) && aligned16(y)) {
    int64_t nvec = flat / VN;
    if (nvec > 0) {
      using V = typename V16<T>::type;
      elementwise_vec_kernel<T, Op>
          <<<stream_grid(h, nvec), kThreads, 0, h->stream>>>(
              nvec, reinterpret_cast<const V *>(x), reinterpret_cast<V *>(y),
              op);
    }
    int64_t tail = flat - nvec * VN;
    if (tail > 0)
      elementwise_strided_kernel<T, Op><<<1, kThreads, 0, h->stream>>>(
          tail, 1, x + nvec * VN, 1, 0, y + nvec * VN, 1, 0, op);
    return launch_status();
  }
  if (contiguous) {
    elementwise_strided_kernel<T, Op>
        <<<stream_grid(h, flat), kThreads, 0, h->stream>>>(flat, 1, x, 1, 0, y,
                                                           1, 0, op);
    return launch_status();
  }
  elementwise_strided_kernel<T, Op>
      <<<stream_grid(h, flat), kThreads, 0, h->stream>>>(n, batch, x, incx, n,
                                                         y, incy, n, op);
  return launch_status();
}

template <typename T, typename Red>
hcblasStatus_t run_reduce(hcblasHandle_t h, int64_t n, int64_t batch,
                          const T *x, int64_t incx, const T *y, int64_t incy,
                          T *result, Red red) {
  constexpr int VN = V16<T>::n;
  T *dres = reinterpret_cast<T *>(h->d_result);
  const int64_t flat = n * batch;
  int grid = stream_grid(h, flat / VN);
  if (grid > kMaxPartials) grid = kMaxPartials;
  const bool contiguous = incx == 1 && (!Red::two || incy == 1);
  if (contiguous && aligned16(x) && (!Red::two || aligned16(y))) {
    reduce_vec_kernel<T, Red><<<grid, kThreads, 0, h->stream>>>(
        flat, x, Red::two ? y : x, h->d_partials, h->d_ticket, dres, red);
  } else if (contiguous) {
    reduce_strided_kernel<T, Red><<<grid, kThreads, 0, h->stream>>>(
        flat, 1, x, 1, 0, Red::two ? y : x, 1, 0, h->d_partials, h->d_ticket,
        dres, red);
  } else {
    reduce_strided_kernel<T, Red><<<grid, kThreads, 0, h->stream>>>(
        n, batch, x, incx, n, Red::two ? y : x, incy, n, h->d_partials,
        h->d_ticket, dres, red);
  }
  hcblasStatus_t st = launch_status();
  if (st != HCBLAS_STATUS_SUCCESS) return st;
  // The API returns a host scalar, so the call completes before returning
  // (same contract as sdot_array.cpp:136-158).
  if (hipMemcpyAsync(h->h_result, dres, sizeof(T), hipMemcpyDeviceToHost,
                     h->stream) != hipSuccess)
    return HCBLAS_STATUS_EXECUTION_FAILED;
  if (hipStreamSynchronize(h->stream) != hipSuccess)
    return HCBLAS_STATUS_EXECUTION_FAILED;
  *result = *reinterpret_cast<T *>(h->h_result);
  return HCBLAS_STATUS_SUCCESS;
}

// ---------------------------------------------------------------------------
// Argument checks shared by the C entry points.
// ---------------------------------------------------------------------------
template <typename T, typename Op>
hcblasStatus_t l1_update(hcblasHandle_t h, int n, const T *x, int incx, T *y,
                         int incy, int batch, Op op) {
  if (!handle_ok(h)) return HCBLAS_STATUS_NOT_INITIALIZED;
  if (batch < 0 || incx == 0 || incy == 0) return HCBLAS_STATUS_INVALID_VALUE;
  if (n <= 0 || batch == 0) return HCBLAS_STATUS_SUCCESS;
  if (x == nullptr || y == nullptr) return HCBLAS_STATUS_INVALID_VALUE;
  return run_elementwise<T, Op>(h, n, batch, x, incx, y, incy, op);
}

template <typename T>
hcblasStatus_t l1_scal(hcblasHandle_t h, int n, const T *alpha, T *x, int incx,
                       int batch) {
  if (!handle_ok(h)) return HCBLAS_STATUS_NOT_INITIALIZED;
  if (batch < 0) return HCBLAS_STATUS_INVALID_VALUE;
  // BLAS: non-positive increment is a no-op for scal.
  if (n <= 0 || batch == 0 || incx <= 0) return HCBLAS_STATUS_SUCCESS;
  if (alpha == nullptr || x == nullptr) return HCBLAS_STATUS_INVALID_VALUE;
  return run_elementwise<T, ScalOp<T>>(h, n, batch, x, incx, x, incx,
                                       ScalOp<T>{*alpha});
}

template <typename T>
hcblasStatus_t l1_cscal(hcblasHandle_t h, int n, T ar, T ai, bool real_alpha,
                        T *x, int incx, int batch) {
  if (!handle_ok(h)) return HCBLAS_STATUS_NOT_INITIALIZED;
  if (batch < 0) return HCBLAS_STATUS_INVALID_VALUE;
  if (n <= 0 || batch == 0 || incx <= 0) return HCBLAS_STATUS_SUCCESS;
  if (x == nullptr) return HCBLAS_STATUS_INVALID_VALUE;
  int64_t total = (int64_t)n * batch;
  cscal_kernel<T><<<stream_grid(h, total), kThreads, 0, h->stream>>>(
      n, batch, x, incx, n, ar, ai, real_alpha);
  return launch_status();
}

template <typename T, typename Red>
hcblasStatus_t l1_reduce(hcblasHandle_t h, int n, const T *x, int incx,
                         const T *y, int incy, T *result, int batch, Red red) {
  if (!handle_ok(h)) return HCBLAS_STATUS_NOT_INITIALIZED;
  if (result == nullptr || batch < 0) return HCBLAS_STATUS_INVALID_VALUE;
  if (Red::two && (incx == 0 || incy == 0)) return HCBLAS_STATUS_INVALID_VALUE;
  // BLAS: n <= 0 gives 0; asum with a non-positive increment gives 0.
  if (n <= 0 || batch == 0 || (!Red::two && incx <= 0)) {
    *result = T(0);
    return HCBLAS_STATUS_SUCCESS;
  }
  if (x == nullptr || (Red::two && y == nullptr))
    return HCBLAS_STATUS_INVALID_VALUE;
  return run_reduce<T, Red>(h, n, batch, x, incx, y, incy, result, red);
}

}  // namespace
}  // namespace hcb

using namespace hcb;

namespace {
template <typename T, bool REAL, typename AT, typename CT>
hcblasStatus_t cscal_entry(hcblasHandle_t handle, int n, const AT *alpha,
                           CT *x, int incx, int batch) {
  if (handle_ok(handle) && alpha == nullptr)
    return HCBLAS_STATUS_INVALID_VALUE;
  const T *a = reinterpret_cast<const T *>(alpha);
  T ar = a ? a[0] : T(0);
  T ai = (a && !REAL) ? a[1] : T(0);
  return l1_cscal<T>(handle, n, ar, ai, REAL, reinterpret_cast<T *>(x), incx,
                     batch);
}
}  // namespace

extern "C" {

// asum -------------------------------------------------------------------
hcblasStatus_t hcblasSasum(hcblasHandle_t handle, int n, const float *x,
                           int incx, float *result) {
  return l1_reduce<float>(handle, n, x, incx, nullptr, 1, result, 1,
                          AsumRed<float>{});
}
hcblasStatus_t hcblasDasum(hcblasHandle_t handle, int n, const double *x,
                           int incx, double *result) {
  return l1_reduce<double>(handle, n, x, incx, nullptr, 1, result, 1,
                           AsumRed<double>{});
}
hcblasStatus_t hcblasSasumBatched(hcblasHandle_t handle, int n, const float *x,
                                  int incx, float *result, int batchCount) {
  return l1_reduce<float>(handle, n, x, incx, nullptr, 1, result, batchCount,
                          AsumRed<float>{});
}
hcblasStatus_t hcblasDasumBatched(hcblasHandle_t handle, int n,
                                  const double *x, int incx, double *result,
                                  int batchCount) {
  return l1_reduce<double>(handle, n, x, incx, nullptr, 1, result, batchCount,
                           AsumRed<double>{});
}

// axpy -------------------------------------------------------------------
hcblasStatus_t hcblasSaxpy(hcblasHandle_t handle, int n, const float *alpha,
                           const float *x, int incx, float *y, int incy) {
  if (handle_ok(handle) && alpha == nullptr) return HCBLAS_STATUS_INVALID_VALUE;
  return l1_update<float>(handle, n, x, incx, y, incy, 1,
                          AxpyOp<float>{alpha ? *alpha : 0.f});
}
hcblasStatus_t hcblasDaxpy(hcblasHandle_t handle, int n, const double *alpha,
                           const double *x, int incx, double *y, int incy) {
  if (handle_ok(handle) && alpha == nullptr) return HCBLAS_STATUS_INVALID_VALUE;
  return l1_update<double>(handle, n, x, incx, y, incy, 1,
                           AxpyOp<double>{alpha ? *alpha : 0.0});
}
hcblasStatus_t hcblasSaxpyBatched(hcblasHandle_t handle, int n,
                                  const float *alpha, const float *x, int incx,
                                  float *y, int incy, int batchCount) {
  if (handle_ok(handle) && alpha == nullptr) return HCBLAS_STATUS_INVALID_VALUE;
  return l1_update<float>(handle, n, x, incx, y, incy, batchCount,
                          AxpyOp<float>{alpha ? *alpha : 0.f});
}
hcblasStatus_t hcblasDaxpyBatched(hcblasHandle_t handle, int n,
                                  const double *alpha, const double *x,
                                  int incx, double *y, int incy,
                                  int batchCount) {
  if (handle_ok(handle) && alpha == nullptr) return HCBLAS_STATUS_INVALID_VALUE;
  return l1_update<double>(handle, n, x, incx, y, incy, batchCount,
                           AxpyOp<double>{alpha ? *alpha : 0.0});
}

// copy -------------------------------------------------------------------
hcblasStatus_t hcblasScopy(hcblasHandle_t handle, int n, const float *x,
                           int incx, float *y, int incy) {
  return l1_update<float>(handle, n, x, incx, y, incy, 1, CopyOp<float>{});
}
hcblasStatus_t hcblasDcopy(hcblasHandle_t handle, int n, const double *x,
                           int incx, double *y, int incy) {
  return l1_update<double>(handle, n, x, incx, y, incy, 1, CopyOp<double>{});
}
hcblasStatus_t hcblasScopyBatched(hcblasHandle_t handle, int n, const float *x,
                                  int incx, float *y, int incy,
                                  int batchCount) {
  return l1_update<float>(handle, n, x, incx, y, incy, batchCount,
                          CopyOp<float>{});
}
hcblasStatus_t hcblasDcopyBatched(hcblasHandle_t handle, int n,
                                  const double *x, int incx, double *y,
                                  int incy, int batchCount) {
  return l1_update<double>(handle, n, x, incx, y, incy, batchCount,
                           CopyOp<double>{});
}

// dot --------------------------------------------------------------------
hcblasStatus_t hcblasSdot(hcblasHandle_t handle, int n, const float *x,
                          int incx, const float *y, int incy, float *result) {
  return l1_reduce<float>(handle, n, x, incx, y, incy, result, 1,
                          DotRed<float>{});
}
hcblasStatus_t hcblasDdot(hcblasHandle_t handle, int n, const double *x,
                          int incx, const double *y, int incy,
                          double *result) {
  return l1_reduce<double>(handle, n, x, incx, y, incy, result, 1,
                           DotRed<double>{});
}
hcblasStatus_t hcblasSdotBatched(hcblasHandle_t handle, int n, const float *x,
                                 int incx, const float *y, int incy,
                                 float *result, int batchCount) {
  return l1_reduce<float>(handle, n, x, incx, y, incy, result, batchCount,
                          DotRed<float>{});
}
hcblasStatus_t hcblasDdotBatched(hcblasHandle_t handle, int n,
                                 const double *x, int incx, const double *y,
                                 int incy, double *result, int batchCount) {
  return l1_reduce<double>(handle, n, x, incx, y, incy, result, batchCount,
                           DotRed<double>{});
}

// scal -------------------------------------------------------------------
hcblasStatus_t hcblasSscal(hcblasHandle_t handle, int n, const float *alpha,
                           float *x, int incx) {
  return l1_scal<float>(handle, n, alpha, x, incx, 1);
}
hcblasStatus_t hcblasDscal(hcblasHandle_t handle, int n, const double *alpha,
                           double *x, int incx) {
  return l1_scal<double>(handle, n, alpha, x, incx, 1);
}
hcblasStatus_t hcblasSscalBatched(hcblasHandle_t handle, int n,
                                  const float *alpha, float *x, int incx,
                                  int batchCount) {
  return l1_scal<float>(handle, n, alpha, x, incx, batchCount);
}
hcblasStatus_t hcblasDscalBatched(hcblasHandle_t handle, int n,
                                  const double *alpha, double *x, int incx,
                                  int batchCount) {
  return l1_scal<double>(handle, n, alpha, x, incx, batchCount);
}


hcblasStatus_t hcblasCscal(hcblasHandle_t handle, int n, const hcComplex *alpha,
                           hcComplex *x, int incx) {
  return cscal_entry<float, false>(handle, n, alpha, x, incx, 1);
}
hcblasStatus_t hcblasZscal(hcblasHandle_t handle, int n,
                           const hcDoubleComplex *alpha, hcDoubleComplex *x,
                           int incx) {
  return cscal_entry<double, false>(handle, n, alpha, x, incx, 1);
}
hcblasStatus_t hcblasCsscal(hcblasHandle_t handle, int n, const float *alpha,
                            hcComplex *x, int incx) {
  return cscal_entry<float, true>(handle, n, alpha, x, incx, 1);
}
hcblasStatus_t hcblasZdscal(hcblasHandle_t handle, int n, const double *alpha,
                            hcDoubleComplex *x, int incx) {
  return cscal_entry<double, true>(handle, n, alpha, x, incx, 1);
}
hcblasStatus_t hcblasCscalBatched(hcblasHandle_t handle, int n,
                                  const hcComplex *alpha, hcComplex *x,
                                  int incx, int batchCount) {
  return cscal_entry<float, false>(handle, n, alpha, x, incx, batchCount);
}
hcblasStatus_t hcblasZscalBatched(hcblasHandle_t handle, int n,
                                  const hcDoubleComplex *alpha,
                                  hcDoubleComplex *x, int incx,
                                  int batchCount) {
  return cscal_entry<double, false>(handle, n, alpha, x, incx, batchCount);
}
hcblasStatus_t hcblasCsscalBatched(hcblasHandle_t handle, int n,
                                   const float *alpha, hcComplex *x, int incx,
                                   int batchCount) {
  return cscal_entry<float, true>(handle, n, alpha, x, incx, batchCount);
}
hcblasStatus_t hcblasZdscalBatched(hcblasHandle_t handle, int n,
                                   const double *alpha, hcDoubleComplex *x,
                                   int incx, int batchCount) {
  return cscal_entry<double, true>(handle, n, alpha, x, incx, batchCount);
}

}  // extern "C"
