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
// take the vector path: one 16-byte vector per thread over a full grid
// (measured on MI355X to stream faster than grid-stride loops), with
// non-temporal loads / stores once the operands exceed the caches. Everything
// else takes a scalar path that honours BLAS increments. dot / asum: each of
// at most 256 workgroups reduces to an FP64 partial, and a one-workgroup
// launch on the same stream sums the partials in index order (deterministic)
// into the handle's pinned host block, whose sequence word the host polls.
#include "common.hpp"

namespace hcb {
namespace {

constexpr int kThreads = 256;
// Operands larger than this (bytes) bypass the caches (non-temporal).
constexpr int64_t kStreamBytes = 64ll << 20;
// Vector-path reductions: loads in flight per thread and maximum grid. A
// sweep on MI355X (sdot, 2^28) measured 7.2 TB/s at 256 workgroups x 4
// vectors in flight and 6.6 TB/s at 512. The finish is a second launch:
// an in-kernel "last workgroup sums" ticket costs an agent-scope release
// (an L2 write-back on this multi-XCD part) per workgroup, 9 us of the
// 2^28 SDOT (307 against 298 us).
constexpr int kReduceUnroll = 4;
constexpr int kReduceGrid = kReducePartials;

typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));

template <typename T>
struct V16;  // native 16-byte vector of T
template <>
struct V16<float> {
  using type = f4v;
  static constexpr int n = 4;
};
template <>
struct V16<double> {
  using type = d2v;
  static constexpr int n = 2;
};

template <bool NT, typename V>
__device__ __forceinline__ V vload(const V *p) {
  if constexpr (NT)
    return __builtin_nontemporal_load(p);
  else
    return *p;
}
template <bool NT, typename V>
__device__ __forceinline__ void vstore(V *p, V v) {
  if constexpr (NT)
    __builtin_nontemporal_store(v, p);
  else
    *p = v;
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

// Vector path: x, y 16-byte aligned, unit stride; one vector per thread
// (grid-stride only beyond kMaxReduceBlocks workgroups).
template <typename T, typename Op, bool NT>
__global__ __launch_bounds__(kThreads) void elementwise_vec_kernel(
    int64_t nvec, const typename V16<T>::type *__restrict__ x,
    typename V16<T>::type *__restrict__ y, Op op) {
  using V = typename V16<T>::type;
  constexpr int VN = V16<T>::n;
  const int64_t stride = (int64_t)gridDim.x * kThreads;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < nvec;
       i += stride) {
    const V xv = vload<NT>(x + i);
    V yv{};
    if (Op::reads_y) yv = vload<NT>(y + i);
    V r;
#pragma unroll
    for (int e = 0; e < VN; ++e) r[e] = op(xv[e], yv[e]);
    vstore<NT>(y + i, r);
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

// Complex scal. Products are rounded separately (contraction is switched
// off in the kernel body), so the result is bit-identical to the CPU
// restatement compiled with -ffp-contract=off:
// re = ar*xr - ai*xi, im = ar*xi + ai*xr. The arithmetic is written in the
// kernel body itself: a helper defined outside the pragma's scope would be
// contracted after inlining.
template <typename T>
__global__ __launch_bounds__(kThreads) void cscal_kernel(int64_t n,
                                                         int64_t batch, T *x,
                                                         int64_t incx,
                                                         int64_t bsx, T ar,
                                                         T ai, bool real_alpha) {
#pragma clang fp contract(off)
  const int64_t total = n * batch;
  const int64_t stride = (int64_t)gridDim.x * kThreads;
  for (int64_t f = (int64_t)blockIdx.x * kThreads + threadIdx.x; f < total;
       f += stride) {
    int64_t b = f / n, i = f - b * n;
    int64_t ix = 2 * (b * bsx + i * incx);
    T xr = x[ix], xi = x[ix + 1];
    if (real_alpha) {
      x[ix] = ar * xr;
      x[ix + 1] = ar * xi;
    } else {
      const T p1 = ar * xr, p2 = ai * xi, p3 = ar * xi, p4 = ai * xr;
      x[ix] = p1 - p2;
      x[ix + 1] = p3 + p4;
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

// Block sum of one double per thread, in a fixed order (valid on thread 0).
template <int TH = kThreads>
__device__ __forceinline__ double block_sum(double v, double *lds) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) lds[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < TH / 64; ++i) r += lds[i];
  }
  return r;
}

// Final sum of the reduction kernels' partials (one per workgroup, at most
// kReduceGrid, written with plain stores; this launch follows on the same
// stream, so the kernel boundary makes them visible). The partials are added
// in workgroup order, so the result does not depend on timing. The result
// goes straight into the handle's pinned host block, then this call's
// sequence number, which the host waits for (release at system scope orders
// the two stores for the host).
template <typename T>
__global__ __launch_bounds__(kThreads) void reduce_finish_kernel(
    int nparts, const double *__restrict__ partials, T *result,
    unsigned long long *done, unsigned long long seq) {
  __shared__ double lds[kThreads / 64];
  const double v = (int)threadIdx.x < nparts ? partials[threadIdx.x] : 0.0;
  const double tot = block_sum(v, lds);
  if (threadIdx.x == 0) {
    *result = (T)tot;
    __hip_atomic_store(done, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// Vector-path reduction: grid-stride over 16-byte vectors of x (and y) with
// U vectors in flight per thread; the scalar tail (< one vector) is added by
// block 0. Each workgroup writes one FP64 partial; reduce_finish_kernel sums
// them.
template <typename T, typename Red, bool NT, int U, int TH = kThreads>
__global__ __launch_bounds__(TH) void reduce_vec_kernel(
    int64_t n, const T *__restrict__ x, const T *__restrict__ y,
    double *partials, Red red) {
  using V = typename V16<T>::type;
  constexpr int VN = V16<T>::n;
  __shared__ double lds[TH / 64];
  const V *xv = reinterpret_cast<const V *>(x);
  const V *yv = reinterpret_cast<const V *>(y);
  const int64_t nvec = n / VN;
  const int64_t stride = (int64_t)gridDim.x * TH;
  T acc[U][VN];
#pragma unroll
  for (int u = 0; u < U; ++u)
#pragma unroll
    for (int e = 0; e < VN; ++e) acc[u][e] = T(0);
  int64_t i = (int64_t)blockIdx.x * TH + threadIdx.x;
  for (; i + (U - 1) * stride < nvec; i += U * stride) {
    V a[U], b[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      a[u] = vload<NT>(xv + i + u * stride);
      if (Red::two) b[u] = vload<NT>(yv + i + u * stride);
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int e = 0; e < VN; ++e) acc[u][e] += red(a[u][e], b[u][e]);
  }
  for (; i < nvec; i += stride) {
    const V a = vload<NT>(xv + i);
    V b{};
    if (Red::two) b = vload<NT>(yv + i);
#pragma unroll
    for (int e = 0; e < VN; ++e) acc[0][e] += red(a[e], b[e]);
  }
  double s = 0.0;
#pragma unroll
  for (int u = 0; u < U; ++u)
#pragma unroll
    for (int e = 0; e < VN; ++e) s += (double)acc[u][e];
  if (blockIdx.x == 0) {
    const int64_t t = nvec * VN + threadIdx.x;
    if (t < n) s += (double)red(x[t], Red::two ? y[t] : T(0));
  }
  const double part = block_sum<TH>(s, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = part;
}

// Strided / batched reduction: element (b, i) at b*bs + blas_index(i, n, inc).
template <typename T, typename Red>
__global__ __launch_bounds__(kThreads) void reduce_strided_kernel(
    int64_t n, int64_t batch, const T *__restrict__ x, int64_t incx,
    int64_t bsx, const T *__restrict__ y, int64_t incy, int64_t bsy,
    double *partials, Red red) {
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
  const double part = block_sum((double)acc, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = part;
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
int full_grid(int64_t work_items) {
  int64_t g = cdiv(work_items, kThreads);
  if (g < 1) g = 1;
  return (int)(g < kMaxReduceBlocks ? g : kMaxReduceBlocks);
}

template <typename T, typename Op>
hcblasStatus_t run_elementwise(hcblasHandle_t h, int64_t n, int64_t batch,
                               const T *x, int64_t incx, int64_t bsx, T *y,
                               int64_t incy, int64_t bsy, Op op) {
  constexpr int VN = V16<T>::n;
  using V = typename V16<T>::type;
  if (n <= 0 || batch <= 0) return HCBLAS_STATUS_SUCCESS;
  // A unit-stride batch with batch stride n is one contiguous vector of
  // n*batch.
  const bool contiguous =
      incx == 1 && incy == 1 && (batch == 1 || (bsx == n && bsy == n));
  const int64_t flat = n * batch;
  if (contiguous && aligned16(x) && aligned16(y)) {
    const int64_t nvec = flat / VN;
    if (nvec > 0) {
      const bool nt = flat * (int64_t)sizeof(T) >= kStreamBytes;
      if (nt)
        elementwise_vec_kernel<T, Op, true>
            <<<full_grid(nvec), kThreads, 0, h->stream>>>(
                nvec, reinterpret_cast<const V *>(x), reinterpret_cast<V *>(y),
                op);
      else
        elementwise_vec_kernel<T, Op, false>
            <<<full_grid(nvec), kThreads, 0, h->stream>>>(
                nvec, reinterpret_cast<const V *>(x), reinterpret_cast<V *>(y),
                op);
    }
    const int64_t tail = flat - nvec * VN;
    if (tail > 0)
      elementwise_strided_kernel<T, Op><<<1, kThreads, 0, h->stream>>>(
          tail, 1, x + nvec * VN, 1, 0, y + nvec * VN, 1, 0, op);
    return launch_status();
  }
  if (contiguous) {
    elementwise_strided_kernel<T, Op>
        <<<full_grid(flat), kThreads, 0, h->stream>>>(flat, 1, x, 1, 0, y, 1,
                                                      0, op);
    return launch_status();
  }
  elementwise_strided_kernel<T, Op>
      <<<full_grid(flat), kThreads, 0, h->stream>>>(n, batch, x, incx, bsx, y,
                                                    incy, bsy, op);
  return launch_status();
}

// Waits until the reduction of call `seq` has published its result in the
// handle's pinned block (see tree_finish).
hcblasStatus_t wait_result(hcblasHandle_t h, unsigned long long seq) {
  const unsigned long long *w =
      reinterpret_cast<const unsigned long long *>(h->h_result + 1);
  for (unsigned spin = 1;; ++spin) {
    if (__atomic_load_n(w, __ATOMIC_ACQUIRE) == seq)
      return HCBLAS_STATUS_SUCCESS;
    if ((spin & 1023u) == 0) {
      const hipError_t q = hipStreamQuery(h->stream);
      if (q == hipSuccess)  // drained: the kernel's stores are visible now
        return __atomic_load_n(w, __ATOMIC_ACQUIRE) == seq
                   ? HCBLAS_STATUS_SUCCESS
                   : HCBLAS_STATUS_EXECUTION_FAILED;
      if (q != hipErrorNotReady) return HCBLAS_STATUS_EXECUTION_FAILED;
    }
    __builtin_ia32_pause();
  }
}

template <typename T, typename Red>
hcblasStatus_t run_reduce(hcblasHandle_t h, int64_t n, int64_t batch,
                          const T *x, int64_t incx, int64_t bsx, const T *y,
                          int64_t incy, int64_t bsy, T *result, Red red) {
  constexpr int VN = V16<T>::n;
  T *dres = reinterpret_cast<T *>(h->h_result_dev);
  unsigned long long *done =
      reinterpret_cast<unsigned long long *>(h->h_result_dev + 1);
  const unsigned long long seq = ++h->result_seq;
  const int64_t flat = n * batch;
  const bool contiguous = incx == 1 && (!Red::two || incy == 1) &&
                          (batch == 1 || (bsx == n && (!Red::two || bsy == n)));
  const T *yy = Red::two ? y : x;
  int nparts = 1;
  if (contiguous && aligned16(x) && (!Red::two || aligned16(y))) {
    // FP32 asum keeps 16 vectors in flight per lane (SASUM 2^28: 6.05
    // against 5.86 TB/s with 8).
    // FP64 asum runs 512-thread workgroups with 4 vectors in flight (DASUM
    // 2^27 151.7 us against 165.2 us for 256 x 8 in a sweep on MI355X; the
    // other forms measured fastest at 256 threads).
    constexpr bool wide = !Red::two && sizeof(T) == 8;
    constexpr int TH = wide ? 2 * kThreads : kThreads;
    constexpr int U = Red::two ? kReduceUnroll
                      : wide   ? kReduceUnroll
                               : 4 * kReduceUnroll;
    const int64_t nvec = flat / VN;
    const int grid = (int)std::max<int64_t>(
        1, std::min<int64_t>(cdiv(nvec, TH * U), kReduceGrid));
    const int64_t bytes = flat * (int64_t)sizeof(T) * (Red::two ? 2 : 1);
    if (bytes >= kStreamBytes)
      reduce_vec_kernel<T, Red, true, U, TH>
          <<<grid, TH, 0, h->stream>>>(flat, x, yy, h->d_partials, red);
    else
      reduce_vec_kernel<T, Red, false, U, TH>
          <<<grid, TH, 0, h->stream>>>(flat, x, yy, h->d_partials, red);
    nparts = grid;
  } else if (contiguous) {
    nparts = (int)std::min<int64_t>(full_grid(flat), kReduceGrid);
    reduce_strided_kernel<T, Red><<<nparts, kThreads, 0, h->stream>>>(
        flat, 1, x, 1, 0, yy, 1, 0, h->d_partials, red);
  } else {
    nparts = (int)std::min<int64_t>(full_grid(flat), kReduceGrid);
    reduce_strided_kernel<T, Red><<<nparts, kThreads, 0, h->stream>>>(
        n, batch, x, incx, bsx, yy, incy, bsy, h->d_partials, red);
  }
  reduce_finish_kernel<T><<<1, kThreads, 0, h->stream>>>(
      nparts, h->d_partials, dres, done, seq);
  hcblasStatus_t st = launch_status();
  if (st != HCBLAS_STATUS_SUCCESS) return st;
  // The API returns a host scalar, so the call waits for its result (same
  // contract as sdot_array.cpp:136-158): it spins on the sequence word the
  // kernel publishes after the result, and returns as soon as it appears,
  // without waiting for the kernel's end-of-kernel signal. If the stream
  // drains without the word appearing, the launch failed.
  st = wait_result(h, seq);
  if (st != HCBLAS_STATUS_SUCCESS) return st;
  *result = *reinterpret_cast<volatile T *>(h->h_result);
  return HCBLAS_STATUS_SUCCESS;
}

// ---------------------------------------------------------------------------
// Argument checks shared by the C entry points.
// ---------------------------------------------------------------------------
// Batch strides: bsx / bsy elements between the vectors of a batch; a
// negative value selects the C API's stride of n elements (hcblas.cpp:376).
inline int64_t batch_stride(int64_t bs, int n) { return bs < 0 ? n : bs; }

template <typename T, typename Op>
hcblasStatus_t l1_update(hcblasHandle_t h, int n, const T *x, int incx, T *y,
                         int incy, int batch, Op op, int64_t bsx = -1,
                         int64_t bsy = -1) {
  if (!handle_ok(h)) return HCBLAS_STATUS_NOT_INITIALIZED;
  if (batch < 0 || incx == 0 || incy == 0) return HCBLAS_STATUS_INVALID_VALUE;
  if (n <= 0 || batch == 0) return HCBLAS_STATUS_SUCCESS;
  if (x == nullptr || y == nullptr) return HCBLAS_STATUS_INVALID_VALUE;
  return run_elementwise<T, Op>(h, n, batch, x, incx, batch_stride(bsx, n), y,
                                incy, batch_stride(bsy, n), op);
}

template <typename T>
hcblasStatus_t l1_scal(hcblasHandle_t h, int n, const T *alpha, T *x, int incx,
                       int batch, int64_t bsx = -1) {
  if (!handle_ok(h)) return HCBLAS_STATUS_NOT_INITIALIZED;
  if (batch < 0) return HCBLAS_STATUS_INVALID_VALUE;
  // BLAS: non-positive increment is a no-op for scal.
  if (n <= 0 || batch == 0 || incx <= 0) return HCBLAS_STATUS_SUCCESS;
  if (alpha == nullptr || x == nullptr) return HCBLAS_STATUS_INVALID_VALUE;
  const int64_t bs = batch_stride(bsx, n);
  return run_elementwise<T, ScalOp<T>>(h, n, batch, x, incx, bs, x, incx, bs,
                                       ScalOp<T>{*alpha});
}

template <typename T>
hcblasStatus_t l1_cscal(hcblasHandle_t h, int n, T ar, T ai, bool real_alpha,
                        T *x, int incx, int batch, int64_t bsx = -1) {
  if (!handle_ok(h)) return HCBLAS_STATUS_NOT_INITIALIZED;
  if (batch < 0) return HCBLAS_STATUS_INVALID_VALUE;
  if (n <= 0 || batch == 0 || incx <= 0) return HCBLAS_STATUS_SUCCESS;
  if (x == nullptr) return HCBLAS_STATUS_INVALID_VALUE;
  int64_t total = (int64_t)n * batch;
  cscal_kernel<T><<<full_grid(total), kThreads, 0, h->stream>>>(
      n, batch, x, incx, batch_stride(bsx, n), ar, ai, real_alpha);
  return launch_status();
}

template <typename T, typename Red>
hcblasStatus_t l1_reduce(hcblasHandle_t h, int n, const T *x, int incx,
                         const T *y, int incy, T *result, int batch, Red red,
                         int64_t bsx = -1, int64_t bsy = -1) {
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
  return run_reduce<T, Red>(h, n, batch, x, incx, batch_stride(bsx, n), y,
                            incy, batch_stride(bsy, n), result, red);
}

}  // namespace
}  // namespace hcb

using namespace hcb;

namespace {
template <typename T, bool REAL, typename AT, typename CT>
hcblasStatus_t cscal_entry(hcblasHandle_t handle, int n, const AT *alpha,
                           CT *x, int incx, int batch, int64_t bsx = -1) {
  if (handle_ok(handle) && alpha == nullptr)
    return HCBLAS_STATUS_INVALID_VALUE;
  const T *a = reinterpret_cast<const T *>(alpha);
  T ar = a ? a[0] : T(0);
  T ai = (a && !REAL) ? a[1] : T(0);
  return l1_cscal<T>(handle, n, ar, ai, REAL, reinterpret_cast<T *>(x), incx,
                     batch, bsx);
}

// Offset forms (Hcblaslibrary overloads): a null base stays null so that the
// argument checks see it.
template <typename P>
P *at(P *p, int64_t off) {
  return p ? p + off : nullptr;
}

// axpy: argument checks, then the netlib quick return for alpha == 0 (y is
// not touched, so Inf / NaN in x do not reach it; saxpy.f, and the
// reference's saxpy_array.cpp:190).
template <typename T>
hcblasStatus_t axpy_entry(hcblasHandle_t handle, int n, const T *alpha,
                          const T *x, int incx, T *y, int incy, int batch,
                          int64_t bsx = -1, int64_t bsy = -1) {
  if (!handle_ok(handle)) return HCBLAS_STATUS_NOT_INITIALIZED;
  if (alpha == nullptr) return HCBLAS_STATUS_INVALID_VALUE;
  if (*alpha == T(0)) {
    if (batch < 0 || incx == 0 || incy == 0) return HCBLAS_STATUS_INVALID_VALUE;
    return HCBLAS_STATUS_SUCCESS;
  }
  return l1_update<T>(handle, n, x, incx, y, incy, batch, AxpyOp<T>{*alpha},
                      bsx, bsy);
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
  return axpy_entry<float>(handle, n, alpha, x, incx, y, incy, 1);
}
hcblasStatus_t hcblasDaxpy(hcblasHandle_t handle, int n, const double *alpha,
                           const double *x, int incx, double *y, int incy) {
  return axpy_entry<double>(handle, n, alpha, x, incx, y, incy, 1);
}
hcblasStatus_t hcblasSaxpyBatched(hcblasHandle_t handle, int n,
                                  const float *alpha, const float *x, int incx,
                                  float *y, int incy, int batchCount) {
  return axpy_entry<float>(handle, n, alpha, x, incx, y, incy, batchCount);
}
hcblasStatus_t hcblasDaxpyBatched(hcblasHandle_t handle, int n,
                                  const double *alpha, const double *x,
                                  int incx, double *y, int incy,
                                  int batchCount) {
  return axpy_entry<double>(handle, n, alpha, x, incx, y, incy, batchCount);
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

// Offset and batch-offset forms (hcblaslib.h:106-135, 319-479) -----------
hcblasStatus_t hcblas_saxpy(hcblasHandle_t handle, int N, const float *alpha,
                            const float *X, int incX, float *Y, int incY,
                            int64_t xOffset, int64_t yOffset) {
  return axpy_entry<float>(handle, N, alpha, at(X, xOffset), incX,
                           at(Y, yOffset), incY, 1);
}
hcblasStatus_t hcblas_saxpy_batched(hcblasHandle_t handle, int N,
                                    const float *alpha, const float *X,
                                    int incX, int64_t X_batchOffset, float *Y,
                                    int incY, int64_t Y_batchOffset,
                                    int64_t xOffset, int64_t yOffset,
                                    int batchSize) {
  if (X_batchOffset < 0 || Y_batchOffset < 0)
    return handle_ok(handle) ? HCBLAS_STATUS_INVALID_VALUE
                             : HCBLAS_STATUS_NOT_INITIALIZED;
  return axpy_entry<float>(handle, N, alpha, at(X, xOffset), incX,
                           at(Y, yOffset), incY, batchSize, X_batchOffset,
                           Y_batchOffset);
}
hcblasStatus_t hcblas_daxpy(hcblasHandle_t handle, int N, const double *alpha,
                            const double *X, int incX, double *Y, int incY,
                            int64_t xOffset, int64_t yOffset) {
  return axpy_entry<double>(handle, N, alpha, at(X, xOffset), incX,
                           at(Y, yOffset), incY, 1);
}
hcblasStatus_t hcblas_daxpy_batched(hcblasHandle_t handle, int N,
                                    const double *alpha, const double *X,
                                    int incX, int64_t X_batchOffset, double *Y,
                                    int incY, int64_t Y_batchOffset,
                                    int64_t xOffset, int64_t yOffset,
                                    int batchSize) {
  if (X_batchOffset < 0 || Y_batchOffset < 0)
    return handle_ok(handle) ? HCBLAS_STATUS_INVALID_VALUE
                             : HCBLAS_STATUS_NOT_INITIALIZED;
  return axpy_entry<double>(handle, N, alpha, at(X, xOffset), incX,
                           at(Y, yOffset), incY, batchSize, X_batchOffset,
                           Y_batchOffset);
}

hcblasStatus_t hcblas_sscal(hcblasHandle_t handle, int N, const float *alpha,
                            float *X, int incX, int64_t xOffset) {
  return l1_scal<float>(handle, N, alpha, at(X, xOffset), incX, 1);
}
hcblasStatus_t hcblas_sscal_batched(hcblasHandle_t handle, int N,
                                    const float *alpha, float *X, int incX,
                                    int64_t xOffset, int64_t X_batchOffset,
                                    int batchSize) {
  if (X_batchOffset < 0)
    return handle_ok(handle) ? HCBLAS_STATUS_INVALID_VALUE
                             : HCBLAS_STATUS_NOT_INITIALIZED;
  return l1_scal<float>(handle, N, alpha, at(X, xOffset), incX, batchSize,
                        X_batchOffset);
}
hcblasStatus_t hcblas_dscal(hcblasHandle_t handle, int N, const double *alpha,
                            double *X, int incX, int64_t xOffset) {
  return l1_scal<double>(handle, N, alpha, at(X, xOffset), incX, 1);
}
hcblasStatus_t hcblas_dscal_batched(hcblasHandle_t handle, int N,
                                    const double *alpha, double *X, int incX,
                                    int64_t xOffset, int64_t X_batchOffset,
                                    int batchSize) {
  if (X_batchOffset < 0)
    return handle_ok(handle) ? HCBLAS_STATUS_INVALID_VALUE
                             : HCBLAS_STATUS_NOT_INITIALIZED;
  return l1_scal<double>(handle, N, alpha, at(X, xOffset), incX, batchSize,
                         X_batchOffset);
}

#define HCB_CSCAL_OFFSET(NAME, T, REAL, AT, CT)                               \
  hcblasStatus_t NAME(hcblasHandle_t handle, int N, const AT *alpha, CT *X,  \
                      int incX, int64_t xOffset) {                            \
    return cscal_entry<T, REAL>(handle, N, alpha, at(X, xOffset), incX, 1);   \
  }                                                                           \
  hcblasStatus_t NAME##_batched(hcblasHandle_t handle, int N, const AT *alpha, \
                                CT *X, int incX, int64_t xOffset,             \
                                int64_t X_batchOffset, int batchSize) {       \
    if (X_batchOffset < 0)                                                    \
      return handle_ok(handle) ? HCBLAS_STATUS_INVALID_VALUE                  \
                               : HCBLAS_STATUS_NOT_INITIALIZED;               \
    return cscal_entry<T, REAL>(handle, N, alpha, at(X, xOffset), incX,       \
                                batchSize, X_batchOffset);                    \
  }
HCB_CSCAL_OFFSET(hcblas_cscal, float, false, hcComplex, hcComplex)
HCB_CSCAL_OFFSET(hcblas_zscal, double, false, hcDoubleComplex, hcDoubleComplex)
HCB_CSCAL_OFFSET(hcblas_csscal, float, true, float, hcComplex)
HCB_CSCAL_OFFSET(hcblas_zdscal, double, true, double, hcDoubleComplex)
#undef HCB_CSCAL_OFFSET

hcblasStatus_t hcblas_scopy(hcblasHandle_t handle, int N, const float *X,
                            int incX, int64_t xOffset, float *Y, int incY,
                            int64_t yOffset) {
  return l1_update<float>(handle, N, at(X, xOffset), incX, at(Y, yOffset),
                          incY, 1, CopyOp<float>{});
}
hcblasStatus_t hcblas_scopy_batched(hcblasHandle_t handle, int N,
                                    const float *X, int incX, int64_t xOffset,
                                    float *Y, int incY, int64_t yOffset,
                                    int64_t X_batchOffset,
                                    int64_t Y_batchOffset, int batchSize) {
  if (X_batchOffset < 0 || Y_batchOffset < 0)
    return handle_ok(handle) ? HCBLAS_STATUS_INVALID_VALUE
                             : HCBLAS_STATUS_NOT_INITIALIZED;
  return l1_update<float>(handle, N, at(X, xOffset), incX, at(Y, yOffset),
                          incY, batchSize, CopyOp<float>{}, X_batchOffset,
                          Y_batchOffset);
}
hcblasStatus_t hcblas_dcopy(hcblasHandle_t handle, int N, const double *X,
                            int incX, int64_t xOffset, double *Y, int incY,
                            int64_t yOffset) {
  return l1_update<double>(handle, N, at(X, xOffset), incX, at(Y, yOffset),
                           incY, 1, CopyOp<double>{});
}
hcblasStatus_t hcblas_dcopy_batched(hcblasHandle_t handle, int N,
                                    const double *X, int incX,
                                    int64_t xOffset, double *Y, int incY,
                                    int64_t yOffset, int64_t X_batchOffset,
                                    int64_t Y_batchOffset, int batchSize) {
  if (X_batchOffset < 0 || Y_batchOffset < 0)
    return handle_ok(handle) ? HCBLAS_STATUS_INVALID_VALUE
                             : HCBLAS_STATUS_NOT_INITIALIZED;
  return l1_update<double>(handle, N, at(X, xOffset), incX, at(Y, yOffset),
                           incY, batchSize, CopyOp<double>{}, X_batchOffset,
                           Y_batchOffset);
}

hcblasStatus_t hcblas_sdot(hcblasHandle_t handle, int N, const float *X,
                           int incX, int64_t xOffset, const float *Y, int incY,
                           int64_t yOffset, float *result) {
  return l1_reduce<float>(handle, N, at(X, xOffset), incX, at(Y, yOffset),
                          incY, result, 1, DotRed<float>{});
}
hcblasStatus_t hcblas_sdot_batched(hcblasHandle_t handle, int N,
                                   const float *X, int incX, int64_t xOffset,
                                   const float *Y, int incY, int64_t yOffset,
                                   float *result, int64_t X_batchOffset,
                                   int64_t Y_batchOffset, int batchSize) {
  if (X_batchOffset < 0 || Y_batchOffset < 0)
    return handle_ok(handle) ? HCBLAS_STATUS_INVALID_VALUE
                             : HCBLAS_STATUS_NOT_INITIALIZED;
  return l1_reduce<float>(handle, N, at(X, xOffset), incX, at(Y, yOffset),
                          incY, result, batchSize, DotRed<float>{},
                          X_batchOffset, Y_batchOffset);
}
hcblasStatus_t hcblas_ddot(hcblasHandle_t handle, int N, const double *X,
                           int incX, int64_t xOffset, const double *Y,
                           int incY, int64_t yOffset, double *result) {
  return l1_reduce<double>(handle, N, at(X, xOffset), incX, at(Y, yOffset),
                           incY, result, 1, DotRed<double>{});
}
hcblasStatus_t hcblas_ddot_batched(hcblasHandle_t handle, int N,
                                   const double *X, int incX, int64_t xOffset,
                                   const double *Y, int incY, int64_t yOffset,
                                   double *result, int64_t X_batchOffset,
                                   int64_t Y_batchOffset, int batchSize) {
  if (X_batchOffset < 0 || Y_batchOffset < 0)
    return handle_ok(handle) ? HCBLAS_STATUS_INVALID_VALUE
                             : HCBLAS_STATUS_NOT_INITIALIZED;
  return l1_reduce<double>(handle, N, at(X, xOffset), incX, at(Y, yOffset),
                           incY, result, batchSize, DotRed<double>{},
                           X_batchOffset, Y_batchOffset);
}

hcblasStatus_t hcblas_sasum(hcblasHandle_t handle, int N, const float *X,
                            int incX, int64_t xOffset, float *result) {
  return l1_reduce<float>(handle, N, at(X, xOffset), incX, nullptr, 1, result,
                          1, AsumRed<float>{});
}
hcblasStatus_t hcblas_sasum_batched(hcblasHandle_t handle, int N,
                                    const float *X, int incX, int64_t xOffset,
                                    float *result, int64_t X_batchOffset,
                                    int batchSize) {
  if (X_batchOffset < 0)
    return handle_ok(handle) ? HCBLAS_STATUS_INVALID_VALUE
                             : HCBLAS_STATUS_NOT_INITIALIZED;
  return l1_reduce<float>(handle, N, at(X, xOffset), incX, nullptr, 1, result,
                          batchSize, AsumRed<float>{}, X_batchOffset);
}
hcblasStatus_t hcblas_dasum(hcblasHandle_t handle, int N, const double *X,
                            int incX, int64_t xOffset, double *result) {
  return l1_reduce<double>(handle, N, at(X, xOffset), incX, nullptr, 1, result,
                           1, AsumRed<double>{});
}
hcblasStatus_t hcblas_dasum_batched(hcblasHandle_t handle, int N,
                                    const double *X, int incX,
                                    int64_t xOffset, double *result,
                                    int64_t X_batchOffset, int batchSize) {
  if (X_batchOffset < 0)
    return handle_ok(handle) ? HCBLAS_STATUS_INVALID_VALUE
                             : HCBLAS_STATUS_NOT_INITIALIZED;
  return l1_reduce<double>(handle, N, at(X, xOffset), incX, nullptr, 1, result,
                           batchSize, AsumRed<double>{}, X_batchOffset);
}

}  // extern "C"
