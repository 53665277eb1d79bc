// Level-2 BLAS on gfx950: {S,D}gemv and {S,D}ger, plain and batched.
//
// Reference kernels replaced (rocmarchive/hcBLAS):
//   gemv_TransA    lib/src/blas/sgemv/sgemv_array.cpp:30-147  (one WG / column)
//   gemv_NoTransA  lib/src/blas/sgemv/sgemv_array.cpp:520-563 (one thread / row)
//   gemv_alpha0_*  lib/src/blas/sgemv/sgemv_array.cpp:709-825
//   ger_HC         lib/src/blas/sger/sger_array.cpp:27-98     (one element / thread)
//
// Every matrix is handled as column-major STORAGE S (R contiguous rows, Cc
// columns, leading dimension ld). Row-major A is column-major A^T, so
//   col-major N, row-major T  ->  "panel" kernel:  y[r] = sum_c S[r,c] x[c]
//   col-major T, row-major N  ->  "column" kernel: y[c] = sum_r S[r,c] x[r]
// with a "band" variant of each for tall operands (a workgroup reads each
// column of its row band as one contiguous run; x or the band's sums stay in
// registers). All stream S once with 16-byte loads and split the reduction
// across workgroups when the output alone cannot fill 256 CUs; split partials
// go to a per-handle workspace and a small finish kernel on the same stream
// sums them in a fixed order, so results are deterministic.
//
// GER updates S[r,c] = fma(v[r], alpha*w[c], S[r,c]) where w is the vector
// that indexes storage columns; with v = x, w = y for column-major and
// v = y, w = x for row-major this is bit-identical to CBLAS (netlib
// "temp = alpha*y(j); a(i,j) += x(i)*temp", row-major by argument swap).
#include "common.hpp"
#include "gemv_plan.hpp"

namespace hcb {
namespace {

constexpr int kThreads = 256;
static_assert(kThreads == kGemvThreads, "plan and kernels agree");

typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));

template <typename T>
struct VecT;
template <>
struct VecT<float> {
  using type = f4v;
  static constexpr int n = 4;
};
template <>
struct VecT<double> {
  using type = d2v;
  static constexpr int n = 2;
};

template <typename T>
__device__ __forceinline__ T vlane(const typename VecT<T>::type &v, int i) {
  return v[i];
}

// 16-byte load, non-temporal when NT (operands far larger than the caches).
template <bool NT, typename V>
__device__ __forceinline__ V vld(const V *p) {
  if constexpr (NT)
    return __builtin_nontemporal_load(p);
  else
    return *p;
}
template <bool NT, typename V>
__device__ __forceinline__ void vst(V *p, V v) {
  if constexpr (NT)
    __builtin_nontemporal_store(v, p);
  else
    *p = v;
}

// Operands at least this large (bytes) are streamed non-temporally.
constexpr int64_t kStreamBytes = 64ll << 20;

__device__ __forceinline__ int64_t bidx(int64_t i, int64_t n, int64_t inc) {
  return inc > 0 ? i * inc : (n - 1 - i) * (-inc);
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <typename T>
__device__ __forceinline__ T fma_t(T a, T b, T c);
template <>
__device__ __forceinline__ float fma_t(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}
template <>
__device__ __forceinline__ double fma_t(double a, double b, double c) {
  return __builtin_fma(a, b, c);
}

struct GemvShape {
  int64_t R, Cc, ld;    // storage rows (contiguous), columns, leading dim
  int64_t incx, incy;   // BLAS increments of x (reduction) and y (output)
  int64_t sA, sx, sy;   // batch strides (elements)
  int splits;           // reduction split count (>= 1)
};

// Final value of one output element: y = alpha*sum + beta*y; beta == 0 does
// not read y (CBLAS: y need not be initialised).
template <typename T>
__device__ __forceinline__ void store_y(T *y, T sum, T alpha, T beta) {
  T r = alpha * sum;
  if (beta != T(0)) r = fma_t(beta, *y, r);
  *y = r;
}

// Split reductions: every workgroup stores its partial sums with plain
// stores to part[split][output], and gemv_finish_kernel, launched next on the
// same stream, adds them (sixteen running sums over the splits k = j mod 16,
// combined as a fixed tree, so sixteen loads per thread are in flight: SGEMV-N
// 16384^2 6846 GB/s against 6753 with eight) and writes y; the order is
// fixed, so results are deterministic. (An in-kernel
// last-arriver reduction needs an agent-scope release per workgroup, which on
// a multi-XCD MI355X is an L2 write-back each: measured 5.1 against 6.4 TB/s
// for SGEMV-N 16384^2 with 1024 workgroups.)
template <typename T>
__global__ __launch_bounds__(kThreads) void gemv_finish_kernel(
    int64_t len_y, int splits, const T *__restrict__ part, int64_t spart,
    T *y, int64_t incy, int64_t sy, T alpha, T beta) {
  const int64_t z = blockIdx.z;
  part += z * spart;
  y += z * sy;
  for (int64_t o = (int64_t)blockIdx.x * kThreads + threadIdx.x; o < len_y;
       o += (int64_t)gridDim.x * kThreads) {
    T s[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) s[j] = T(0);
    int k = 0;
    for (; k + 16 <= splits; k += 16) {
#pragma unroll
      for (int j = 0; j < 16; ++j) s[j] += part[(int64_t)(k + j) * len_y + o];
    }
    for (int j = 0; k < splits; ++k, ++j) s[j] += part[(int64_t)k * len_y + o];
#pragma unroll
    for (int w = 8; w >= 1; w >>= 1)
#pragma unroll
      for (int j = 0; j < w; ++j) s[j] = s[j] + s[j + w];
    store_y(&y[bidx(o, len_y, incy)], s[0], alpha, beta);
  }
}

// ---------------------------------------------------------------------------
// Panel kernel: y[r] = alpha * sum_c S[r,c] x[c] + beta*y[r].
// Workgroup = 4 waves over the same 64*VN rows (one 16-byte vector per lane);
// the waves interleave the workgroup's columns (c = c0 + wave + 4u) with U
// loads in flight per lane, and LDS sums the four waves. blockIdx.y splits the
// columns across workgroups. (Measured on MI355X: 6.4 TB/s at 16384^2 fp32,
// against 3.5 TB/s for one wave per row band walking all columns.)
// ---------------------------------------------------------------------------
template <typename T, bool VEC, bool NT>
__global__ __launch_bounds__(kThreads) void gemv_panel_kernel(
    GemvShape s, const T *__restrict__ A, const T *__restrict__ x, T *y,
    T alpha, T beta, T *part) {
  constexpr int VN = VEC ? VecT<T>::n : 1;
  constexpr int ROWS = 64 * VN;
  constexpr int U = 8;
  __shared__ T red[4][ROWS];
  const int64_t z = blockIdx.z;
  A += z * s.sA;
  x += z * s.sx;
  y += z * s.sy;
  // wave is uniform across the wave: readfirstlane lets the compiler fetch
  // x[c] with scalar loads instead of a vector load per A load.
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int64_t rb = (int64_t)blockIdx.x * ROWS;
  const int64_t r0 = rb + (int64_t)lane * VN;
  const int split = blockIdx.y;
  const int64_t cper = cdiv_dev(s.Cc, s.splits);
  const int64_t c_begin = split * cper;
  const int64_t c_end = c_begin + cper < s.Cc ? c_begin + cper : s.Cc;

  T acc[VN];
#pragma unroll
  for (int e = 0; e < VN; ++e) acc[e] = T(0);

  if (r0 < s.R) {  // on the vector path R % VN == 0
    const T *col = A + r0;
    int64_t c = c_begin + wave;
    if constexpr (VEC) {
      using V = typename VecT<T>::type;
      for (; c + 4 * (U - 1) < c_end; c += 4 * U) {
        V a[U];
        T xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          a[u] = vld<NT>(reinterpret_cast<const V *>(col + (c + 4 * u) * s.ld));
          xv[u] = x[bidx(c + 4 * u, s.Cc, s.incx)];
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int e = 0; e < VN; ++e) acc[e] = fma_t(a[u][e], xv[u], acc[e]);
      }
      for (; c < c_end; c += 4) {
        const V a = vld<NT>(reinterpret_cast<const V *>(col + c * s.ld));
        const T xv = x[bidx(c, s.Cc, s.incx)];
#pragma unroll
        for (int e = 0; e < VN; ++e) acc[e] = fma_t(a[e], xv, acc[e]);
      }
    } else {
      for (; c < c_end; c += 4)
        acc[0] = fma_t(col[c * s.ld], x[bidx(c, s.Cc, s.incx)], acc[0]);
    }
  }
#pragma unroll
  for (int e = 0; e < VN; ++e) red[wave][lane * VN + e] = acc[e];
  __syncthreads();
  // Four-wave sum, in wave order, into red[0].
  for (int i = threadIdx.x; i < ROWS; i += kThreads)
    red[0][i] = ((red[0][i] + red[1][i]) + red[2][i]) + red[3][i];
  __syncthreads();
  if (s.splits == 1) {
    for (int i = threadIdx.x; i < ROWS; i += kThreads)
      if (rb + i < s.R)
        store_y(&y[bidx(rb + i, s.R, s.incy)], red[0][i], alpha, beta);
    return;
  }
  T *pp = part + (z * s.splits + split) * s.R;
  for (int i = threadIdx.x; i < ROWS; i += kThreads)
    if (rb + i < s.R) pp[rb + i] = red[0][i];
}

// ---------------------------------------------------------------------------
// Band-N kernel: the panel kernel's job for operands with R >= 1024*VN*4.
// Workgroup = (one band of XR*256*VN rows, one block of columns): the band's
// sums stay in registers (XR vectors per thread) while the workgroup walks
// its columns UC at a time, each column's band being one contiguous run
// (16 KiB) read by all 256 threads; x[c] is wave-uniform (scalar loads). The
// loads of UC columns are all issued before their FMAs (a scheduling barrier
// keeps the compiler from interleaving them one by one). Column blocks are
// sized so that bands x blocks is about one workgroup per CU with at least
// 256 columns per block, so the partials cost at most 2/256 of the traffic.
// Measured on MI355X at 16384^2: 155 us fp32 (6.9 TB/s) and 308 us fp64
// against 166 / 315 us for the panel kernel. Rows of a partial last band
// load a clamped row (the band's first) and are never stored.
// ---------------------------------------------------------------------------

template <typename T, bool NT>
__global__ __launch_bounds__(kThreads) void gemv_bandn_kernel(
    GemvShape s, const T *__restrict__ A, const T *__restrict__ x, T *y,
    T alpha, T beta, T *part) {
  using V = typename VecT<T>::type;
  constexpr int VN = VecT<T>::n;
  constexpr int XR = kBandNRows, UC = kBandNCols;
  constexpr int64_t STEP = (int64_t)kThreads * VN;
  const int64_t z = blockIdx.z;
  A += z * s.sA;
  x += z * s.sx;
  y += z * s.sy;
  // blockIdx.x walks the column blocks and blockIdx.y the bands: consecutive
  // workgroups (spread over the XCDs in turn) then share a band (measured
  // 155 against 164 us at 16384^2 fp32 with the axes swapped).
  const int64_t r_begin = (int64_t)blockIdx.y * XR * STEP;
  const int64_t rt = r_begin + (int64_t)threadIdx.x * VN;
  int64_t roff[XR];
#pragma unroll
  for (int u = 0; u < XR; ++u) {
    const int64_t r = rt + u * STEP;
    roff[u] = r < s.R ? r : r_begin;  // R % VN == 0: wholly in or out
  }
  const int split = blockIdx.x;
  const int64_t cper = cdiv_dev(s.Cc, s.splits);
  const int64_t c_begin = split * cper;
  const int64_t c_end = c_begin + cper < s.Cc ? c_begin + cper : s.Cc;
  T acc[XR][VN];
#pragma unroll
  for (int u = 0; u < XR; ++u)
#pragma unroll
    for (int e = 0; e < VN; ++e) acc[u][e] = T(0);
  int64_t c = c_begin;
  for (; c + UC <= c_end; c += UC) {
    V a[UC][XR];
    T xv[UC];
#pragma unroll
    for (int j = 0; j < UC; ++j) {
      xv[j] = x[bidx(c + j, s.Cc, s.incx)];
      const T *col = A + (c + j) * s.ld;
#pragma unroll
      for (int u = 0; u < XR; ++u)
        a[j][u] = vld<NT>(reinterpret_cast<const V *>(col + roff[u]));
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < UC; ++j)
#pragma unroll
      for (int u = 0; u < XR; ++u)
#pragma unroll
        for (int e = 0; e < VN; ++e)
          acc[u][e] = fma_t(vlane<T>(a[j][u], e), xv[j], acc[u][e]);
  }
  for (; c < c_end; ++c) {
    const T xv = x[bidx(c, s.Cc, s.incx)];
    const T *col = A + c * s.ld;
#pragma unroll
    for (int u = 0; u < XR; ++u) {
      const V a = vld<NT>(reinterpret_cast<const V *>(col + roff[u]));
#pragma unroll
      for (int e = 0; e < VN; ++e) acc[u][e] = fma_t(vlane<T>(a, e), xv, acc[u][e]);
    }
  }
#pragma unroll
  for (int u = 0; u < XR; ++u) {
    const int64_t r = rt + u * STEP;
    if (r >= s.R) continue;
    if (s.splits == 1) {
#pragma unroll
      for (int e = 0; e < VN; ++e)
        store_y(&y[bidx(r + e, s.R, s.incy)], acc[u][e], alpha, beta);
    } else {
      V v;
#pragma unroll
      for (int e = 0; e < VN; ++e) v[e] = acc[u][e];
      *reinterpret_cast<V *>(part + (z * s.splits + split) * s.R + r) = v;
    }
  }
}

// ---------------------------------------------------------------------------
// Column kernel: y[c] = alpha * sum_r S[r,c] x[r] + beta*y[c].
// Workgroup = 4 waves x CPW columns; each wave reduces CPW columns over its
// row range with 16-byte loads, reusing each x load for the CPW columns.
// blockIdx.y selects a contiguous range of rows (split).
// ---------------------------------------------------------------------------

template <typename T, bool VEC, bool NT>
__global__ __launch_bounds__(kThreads) void gemv_column_kernel(
    GemvShape s, const T *__restrict__ A, const T *__restrict__ x, T *y,
    T alpha, T beta, T *part) {
  constexpr int VN = VEC ? VecT<T>::n : 1;
  const int64_t z = blockIdx.z;
  A += z * s.sA;
  x += z * s.sx;
  y += z * s.sy;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane_id = threadIdx.x & 63;
  const int64_t cbase = (int64_t)blockIdx.x * kColsPerBlock + wave * kCPW;
  const int split = blockIdx.y;
  // Row range of this split, a multiple of the 64*VN step.
  const int64_t step = 64 * VN;
  const int64_t rsteps = cdiv_dev(s.R, step);
  const int64_t sper = cdiv_dev(rsteps, s.splits);
  const int64_t r_begin = split * sper * step;
  int64_t r_end = r_begin + sper * step;
  if (r_end > s.R) r_end = s.R;

  T acc[kCPW][VN];
#pragma unroll
  for (int j = 0; j < kCPW; ++j)
#pragma unroll
    for (int e = 0; e < VN; ++e) acc[j][e] = T(0);

  const T *cols[kCPW];
  bool colok[kCPW];
#pragma unroll
  for (int j = 0; j < kCPW; ++j) {
    colok[j] = cbase + j < s.Cc;
    cols[j] = A + (colok[j] ? (cbase + j) : 0) * s.ld;
  }

  if (VEC) {
    using V = typename VecT<T>::type;
    const bool xunit = s.incx == 1;
    int64_t r = r_begin + lane_id * VN;
    if (xunit) {
      // Two row steps per iteration: 2 * kCPW loads of A in flight per lane.
      for (; r + step < r_end; r += 2 * step) {
        const V t0 = *reinterpret_cast<const V *>(x + r);
        const V t1 = *reinterpret_cast<const V *>(x + r + step);
        V a0[kCPW], a1[kCPW];
#pragma unroll
        for (int j = 0; j < kCPW; ++j) {
          a0[j] = vld<NT>(reinterpret_cast<const V *>(cols[j] + r));
          a1[j] = vld<NT>(reinterpret_cast<const V *>(cols[j] + r + step));
        }
#pragma unroll
        for (int j = 0; j < kCPW; ++j) {
          if (!colok[j]) continue;
#pragma unroll
          for (int e = 0; e < VN; ++e)
            acc[j][e] = fma_t(vlane<T>(a0[j], e), vlane<T>(t0, e), acc[j][e]);
#pragma unroll
          for (int e = 0; e < VN; ++e)
            acc[j][e] = fma_t(vlane<T>(a1[j], e), vlane<T>(t1, e), acc[j][e]);
        }
      }
    }
    for (; r < r_end; r += step) {
      // On this path R % VN == 0, so a vector never straddles R.
      T xv[VN];
      if (xunit) {
        const V t = *reinterpret_cast<const V *>(x + r);
#pragma unroll
        for (int e = 0; e < VN; ++e) xv[e] = vlane<T>(t, e);
      } else {
#pragma unroll
        for (int e = 0; e < VN; ++e) xv[e] = x[bidx(r + e, s.R, s.incx)];
      }
#pragma unroll
      for (int j = 0; j < kCPW; ++j) {
        if (!colok[j]) continue;
        const V a = vld<NT>(reinterpret_cast<const V *>(cols[j] + r));
#pragma unroll
        for (int e = 0; e < VN; ++e) acc[j][e] = fma_t(vlane<T>(a, e), xv[e], acc[j][e]);
      }
    }
  } else {
    for (int64_t r = r_begin + lane_id; r < r_end; r += step) {
      T xv = x[bidx(r, s.R, s.incx)];
#pragma unroll
      for (int j = 0; j < kCPW; ++j)
        if (colok[j]) acc[j][0] = fma_t(cols[j][r], xv, acc[j][0]);
    }
  }

  T sums[kCPW];
#pragma unroll
  for (int j = 0; j < kCPW; ++j) {
    T t = T(0);
#pragma unroll
    for (int e = 0; e < VN; ++e) t += acc[j][e];
    sums[j] = wave_sum(t);
  }

  if (s.splits == 1) {
    if (lane_id == 0) {
#pragma unroll
      for (int j = 0; j < kCPW; ++j)
        if (colok[j])
          store_y(&y[bidx(cbase + j, s.Cc, s.incy)], sums[j], alpha, beta);
    }
    return;
  }
  if (lane_id == 0) {
    T *pp = part + (z * s.splits + split) * s.Cc;
#pragma unroll
    for (int j = 0; j < kCPW; ++j)
      if (colok[j]) pp[cbase + j] = sums[j];
  }
}

// ---------------------------------------------------------------------------
// Band kernel: the column kernel's job for tall operands (R >= XR*256*VN).
// Workgroup = (CPB consecutive columns, one band of XR*256*VN rows). The
// band's x is loaded once into registers (thread t keeps the vectors at rows
// t*VN + u*256*VN, u < XR) and each column is then one contiguous run of the
// band read by all 256 threads, XR 16-byte loads per thread in flight with no
// x traffic. The CPB per-thread sums are reduced once at the end (wave
// shuffles, then the four waves in order). Measured on MI355X at 16384^2:
// 6.9 TB/s fp32 and 7.0 TB/s fp64, against 6.2 / 6.5 for the column kernel,
// whose four columns per wave and per-step x loads cost bandwidth.
// Bands past the first full ones (R not a multiple of the band) load from a
// clamped row (the band's first) with x = 0, so every load stays unpredicated.
// ---------------------------------------------------------------------------

template <typename T, int XR, bool NT>
__global__ __launch_bounds__(kThreads) void gemv_band_kernel(
    GemvShape s, const T *__restrict__ A, const T *__restrict__ x, T *y,
    T alpha, T beta, T *part) {
  using V = typename VecT<T>::type;
  constexpr int VN = VecT<T>::n;
  constexpr int CPB = kBandCols;
  constexpr int64_t STEP = (int64_t)kThreads * VN;
  __shared__ T red[4][CPB];
  const int64_t z = blockIdx.z;
  A += z * s.sA;
  x += z * s.sx;
  y += z * s.sy;
  const int split = blockIdx.y;
  const int64_t r_begin = (int64_t)split * XR * STEP;
  const int64_t rt = r_begin + (int64_t)threadIdx.x * VN;
  const bool full = r_begin + XR * STEP <= s.R;
  V xr[XR];
#pragma unroll
  for (int u = 0; u < XR; ++u) {
    const int64_t r = rt + u * STEP;
    if (r >= s.R) {  // R % VN == 0: a vector is wholly in or out
      xr[u] = V{};
    } else if (s.incx == 1) {
      xr[u] = *reinterpret_cast<const V *>(x + r);
    } else {
#pragma unroll
      for (int e = 0; e < VN; ++e) xr[u][e] = x[bidx(r + e, s.R, s.incx)];
    }
  }
  const int64_t c0 = (int64_t)blockIdx.x * CPB;
  T sums[CPB];
#pragma unroll
  for (int j = 0; j < CPB; ++j) {
    const int64_t c = c0 + j;
    T t = T(0);
    if (c < s.Cc) {
      const T *col = A + c * s.ld;
      V a[XR];
      if (full) {
#pragma unroll
        for (int u = 0; u < XR; ++u)
          a[u] = vld<NT>(reinterpret_cast<const V *>(col + rt + u * STEP));
      } else {
#pragma unroll
        for (int u = 0; u < XR; ++u) {
          const int64_t r = rt + u * STEP;
          a[u] = vld<NT>(
              reinterpret_cast<const V *>(col + (r < s.R ? r : r_begin)));
        }
      }
      T acc[VN];
#pragma unroll
      for (int e = 0; e < VN; ++e) acc[e] = T(0);
#pragma unroll
      for (int u = 0; u < XR; ++u)
#pragma unroll
        for (int e = 0; e < VN; ++e)
          acc[e] = fma_t(vlane<T>(a[u], e), vlane<T>(xr[u], e), acc[e]);
#pragma unroll
      for (int e = 0; e < VN; ++e) t += acc[e];
    }
    sums[j] = t;
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int j = 0; j < CPB; ++j) {
    const T v = wave_sum(sums[j]);
    if (lane == 0) red[wave][j] = v;
  }
  __syncthreads();
  if (threadIdx.x >= CPB) return;
  const int j = threadIdx.x;
  const int64_t c = c0 + j;
  if (c >= s.Cc) return;
  const T v = ((red[0][j] + red[1][j]) + red[2][j]) + red[3][j];
  if (s.splits == 1)
    store_y(&y[bidx(c, s.Cc, s.incy)], v, alpha, beta);
  else
    part[(z * s.splits + split) * s.Cc + c] = v;
}

// alpha == 0 (or an empty reduction): y = beta*y without reading A or x
// (reference gemv_alpha0_col, sgemv_array.cpp:709-734).
template <typename T>
__global__ __launch_bounds__(kThreads) void gemv_scale_y_kernel(
    int64_t len, int64_t incy, int64_t sy, T *y, T beta) {
  y += (int64_t)blockIdx.z * sy;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < len;
       i += (int64_t)gridDim.x * kThreads) {
    T *p = &y[bidx(i, len, incy)];
    *p = beta == T(0) ? T(0) : beta * *p;
  }
}

// ---------------------------------------------------------------------------
// GER: S[r,c] = fma(v[r], alpha*w[c], S[r,c]).
// Each thread owns VN consecutive rows (one 16-byte vector) and walks kGerCols
// columns, so v stays in registers and every store is a full 16-byte store.
// ---------------------------------------------------------------------------
constexpr int kGerCols = 16;

template <typename T, bool VEC, bool NT>
__global__ __launch_bounds__(kThreads) void ger_kernel(
    int64_t R, int64_t Cc, T *A, int64_t ld, int64_t sA,
    const T *__restrict__ v, int64_t incv, int64_t sv,
    const T *__restrict__ w, int64_t incw, int64_t sw, T alpha) {
  constexpr int VN = VEC ? VecT<T>::n : 1;
  const int64_t z = blockIdx.z;
  A += z * sA;
  v += z * sv;
  w += z * sw;
  const int64_t r0 = ((int64_t)blockIdx.x * kThreads + threadIdx.x) * VN;
  if (r0 >= R) return;
  const int64_t c0 = (int64_t)blockIdx.y * kGerCols;
  T vr[VN];
#pragma unroll
  for (int e = 0; e < VN; ++e)
    vr[e] = (r0 + e < R) ? v[bidx(r0 + e, R, incv)] : T(0);
  const int64_t cend = c0 + kGerCols < Cc ? c0 + kGerCols : Cc;
  if (VEC) {
    using V = typename VecT<T>::type;
    if (cend - c0 == kGerCols) {
      // Every load is issued before the first store, so kGerCols loads are
      // in flight per lane (the stores would otherwise order each next load
      // behind them: the compiler cannot prove the columns disjoint).
      V a[kGerCols];
#pragma unroll
      for (int j = 0; j < kGerCols; ++j) {
        const V *p = reinterpret_cast<const V *>(A + r0 + (c0 + j) * ld);
        if constexpr (NT)
          a[j] = __builtin_nontemporal_load(p);
        else
          a[j] = *p;
      }
#pragma unroll
      for (int j = 0; j < kGerCols; ++j) {
        const T t = alpha * w[bidx(c0 + j, Cc, incw)];
#pragma unroll
        for (int e = 0; e < VN; ++e) a[j][e] = fma_t(vr[e], t, a[j][e]);
        V *p = reinterpret_cast<V *>(A + r0 + (c0 + j) * ld);
        if constexpr (NT)
          __builtin_nontemporal_store(a[j], p);
        else
          *p = a[j];
      }
      return;
    }
    for (int64_t c = c0; c < cend; ++c) {
      const T t = alpha * w[bidx(c, Cc, incw)];
      V *p = reinterpret_cast<V *>(A + r0 + c * ld);
      V a = *p;
#pragma unroll
      for (int e = 0; e < VN; ++e) a[e] = fma_t(vr[e], t, a[e]);
      *p = a;
    }
  } else {
    for (int64_t c = c0; c < cend; ++c) {
      const T t = alpha * w[bidx(c, Cc, incw)];
      T *p = A + r0 + c * ld;
      *p = fma_t(vr[0], t, *p);
    }
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
template <typename T>
hcblasStatus_t run_gemv(hcblasHandle_t h, bool panel, int64_t R, int64_t Cc,
                        T alpha, const T *A, int64_t ld, const T *x,
                        int64_t incx, T beta, T *y, int64_t incy, int64_t sA,
                        int64_t sx, int64_t sy, int64_t batch,
                        void *fixed_part = nullptr, size_t fixed_bytes = 0) {
  const int64_t len_y = panel ? R : Cc;
  const int64_t len_x = panel ? Cc : R;
  if (len_y == 0 || batch == 0) return HCBLAS_STATUS_SUCCESS;
  if (alpha == T(0) || len_x == 0) {
    if (beta == T(1)) return HCBLAS_STATUS_SUCCESS;
    dim3 g((unsigned)std::min<int64_t>(cdiv(len_y, kThreads), 4096), 1,
           (unsigned)batch);
    gemv_scale_y_kernel<T><<<g, kThreads, 0, h->stream>>>(len_y, incy, sy, y,
                                                          beta);
    return launch_status();
  }
  constexpr int VN = VecT<T>::n;
  // 16-byte path: aligned base, leading dimension and rows a multiple of VN,
  // and batch strides that keep every problem's A (and x, where the column
  // kernels read it as vectors) 16-byte aligned.
  const bool xvec = !panel && incx == 1;
  const bool vec = aligned16(A) && (ld % VN == 0) && (R % VN == 0) &&
                   (!xvec || aligned16(x)) &&
                   (batch == 1 || (sA % VN == 0 && (!xvec || sx % VN == 0)));
  const bool nt = (int64_t)sizeof(T) * R * Cc * batch >= kStreamBytes;
  GemvShape s{R, Cc, ld, incx, incy, sA, sx, sy, 1};
  const GemvPlan plan =
      plan_gemv(h->num_cus, panel, R, Cc, vec, VN, sizeof(T), batch);
  s.splits = plan.splits;
  const int64_t oblocks = plan.oblocks;
  const bool bandn = plan.kernel == GemvKernel::kBandN;
  const int band_xr = plan.kernel == GemvKernel::kBand16  ? 16
                      : plan.kernel == GemvKernel::kBand8 ? 8
                                                          : 0;
  T *part = nullptr;
  if (s.splits > 1) {
    if (fixed_part) {
      // Allocation-free entry: the same split (so the same summation order
      // and results as the workspace path) or an error, never fewer splits.
      if (plan.partial_bytes > fixed_bytes) return HCBLAS_STATUS_INVALID_VALUE;
      part = static_cast<T *>(fixed_part);
    } else {
      if (ensure_workspace(h->gemv_partials, plan.partial_bytes) != hipSuccess)
        return HCBLAS_STATUS_ALLOC_FAILED;
      part = static_cast<T *>(h->gemv_partials.ptr);
    }
  }
  for (int64_t b0 = 0; b0 < batch; b0 += 65535) {
    int64_t nb = std::min<int64_t>(65535, batch - b0);
    dim3 grid((unsigned)oblocks, (unsigned)s.splits, (unsigned)nb);
    if (bandn) grid = dim3((unsigned)s.splits, (unsigned)oblocks, (unsigned)nb);
    const T *Ab = A + b0 * sA;
    const T *xb = x + b0 * sx;
    T *yb = y + b0 * sy;
    if (bandn) {
      if (nt)
        gemv_bandn_kernel<T, true><<<grid, kThreads, 0, h->stream>>>(
            s, Ab, xb, yb, alpha, beta, part);
      else
        gemv_bandn_kernel<T, false><<<grid, kThreads, 0, h->stream>>>(
            s, Ab, xb, yb, alpha, beta, part);
    } else if (panel) {
      if (vec && nt)
        gemv_panel_kernel<T, true, true><<<grid, kThreads, 0, h->stream>>>(
            s, Ab, xb, yb, alpha, beta, part);
      else if (vec)
        gemv_panel_kernel<T, true, false><<<grid, kThreads, 0, h->stream>>>(
            s, Ab, xb, yb, alpha, beta, part);
      else
        gemv_panel_kernel<T, false, false><<<grid, kThreads, 0, h->stream>>>(
            s, Ab, xb, yb, alpha, beta, part);
    } else if (band_xr == 16) {
      if (nt)
        gemv_band_kernel<T, 16, true><<<grid, kThreads, 0, h->stream>>>(
            s, Ab, xb, yb, alpha, beta, part);
      else
        gemv_band_kernel<T, 16, false><<<grid, kThreads, 0, h->stream>>>(
            s, Ab, xb, yb, alpha, beta, part);
    } else if (band_xr == 8) {
      if (nt)
        gemv_band_kernel<T, 8, true><<<grid, kThreads, 0, h->stream>>>(
            s, Ab, xb, yb, alpha, beta, part);
      else
        gemv_band_kernel<T, 8, false><<<grid, kThreads, 0, h->stream>>>(
            s, Ab, xb, yb, alpha, beta, part);
    } else {
      if (vec && nt)
        gemv_column_kernel<T, true, true><<<grid, kThreads, 0, h->stream>>>(
            s, Ab, xb, yb, alpha, beta, part);
      else if (vec)
        gemv_column_kernel<T, true, false><<<grid, kThreads, 0, h->stream>>>(
            s, Ab, xb, yb, alpha, beta, part);
      else
        gemv_column_kernel<T, false, false><<<grid, kThreads, 0, h->stream>>>(
            s, Ab, xb, yb, alpha, beta, part);
    }
    if (s.splits > 1) {
      dim3 fg((unsigned)std::min<int64_t>(cdiv(len_y, kThreads), 1024), 1,
              (unsigned)nb);
      gemv_finish_kernel<T><<<fg, kThreads, 0, h->stream>>>(
          len_y, s.splits, part, (int64_t)s.splits * len_y, yb, incy, sy,
          alpha, beta);
    }
    hcblasStatus_t st = launch_status();
    if (st != HCBLAS_STATUS_SUCCESS) return st;
  }
  return HCBLAS_STATUS_SUCCESS;
}

// Common argument handling for gemv in the reference's (order, trans) terms.
template <typename T>
hcblasStatus_t gemv_entry(hcblasHandle_t h, hcblasOrder_t order,
                          hcblasOperation_t trans, int M, int N,
                          const T *alpha, const T *A, int64_t aOffset,
                          int64_t lda, const T *x, int64_t xOffset,
                          int64_t incx, const T *beta, T *y, int64_t yOffset,
                          int64_t incy, int64_t batch, bool ref_batch_strides,
                          int64_t sA_in = 0, int64_t sx_in = 0,
                          int64_t sy_in = 0, bool noalloc = false) {
  if (!handle_ok(h)) return HCBLAS_STATUS_NOT_INITIALIZED;
  if (M < 0 || N < 0 || incx == 0 || incy == 0 || batch < 0)
    return HCBLAS_STATUS_INVALID_VALUE;
  if (order != HCBLAS_ROW_MAJOR && order != HCBLAS_COL_MAJOR)
    return HCBLAS_STATUS_INVALID_VALUE;
  if (canon_op(trans) < 0) return HCBLAS_STATUS_INVALID_VALUE;
  const bool notrans = canon_op(trans) == HCBLAS_OP_N;
  const int64_t len_x = notrans ? N : M, len_y = notrans ? M : N;
  // Storage is column-major R x Cc.
  const bool col = order == HCBLAS_COL_MAJOR;
  const int64_t R = col ? M : N, Cc = col ? N : M;
  if (lda < std::max<int64_t>(1, R)) return HCBLAS_STATUS_INVALID_VALUE;
  if (M == 0 || N == 0 || batch == 0) {
    if (len_y == 0 || batch == 0) return HCBLAS_STATUS_SUCCESS;
  }
  if (alpha == nullptr || beta == nullptr || y == nullptr)
    return HCBLAS_STATUS_INVALID_VALUE;
  if (*alpha != T(0) && len_x > 0 && (A == nullptr || x == nullptr))
    return HCBLAS_STATUS_INVALID_VALUE;
  // Batch strides as in hcblas.cpp:1008-1019 (x by len(x), y by len(y), A by
  // M*N), or the caller's explicit batch offsets.
  int64_t sA = ref_batch_strides ? (int64_t)M * N : sA_in;
  int64_t sx = ref_batch_strides ? len_x : sx_in;
  int64_t sy = ref_batch_strides ? len_y : sy_in;
  const bool panel = col == notrans;
  return run_gemv<T>(h, panel, R, Cc, *alpha, A ? A + aOffset : nullptr, lda,
                     x ? x + xOffset : nullptr, incx, *beta, y + yOffset, incy,
                     sA, sx, sy, batch, noalloc ? h->noalloc_scratch : nullptr,
                     noalloc ? kNoAllocScratchBytes : 0);
}

template <typename T>
hcblasStatus_t ger_entry(hcblasHandle_t h, hcblasOrder_t order, int M, int N,
                         const T *alpha, const T *x, int64_t xOffset,
                         int64_t incx, const T *y, int64_t yOffset,
                         int64_t incy, T *A, int64_t aOffset, int64_t lda,
                         int64_t batch, bool ref_batch_strides = true,
                         int64_t sA_in = 0, int64_t sx_in = 0,
                         int64_t sy_in = 0) {
  if (!handle_ok(h)) return HCBLAS_STATUS_NOT_INITIALIZED;
  if (M < 0 || N < 0 || incx == 0 || incy == 0 || batch < 0)
    return HCBLAS_STATUS_INVALID_VALUE;
  if (order != HCBLAS_ROW_MAJOR && order != HCBLAS_COL_MAJOR)
    return HCBLAS_STATUS_INVALID_VALUE;
  const bool col = order == HCBLAS_COL_MAJOR;
  const int64_t R = col ? M : N, Cc = col ? N : M;
  if (lda < std::max<int64_t>(1, R)) return HCBLAS_STATUS_INVALID_VALUE;
  if (M == 0 || N == 0 || batch == 0) return HCBLAS_STATUS_SUCCESS;
  if (alpha == nullptr || x == nullptr || y == nullptr || A == nullptr)
    return HCBLAS_STATUS_INVALID_VALUE;
  if (*alpha == T(0)) return HCBLAS_STATUS_SUCCESS;  // netlib quick return
  // Batch strides as in hcblas.cpp:1160-1167 (x by M, y by N, A by M*N), or
  // the caller's explicit batch offsets.
  const bool multi = batch > 1;
  const int64_t sA = !multi ? 0 : ref_batch_strides ? (int64_t)M * N : sA_in;
  const int64_t sxv = !multi ? 0 : ref_batch_strides ? M : sx_in;
  const int64_t syv = !multi ? 0 : ref_batch_strides ? N : sy_in;
  const T *v = col ? x + xOffset : y + yOffset;
  const T *w = col ? y + yOffset : x + xOffset;
  const int64_t incv = col ? incx : incy, incw = col ? incy : incx;
  const int64_t sv = col ? sxv : syv, sw = col ? syv : sxv;
  T *Ab = A + aOffset;
  constexpr int VN = VecT<T>::n;
  const bool vec = aligned16(Ab) && (lda % VN == 0) && (R % VN == 0) &&
                   (batch == 1 || sA % VN == 0);
  const int64_t rows_per_block = (int64_t)kThreads * (vec ? VN : 1);
  // A matrix larger than the 256 MB last-level cache (MALL) streams through
  // with non-temporal accesses: caching it cannot keep it for the next call
  // and only adds write-back traffic (DGER 8192^2: 6.09 against 5.6 TB/s).
  // One that fits keeps the cached path, which the next call re-reads from
  // the MALL (SGER 8192^2: 6.87 cached against 6.10 non-temporal).
  const bool nt =
      (int64_t)lda * Cc * batch * (int64_t)sizeof(T) > (256ll << 20);
  for (int64_t b0 = 0; b0 < batch; b0 += 65535) {
    int64_t nb = std::min<int64_t>(65535, batch - b0);
    dim3 grid((unsigned)cdiv(R, rows_per_block),
              (unsigned)cdiv(Cc, kGerCols), (unsigned)nb);
    if (vec && nt)
      ger_kernel<T, true, true><<<grid, kThreads, 0, h->stream>>>(
          R, Cc, Ab + b0 * sA, lda, sA, v + b0 * sv, incv, sv, w + b0 * sw,
          incw, sw, *alpha);
    else if (vec)
      ger_kernel<T, true, false><<<grid, kThreads, 0, h->stream>>>(
          R, Cc, Ab + b0 * sA, lda, sA, v + b0 * sv, incv, sv, w + b0 * sw,
          incw, sw, *alpha);
    else
      ger_kernel<T, false, false><<<grid, kThreads, 0, h->stream>>>(
          R, Cc, Ab + b0 * sA, lda, sA, v + b0 * sv, incv, sv, w + b0 * sw,
          incw, sw, *alpha);
    hcblasStatus_t st = launch_status();
    if (st != HCBLAS_STATUS_SUCCESS) return st;
  }
  return HCBLAS_STATUS_SUCCESS;
}

// Adjoint of A_out = alpha * x * y^T + A_in (no reference counterpart): for
// a cotangent dA (M x N in the given order) the GEMV pair
//   dx = alpha * dA * y,   dy = alpha * dA^T * x
// (dA_in = dA is the identity and needs no work). dx / dy are written with
// unit stride; either may be NULL.
template <typename T>
hcblasStatus_t ger_backward_entry(hcblasHandle_t h, hcblasOrder_t order,
                                         int M, int N, const T *alpha,
                                         const T *x, int64_t incx, const T *y,
                                         int64_t incy, const T *dA,
                                         int64_t ldda, T *dx, T *dy) {
  if (!handle_ok(h)) return HCBLAS_STATUS_NOT_INITIALIZED;
  const T zero = T(0);
  hcblasStatus_t st = HCBLAS_STATUS_SUCCESS;
  if (dx) {
    st = gemv_entry<T>(h, order, HCBLAS_OP_N, M, N, alpha, dA, 0, ldda, y, 0,
                       incy, &zero, dx, 0, 1, 1, false);
    if (st != HCBLAS_STATUS_SUCCESS) return st;
  }
  if (dy)
    st = gemv_entry<T>(h, order, HCBLAS_OP_T, M, N, alpha, dA, 0, ldda, x, 0,
                       incx, &zero, dy, 0, 1, 1, false);
  return st;
}

}  // namespace
}  // namespace hcb

using namespace hcb;

extern "C" {

hcblasStatus_t hcblasSgemv(hcblasHandle_t handle, hcblasOperation_t trans,
                           int m, int n, const float *alpha, const float *A,
                           int lda, const float *x, int incx, const float *beta,
                           float *y, int incy) {
  return gemv_entry<float>(handle, HCBLAS_COL_MAJOR, trans, m, n, alpha, A, 0,
                           lda, x, 0, incx, beta, y, 0, incy, 1, false);
}
hcblasStatus_t hcblasDgemv(hcblasHandle_t handle, hcblasOperation_t trans,
                           int m, int n, const double *alpha, const double *A,
                           int lda, const double *x, int incx,
                           const double *beta, double *y, int incy) {
  return gemv_entry<double>(handle, HCBLAS_COL_MAJOR, trans, m, n, alpha, A, 0,
                            lda, x, 0, incx, beta, y, 0, incy, 1, false);
}
hcblasStatus_t hcblasSgemvBatched(hcblasHandle_t handle,
                                  hcblasOperation_t trans, int m, int n,
                                  const float *alpha, const float *A, int lda,
                                  const float *x, int incx, const float *beta,
                                  float *y, int incy, int batchCount) {
  return gemv_entry<float>(handle, HCBLAS_COL_MAJOR, trans, m, n, alpha, A, 0,
                           lda, x, 0, incx, beta, y, 0, incy, batchCount, true);
}
hcblasStatus_t hcblasDgemvBatched(hcblasHandle_t handle,
                                  hcblasOperation_t trans, int m, int n,
                                  const double *alpha, const double *A,
                                  int lda, const double *x, int incx,
                                  const double *beta, double *y, int incy,
                                  int batchCount) {
  return gemv_entry<double>(handle, HCBLAS_COL_MAJOR, trans, m, n, alpha, A, 0,
                            lda, x, 0, incx, beta, y, 0, incy, batchCount,
                            true);
}
hcblasStatus_t hcblas_sgemv(hcblasHandle_t handle, hcblasOrder_t order,
                            hcblasOperation_t type, int M, int N,
                            const float *alpha, const float *A,
                            int64_t aOffset, int64_t lda, const float *X,
                            int64_t xOffset, int64_t incX, const float *beta,
                            float *Y, int64_t yOffset, int64_t incY) {
  return gemv_entry<float>(handle, order, type, M, N, alpha, A, aOffset, lda,
                           X, xOffset, incX, beta, Y, yOffset, incY, 1, false);
}
hcblasStatus_t hcblas_dgemv(hcblasHandle_t handle, hcblasOrder_t order,
                            hcblasOperation_t type, int M, int N,
                            const double *alpha, const double *A,
                            int64_t aOffset, int64_t lda, const double *X,
                            int64_t xOffset, int64_t incX, const double *beta,
                            double *Y, int64_t yOffset, int64_t incY) {
  return gemv_entry<double>(handle, order, type, M, N, alpha, A, aOffset, lda,
                            X, xOffset, incX, beta, Y, yOffset, incY, 1, false);
}

hcblasStatus_t hcblasSger(hcblasHandle_t handle, int m, int n,
                          const float *alpha, const float *x, int incx,
                          const float *y, int incy, float *A, int lda) {
  return ger_entry<float>(handle, HCBLAS_COL_MAJOR, m, n, alpha, x, 0, incx, y,
                          0, incy, A, 0, lda, 1);
}
hcblasStatus_t hcblasDger(hcblasHandle_t handle, int m, int n,
                          const double *alpha, const double *x, int incx,
                          const double *y, int incy, double *A, int lda) {
  return ger_entry<double>(handle, HCBLAS_COL_MAJOR, m, n, alpha, x, 0, incx, y,
                           0, incy, A, 0, lda, 1);
}
hcblasStatus_t hcblasSgerBatched(hcblasHandle_t handle, int m, int n,
                                 const float *alpha, const float *x, int incx,
                                 const float *y, int incy, float *A, int lda,
                                 int batchCount) {
  return ger_entry<float>(handle, HCBLAS_COL_MAJOR, m, n, alpha, x, 0, incx, y,
                          0, incy, A, 0, lda, batchCount);
}
hcblasStatus_t hcblasDgerBatched(hcblasHandle_t handle, int m, int n,
                                 const double *alpha, const double *x,
                                 int incx, const double *y, int incy,
                                 double *A, int lda, int batchCount) {
  return ger_entry<double>(handle, HCBLAS_COL_MAJOR, m, n, alpha, x, 0, incx,
                           y, 0, incy, A, 0, lda, batchCount);
}
hcblasStatus_t hcblas_sger(hcblasHandle_t handle, hcblasOrder_t order, int M,
                           int N, const float *alpha, const float *X,
                           int64_t xOffset, int64_t incX, const float *Y,
                           int64_t yOffset, int64_t incY, float *A,
                           int64_t aOffset, int64_t lda) {
  return ger_entry<float>(handle, order, M, N, alpha, X, xOffset, incX, Y,
                          yOffset, incY, A, aOffset, lda, 1);
}
hcblasStatus_t hcblas_dger(hcblasHandle_t handle, hcblasOrder_t order, int M,
                           int N, const double *alpha, const double *X,
                           int64_t xOffset, int64_t incX, const double *Y,
                           int64_t yOffset, int64_t incY, double *A,
                           int64_t aOffset, int64_t lda) {
  return ger_entry<double>(handle, order, M, N, alpha, X, xOffset, incX, Y,
                           yOffset, incY, A, aOffset, lda, 1);
}

hcblasStatus_t hcblas_sgemv_batched(
    hcblasHandle_t handle, hcblasOrder_t order, hcblasOperation_t type, int M,
    int N, const float *alpha, const float *A, int64_t aOffset,
    int64_t A_batchOffset, int64_t lda, const float *X, int64_t xOffset,
    int64_t X_batchOffset, int64_t incX, const float *beta, float *Y,
    int64_t yOffset, int64_t Y_batchOffset, int64_t incY, int batchSize) {
  return gemv_entry<float>(handle, order, type, M, N, alpha, A, aOffset, lda,
                           X, xOffset, incX, beta, Y, yOffset, incY, batchSize,
                           false, A_batchOffset, X_batchOffset, Y_batchOffset);
}
hcblasStatus_t hcblas_dgemv_batched(
    hcblasHandle_t handle, hcblasOrder_t order, hcblasOperation_t type, int M,
    int N, const double *alpha, const double *A, int64_t aOffset,
    int64_t A_batchOffset, int64_t lda, const double *X, int64_t xOffset,
    int64_t X_batchOffset, int64_t incX, const double *beta, double *Y,
    int64_t yOffset, int64_t Y_batchOffset, int64_t incY, int batchSize) {
  return gemv_entry<double>(handle, order, type, M, N, alpha, A, aOffset, lda,
                            X, xOffset, incX, beta, Y, yOffset, incY,
                            batchSize, false, A_batchOffset, X_batchOffset,
                            Y_batchOffset);
}
hcblasStatus_t hcblas_sger_batched(
    hcblasHandle_t handle, hcblasOrder_t order, int M, int N,
    const float *alpha, const float *X, int64_t xOffset,
    int64_t X_batchOffset, int64_t incX, const float *Y, int64_t yOffset,
    int64_t Y_batchOffset, int64_t incY, float *A, int64_t aOffset,
    int64_t A_batchOffset, int64_t lda, int batchSize) {
  return ger_entry<float>(handle, order, M, N, alpha, X, xOffset, incX, Y,
                          yOffset, incY, A, aOffset, lda, batchSize, false,
                          A_batchOffset, X_batchOffset, Y_batchOffset);
}
hcblasStatus_t hcblas_dger_batched(
    hcblasHandle_t handle, hcblasOrder_t order, int M, int N,
    const double *alpha, const double *X, int64_t xOffset,
    int64_t X_batchOffset, int64_t incX, const double *Y, int64_t yOffset,
    int64_t Y_batchOffset, int64_t incY, double *A, int64_t aOffset,
    int64_t A_batchOffset, int64_t lda, int batchSize) {
  return ger_entry<double>(handle, order, M, N, alpha, X, xOffset, incX, Y,
                           yOffset, incY, A, aOffset, lda, batchSize, false,
                           A_batchOffset, X_batchOffset, Y_batchOffset);
}

hcblasStatus_t hcblas_gemv_noalloc(hcblasHandle_t handle,
                                   hcblasDatatype_t dtype, hcblasOrder_t order,
                                   hcblasOperation_t type, int M, int N,
                                   const void *alpha, const void *A,
                                   int64_t aOffset, int64_t lda, const void *X,
                                   int64_t xOffset, int64_t incX,
                                   const void *beta, void *Y, int64_t yOffset,
                                   int64_t incY) {
  if (!handle_ok(handle)) return HCBLAS_STATUS_NOT_INITIALIZED;
  if (dtype == HCBLAS_R_32F)
    return gemv_entry<float>(
        handle, order, type, M, N, static_cast<const float *>(alpha),
        static_cast<const float *>(A), aOffset, lda,
        static_cast<const float *>(X), xOffset, incX,
        static_cast<const float *>(beta), static_cast<float *>(Y), yOffset,
        incY, 1, false, 0, 0, 0, true);
  if (dtype == HCBLAS_R_64F)
    return gemv_entry<double>(
        handle, order, type, M, N, static_cast<const double *>(alpha),
        static_cast<const double *>(A), aOffset, lda,
        static_cast<const double *>(X), xOffset, incX,
        static_cast<const double *>(beta), static_cast<double *>(Y), yOffset,
        incY, 1, false, 0, 0, 0, true);
  return HCBLAS_STATUS_INVALID_VALUE;
}

hcblasStatus_t hcblas_sger_backward(hcblasHandle_t handle,
                                    hcblasOrder_t order, int M, int N,
                                    const float *alpha, const float *X,
                                    int64_t incX, const float *Y, int64_t incY,
                                    const float *dA, int64_t ldda, float *dX,
                                    float *dY) {
  return ger_backward_entry<float>(handle, order, M, N, alpha, X, incX, Y,
                                   incY, dA, ldda, dX, dY);
}
hcblasStatus_t hcblas_dger_backward(hcblasHandle_t handle,
                                    hcblasOrder_t order, int M, int N,
                                    const double *alpha, const double *X,
                                    int64_t incX, const double *Y,
                                    int64_t incY, const double *dA,
                                    int64_t ldda, double *dX, double *dY) {
  return ger_backward_entry<double>(handle, order, M, N, alpha, X, incX, Y,
                                    incY, dA, ldda, dX, dY);
}

}  // extern "C"
