// GEMM entry points: S/D/H/C/Z gemm, pointer-array and strided batches, with
// explicit storage order and per-matrix offsets.
//
// Reference path replaced (rocmarchive/hcBLAS):
//   hcblasSgemm ... hcblasZgemmBatched   lib/src/hcblas.cpp:1304-1675
//   Hcblaslibrary::hcblas_sgemm          lib/src/blas/sgemm/sgemm_array_wrapper.cpp:296-364
//   gemm_HC order/trans dispatch         sgemm_array_wrapper.cpp:27-154
//   gemm_alpha0_col / _row               sgemm_array_wrapper.cpp:157-293
//   cgemm / zgemm wrappers               lib/src/blas/cgemm/cgemm_array_wrapper.cpp:212-371
//
// Every call is reduced to ONE column-major form (row-major C = op(A)op(B)
// is column-major C^T = op(B)^T op(A)^T, i.e. an operand swap), so all eight
// order x transA x transB cases run the same MFMA kernel family; only the
// operand orientations (k-contiguous or not) differ.
//
// Complex GEMM runs on the real MFMA kernels through an exact real embedding:
// with a = alpha*op(A)(m,k) the 2M x 2K real matrix
//     [ re a  -im a ]
//     [ im a   re a ]
// times the 2K x N real view of op(B) (rows alternate re/im) gives the 2M x N
// real view of the interleaved complex product, i.e. C itself with
// ldc_real = 2*ldc. This is the 4-real-product method (8MNK flops); alpha is
// folded into the A embedding (netlib folds it into B: temp = alpha*B(l,j)).
#include <cstring>

#include "gemm_common.hpp"
#include "gemm_launch.hpp"

namespace hcb {

namespace {


// Checks dimensions and leading dimensions (CBLAS rules) and builds the
// column-major canonical problem. Returns SUCCESS or INVALID_VALUE.
hcblasStatus_t canonicalize(hcblasOrder_t order, hcblasOperation_t ta,
                            hcblasOperation_t tb, int M, int N, int K,
                            const void *A, int64_t lda, int64_t aOff,
                            int64_t sA, const void *const *Ap, const void *B,
                            int64_t ldb, int64_t bOff, int64_t sB,
                            const void *const *Bp, int64_t ldc, Canon &c) {
  if (M < 0 || N < 0 || K < 0) return HCBLAS_STATUS_INVALID_VALUE;
  if (canon_op(ta) < 0 || canon_op(tb) < 0) return HCBLAS_STATUS_INVALID_VALUE;
  ta = static_cast<hcblasOperation_t>(canon_op(ta));
  tb = static_cast<hcblasOperation_t>(canon_op(tb));
  if (order != HCBLAS_ROW_MAJOR && order != HCBLAS_COL_MAJOR)
    return HCBLAS_STATUS_INVALID_VALUE;
  const bool col = order == HCBLAS_COL_MAJOR;
  const bool an = ta == HCBLAS_OP_N, bn = tb == HCBLAS_OP_N;
  // Rows of the stored arrays (the dimension lda must cover).
  int64_t a_rows, b_rows, c_rows;
  if (col) {
    a_rows = an ? M : K;
    b_rows = bn ? K : N;
    c_rows = M;
  } else {
    a_rows = an ? K : M;
    b_rows = bn ? N : K;
    c_rows = N;
  }
  if (lda < std::max<int64_t>(1, a_rows) ||
      ldb < std::max<int64_t>(1, b_rows) ||
      ldc < std::max<int64_t>(1, c_rows))
    return HCBLAS_STATUS_INVALID_VALUE;
  if (col) {
    c.M = M;
    c.N = N;
    c.K = K;
    c.a_k = !an;
    c.b_k = bn;
    c.conj_a = ta == HCBLAS_OP_C;
    c.conj_b = tb == HCBLAS_OP_C;
    c.A = A;
    c.lda = lda;
    c.aOff = aOff;
    c.strideA = sA;
    c.Aptrs = Ap;
    c.B = B;
    c.ldb = ldb;
    c.bOff = bOff;
    c.strideB = sB;
    c.Bptrs = Bp;
  } else {
    c.M = N;
    c.N = M;
    c.K = K;
    c.a_k = !bn;
    c.b_k = an;
    c.conj_a = tb == HCBLAS_OP_C;
    c.conj_b = ta == HCBLAS_OP_C;
    c.A = B;
    c.lda = ldb;
    c.aOff = bOff;
    c.strideA = sB;
    c.Aptrs = Bp;
    c.B = A;
    c.ldb = lda;
    c.bOff = aOff;
    c.strideB = sA;
    c.Bptrs = Ap;
  }
  return HCBLAS_STATUS_SUCCESS;
}

// ---------------------------------------------------------------------------
// C = beta*C (alpha == 0 or K == 0), beta == 0 writes zeros without reading
// (reference gemm_alpha0_col, sgemm_array_wrapper.cpp:157-210, but NaN in C
// propagates through beta != 0 as in CBLAS).
// ---------------------------------------------------------------------------
template <typename OutT, typename S>
__global__ __launch_bounds__(256) void scale_c_kernel(int M, int N, OutT *C,
                                                      int64_t ldc,
                                                      int64_t strideC,
                                                      void *const *Cptrs,
                                                      int64_t cOff, S beta) {
  OutT *Cb = Cptrs ? static_cast<OutT *>(Cptrs[blockIdx.z]) + cOff
                   : C + cOff + (int64_t)blockIdx.z * strideC;
  const int64_t total = (int64_t)M * N;
  for (int64_t f = (int64_t)blockIdx.x * 256 + threadIdx.x; f < total;
       f += (int64_t)gridDim.x * 256) {
    const int64_t j = f / M, i = f - j * M;
    OutT *p = Cb + j * ldc + i;
    if (beta == S(0))
      *p = to_out<OutT>(S(0));
    else
      *p = to_out<OutT>(beta * from_out<S>(*p));
  }
}

// Complex beta scaling of an interleaved M x N matrix (products rounded
// separately, as in complex scal).
template <typename T>
__global__ __launch_bounds__(256) void scale_c_complex_kernel(
    int M, int N, T *C, int64_t ldc, int64_t strideC, void *const *Cptrs,
    int64_t cOff, T br, T bi) {
#pragma clang fp contract(off)
  T *Cb = Cptrs ? static_cast<T *>(Cptrs[blockIdx.z]) + 2 * cOff
                : C + 2 * (cOff + (int64_t)blockIdx.z * strideC);
  const int64_t total = (int64_t)M * N;
  for (int64_t f = (int64_t)blockIdx.x * 256 + threadIdx.x; f < total;
       f += (int64_t)gridDim.x * 256) {
    const int64_t j = f / M, i = f - j * M;
    T *p = Cb + 2 * (j * ldc + i);
    if (br == T(0) && bi == T(0)) {
      p[0] = T(0);
      p[1] = T(0);
    } else {
      T xr = p[0], xi = p[1];
      T pr1 = br * xr, pr2 = bi * xi, pi1 = br * xi, pi2 = bi * xr;
      p[0] = pr1 - pr2;
      p[1] = pi1 + pi2;
    }
  }
}

// Complex embedding of A' (M x K, orientation a_k, conj) with alpha folded in:
// writes the k-contiguous real matrix Ar (2M rows, row stride ldr >= 2K).
// 32x32 complex tiles are staged through LDS so both the read (along the
// source's contiguous dimension) and the write (along k) are coalesced.
template <typename T>
__global__ __launch_bounds__(256) void complex_embed_a_kernel(
    int M, int K, const T *A, int64_t lda, int64_t aOff, int64_t strideA,
    const void *const *Aptrs, bool a_k, bool conj, T ar, T ai, T *Ar,
    int64_t ldr, int64_t strideR) {
#pragma clang fp contract(off)
  __shared__ T tre[32][33], tim[32][33];
  const int z = blockIdx.z;
  const T *Ab = Aptrs ? static_cast<const T *>(Aptrs[z]) + 2 * aOff
                      : A + 2 * (aOff + (int64_t)z * strideA);
  T *Rb = Ar + (int64_t)z * strideR;
  const int m0 = blockIdx.y * 32, k0 = blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  // Load: tile index (m, k) kept as tre[m - m0][k - k0].
  for (int r = ty; r < 32; r += 8) {
    int m, k;
    if (a_k) {
      m = m0 + r;
      k = k0 + tx;  // contiguous along k
    } else {
      m = m0 + tx;
      k = k0 + r;  // contiguous along m
    }
    T re = T(0), im = T(0);
    if (m < M && k < K) {
      const T *src = a_k ? Ab + 2 * ((int64_t)m * lda + k)
                         : Ab + 2 * ((int64_t)k * lda + m);
      re = src[0];
      im = conj ? -src[1] : src[1];
    }
    // alpha * a, products rounded separately.
    T pr1 = ar * re, pr2 = ai * im, pi1 = ar * im, pi2 = ai * re;
    const int mm = a_k ? r : tx, kk = a_k ? tx : r;
    tre[mm][kk] = pr1 - pr2;
    tim[mm][kk] = pi1 + pi2;
  }
  __syncthreads();
  // Store rows 2m and 2m+1, columns 2k and 2k+1, contiguous along k.
  for (int r = ty; r < 32; r += 8) {
    const int m = m0 + r, k = k0 + tx;
    if (m >= M || k >= K) continue;
    const T re = tre[r][tx], im = tim[r][tx];
    T *row0 = Rb + (int64_t)(2 * m) * ldr + 2 * k;
    T *row1 = row0 + ldr;
    row0[0] = re;
    row0[1] = -im;
    row1[0] = im;
    row1[1] = re;
  }
}

// Real view of op(B) packed k-contiguous: Br (N rows, row stride ldr >= 2K),
// Br[n][2k + c] = component c of B'(k, n) (conjugated if requested).
template <typename T>
__global__ __launch_bounds__(256) void complex_pack_b_kernel(
    int K, int N, const T *B, int64_t ldb, int64_t bOff, int64_t strideB,
    const void *const *Bptrs, bool b_k, bool conj, T *Br, int64_t ldr,
    int64_t strideR) {
  __shared__ T tre[32][33], tim[32][33];
  const int z = blockIdx.z;
  const T *Bb = Bptrs ? static_cast<const T *>(Bptrs[z]) + 2 * bOff
                      : B + 2 * (bOff + (int64_t)z * strideB);
  T *Rb = Br + (int64_t)z * strideR;
  const int n0 = blockIdx.y * 32, k0 = blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int r = ty; r < 32; r += 8) {
    int n, k;
    if (b_k) {
      n = n0 + r;
      k = k0 + tx;
    } else {
      n = n0 + tx;
      k = k0 + r;
    }
    T re = T(0), im = T(0);
    if (n < N && k < K) {
      const T *src = b_k ? Bb + 2 * ((int64_t)n * ldb + k)
                         : Bb + 2 * ((int64_t)k * ldb + n);
      re = src[0];
      im = conj ? -src[1] : src[1];
    }
    const int nn = b_k ? r : tx, kk = b_k ? tx : r;
    tre[nn][kk] = re;
    tim[nn][kk] = im;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int n = n0 + r, k = k0 + tx;
    if (n >= N || k >= K) continue;
    T *dst = Rb + (int64_t)n * ldr + 2 * k;
    dst[0] = tre[r][tx];
    dst[1] = tim[r][tx];
  }
}

inline int grid_for(int64_t work, int num_cus) {
  int64_t g = cdiv(work, 256);
  int64_t cap = (int64_t)num_cus * 8;
  return (int)std::max<int64_t>(1, std::min(g, cap));
}

template <typename T>
constexpr int vec_of() {
  return 16 / (int)sizeof(T);
}

bool ptr_aligned(const void *p, int64_t off_elems, int elem) {
  return ((reinterpret_cast<uintptr_t>(p) + (uintptr_t)(off_elems * elem)) &
          15u) == 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// Real / half GEMM driver.
// ---------------------------------------------------------------------------
template <Elem E, typename OutT>
hcblasStatus_t gemm_real(hcblasHandle_t h, const Canon &c, double alpha,
                         double beta, void *C, int64_t ldc, int64_t cOff,
                         int64_t strideC, void *const *Cptrs, int64_t batch) {
  using T = typename ElemT<E>::T;
  using Acc = typename ElemT<E>::Acc;
  if (c.M == 0 || c.N == 0 || batch == 0) return HCBLAS_STATUS_SUCCESS;
  if (C == nullptr && Cptrs == nullptr) return HCBLAS_STATUS_INVALID_VALUE;
  if (alpha == 0.0 || c.K == 0) {
    if (beta == 1.0) return HCBLAS_STATUS_SUCCESS;
    for (int64_t b0 = 0; b0 < batch; b0 += 65535) {
      const int64_t nb = std::min<int64_t>(65535, batch - b0);
      dim3 g(grid_for((int64_t)c.M * c.N, h->num_cus), 1, (unsigned)nb);
      scale_c_kernel<OutT, Acc><<<g, 256, 0, h->stream>>>(
          c.M, c.N, Cptrs ? nullptr : static_cast<OutT *>(C) + b0 * strideC,
          ldc, strideC, Cptrs ? Cptrs + b0 : nullptr, cOff, (Acc)beta);
      hcblasStatus_t st = launch_status();
      if (st != HCBLAS_STATUS_SUCCESS) return st;
    }
    return HCBLAS_STATUS_SUCCESS;
  }
  if ((c.A == nullptr && c.Aptrs == nullptr) ||
      (c.B == nullptr && c.Bptrs == nullptr))
    return HCBLAS_STATUS_INVALID_VALUE;
  GemmParams p{};
  p.M = c.M;
  p.N = c.N;
  p.K = c.K;
  p.A = c.A;
  p.lda = c.lda;
  p.strideA = c.strideA;
  p.Aptrs = c.Aptrs;
  p.aOff = c.aOff;
  p.B = c.B;
  p.ldb = c.ldb;
  p.strideB = c.strideB;
  p.Bptrs = c.Bptrs;
  p.bOff = c.bOff;
  p.C = C;
  p.ldc = ldc;
  p.strideC = strideC;
  p.Cptrs = Cptrs;
  p.cOff = cOff;
  p.alpha = alpha;
  p.beta = beta;
  // 16-byte global loads need every matrix base, offset and leading dimension
  // (and batch stride) aligned to one vector. For pointer arrays the host
  // checks offsets and leading dimensions only; the kernel checks each
  // problem's bases and takes the element-wise loaders for a misaligned one.
  constexpr int V = vec_of<T>();
  // An r-contiguous operand also needs its row count to be a whole number
  // of vectors, so that a vector is wholly inside or outside the rows.
  bool vload = c.lda % V == 0 && c.ldb % V == 0 && c.aOff % V == 0 &&
               c.bOff % V == 0 && (c.a_k || c.M % V == 0) &&
               (c.b_k || c.N % V == 0);
  if (!c.Aptrs) vload = vload && ptr_aligned(c.A, c.aOff, sizeof(T)) &&
                        (batch == 1 || c.strideA % V == 0);
  if (!c.Bptrs) vload = vload && ptr_aligned(c.B, c.bOff, sizeof(T)) &&
                        (batch == 1 || c.strideB % V == 0);
  if (batch == 1 && vload && !c.Aptrs && !c.Bptrs && !Cptrs) {
    int ks = 0;
    const int S = splitk_count(c.M, c.N, c.K, h->num_cus,
                               E == Elem::F64 ? 16 : E == Elem::F32 ? 32 : 64,
                               ks);
    if (S > 1) {
      const size_t bytes = sizeof(Acc) * (size_t)S * c.M * c.N;
      if (ensure_workspace(h->kpart_ws, bytes) != hipSuccess)
        return HCBLAS_STATUS_ALLOC_FAILED;
      return launch_gemm_splitk<E, OutT>(h->stream, h->num_cus, p, c.a_k,
                                         c.b_k, S, ks,
                                         static_cast<Acc *>(h->kpart_ws.ptr));
    }
  }
  return launch_gemm<E, OutT>(h->stream, h->num_cus, p, c.a_k, c.b_k, vload,
                              batch);
}

// ---------------------------------------------------------------------------
// Complex GEMM driver (real embedding).
// ---------------------------------------------------------------------------
template <typename T>
hcblasStatus_t gemm_complex(hcblasHandle_t h, const Canon &c, T ar, T ai,
                            T br, T bi, void *C, int64_t ldc, int64_t cOff,
                            int64_t strideC, void *const *Cptrs,
                            int64_t batch) {
  constexpr Elem E = sizeof(T) == 4 ? Elem::F32 : Elem::F64;
  if (c.M == 0 || c.N == 0 || batch == 0) return HCBLAS_STATUS_SUCCESS;
  if (C == nullptr && Cptrs == nullptr) return HCBLAS_STATUS_INVALID_VALUE;
  const bool alpha0 = ar == T(0) && ai == T(0);
  const bool beta_real = bi == T(0);
  // beta*C first when beta is not real or when no product follows.
  if (alpha0 || c.K == 0 || !beta_real) {
    if (!(br == T(1) && bi == T(0))) {
      for (int64_t b0 = 0; b0 < batch; b0 += 65535) {
        const int64_t nb = std::min<int64_t>(65535, batch - b0);
        dim3 g(grid_for((int64_t)c.M * c.N, h->num_cus), 1, (unsigned)nb);
        scale_c_complex_kernel<T><<<g, 256, 0, h->stream>>>(
            c.M, c.N,
            Cptrs ? nullptr : static_cast<T *>(C) + 2 * b0 * strideC, ldc,
            strideC, Cptrs ? Cptrs + b0 : nullptr, cOff, br, bi);
        hcblasStatus_t st = launch_status();
        if (st != HCBLAS_STATUS_SUCCESS) return st;
      }
    }
    if (alpha0 || c.K == 0) return HCBLAS_STATUS_SUCCESS;
    br = T(1);
    bi = T(0);
  }
  if ((c.A == nullptr && c.Aptrs == nullptr) ||
      (c.B == nullptr && c.Bptrs == nullptr))
    return HCBLAS_STATUS_INVALID_VALUE;

  // Workspace: Ar (2M x ldr per batch), Br (N x ldr per batch) if packed.
  const int64_t ldr = (2 * (int64_t)c.K + 15) / 16 * 16;
  const int64_t a_elems = 2 * (int64_t)c.M * ldr;
  const bool pack_b = !c.b_k || c.conj_b;
  const int64_t b_elems = pack_b ? (int64_t)c.N * ldr : 0;
  const int64_t nb_ws = batch;
  size_t need = sizeof(T) * (size_t)(a_elems + b_elems) * (size_t)nb_ws;
  if (ensure_workspace(h->gemm_ws, need) != hipSuccess)
    return HCBLAS_STATUS_ALLOC_FAILED;
  T *Ar = static_cast<T *>(h->gemm_ws.ptr);
  T *Br = Ar + a_elems * nb_ws;

  for (int64_t b0 = 0; b0 < batch; b0 += 65535) {
    const int64_t nb = std::min<int64_t>(65535, batch - b0);
    dim3 ga((unsigned)cdiv(c.K, 32), (unsigned)cdiv(c.M, 32), (unsigned)nb);
    complex_embed_a_kernel<T><<<ga, 256, 0, h->stream>>>(
        c.M, c.K,
        c.Aptrs ? nullptr : static_cast<const T *>(c.A) + 2 * b0 * c.strideA,
        c.lda, c.aOff, c.strideA, c.Aptrs ? c.Aptrs + b0 : nullptr, c.a_k,
        c.conj_a, ar, ai, Ar + b0 * a_elems, ldr, a_elems);
    hcblasStatus_t st = launch_status();
    if (st != HCBLAS_STATUS_SUCCESS) return st;
    if (pack_b) {
      dim3 gb((unsigned)cdiv(c.K, 32), (unsigned)cdiv(c.N, 32), (unsigned)nb);
      complex_pack_b_kernel<T><<<gb, 256, 0, h->stream>>>(
          c.K, c.N,
          c.Bptrs ? nullptr
                  : static_cast<const T *>(c.B) + 2 * b0 * c.strideB,
          c.ldb, c.bOff, c.strideB, c.Bptrs ? c.Bptrs + b0 : nullptr, c.b_k,
          c.conj_b, Br + b0 * b_elems, ldr, b_elems);
      st = launch_status();
      if (st != HCBLAS_STATUS_SUCCESS) return st;
    }
  }

  Canon r{};
  r.M = 2 * c.M;
  r.N = c.N;
  r.K = 2 * c.K;
  r.a_k = true;
  r.b_k = true;
  r.A = Ar;
  r.lda = ldr;
  r.aOff = 0;
  r.strideA = a_elems;
  r.Aptrs = nullptr;
  if (pack_b) {
    r.B = Br;
    r.ldb = ldr;
    r.bOff = 0;
    r.strideB = b_elems;
    r.Bptrs = nullptr;
  } else {
    // Interleaved k-contiguous complex B is already the real view.
    r.B = c.B;
    r.ldb = 2 * c.ldb;
    r.bOff = 2 * c.bOff;
    r.strideB = 2 * c.strideB;
    r.Bptrs = c.Bptrs;
  }
  // Pointer-array C: the kernel adds cOff in real elements. The embedded
  // real problem follows the handle's math mode (split CGEMM, Ozaki ZGEMM).
  if constexpr (E == Elem::F32) {
    if (h->math_mode & HCBLAS_OZAKI_CRT_F32_MATH)
      return gemm_ozaki_crt_f32(h, r, 1.f, (float)br, static_cast<float *>(C),
                                2 * ldc, 2 * cOff, 2 * strideC, Cptrs, batch);
    if (h->math_mode & HCBLAS_SPLIT_BF16_MATH)
      return gemm_split_bf16(h, r, 1.f, (float)br, static_cast<float *>(C),
                             2 * ldc, 2 * cOff, 2 * strideC, Cptrs, batch);
  } else {
    if (h->math_mode & HCBLAS_OZAKI_CRT_MATH)
      return gemm_ozaki_crt(h, r, 1.0, (double)br, static_cast<double *>(C),
                            2 * ldc, 2 * cOff, 2 * strideC, Cptrs, batch);
    if (h->math_mode & HCBLAS_OZAKI_INT8_MATH)
      return gemm_ozaki_int8(h, r, 1.0, (double)br, static_cast<double *>(C),
                             2 * ldc, 2 * cOff, 2 * strideC, Cptrs, batch);
  }
  return gemm_real<E, T>(h, r, 1.0, (double)br, C, 2 * ldc, 2 * cOff,
                         2 * strideC, Cptrs, batch);
}

template hcblasStatus_t gemm_real<Elem::F32, float>(hcblasHandle_t,
                                                    const Canon &, double,
                                                    double, void *, int64_t,
                                                    int64_t, int64_t,
                                                    void *const *, int64_t);
template hcblasStatus_t gemm_real<Elem::F64, double>(hcblasHandle_t,
                                                     const Canon &, double,
                                                     double, void *, int64_t,
                                                     int64_t, int64_t,
                                                     void *const *, int64_t);

}  // namespace hcb

// ---------------------------------------------------------------------------
// Adjoint helpers
// ---------------------------------------------------------------------------
namespace hcb {
namespace {

// dst = beta * src over a batch of rows x cols matrices stored with
// (element (i, j) at i * rs + j * cs); beta == 0 writes zeros.
template <typename T>
__global__ __launch_bounds__(256) void scaled_copy_kernel(
    int rows, int cols, const T *src, int64_t src_rs, int64_t src_cs,
    int64_t src_stride, T *dst, int64_t dst_rs, int64_t dst_cs,
    int64_t dst_stride, T beta) {
  const T *s = src + (int64_t)blockIdx.z * src_stride;
  T *d = dst + (int64_t)blockIdx.z * dst_stride;
  const int64_t total = (int64_t)rows * cols;
  for (int64_t f = (int64_t)blockIdx.x * 256 + threadIdx.x; f < total;
       f += (int64_t)gridDim.x * 256) {
    int64_t i, j;
    if (dst_rs == 1) {  // column-major: i fastest
      j = f / rows;
      i = f - j * rows;
    } else {
      i = f / cols;
      j = f - i * cols;
    }
    d[i * dst_rs + j * dst_cs] = beta == T(0) ? T(0) : beta * s[i * src_rs + j * src_cs];
  }
}

}  // namespace
}  // namespace hcb

// ---------------------------------------------------------------------------
// Entry points
// ---------------------------------------------------------------------------
using namespace hcb;

namespace {

inline float half_to_float(hcblasHalf hb) {
  _Float16 f;
  std::memcpy(&f, &hb, 2);
  return (float)f;
}

// Dispatches a canonical problem of data type dt; alpha/beta point to the
// type's scalar (float for the 16-bit types).
hcblasStatus_t dispatch(hcblasHandle_t h, hcblasDatatype_t dt, const Canon &c,
                        const void *alpha, const void *beta, void *C,
                        int64_t ldc, int64_t cOff, int64_t strideC,
                        void *const *Cptrs, int64_t batch, int mode) {
  switch (dt) {
    case HCBLAS_R_32F: {
      const float a = *static_cast<const float *>(alpha);
      const float b = *static_cast<const float *>(beta);
      if (mode & HCBLAS_OZAKI_CRT_F32_MATH)
        return gemm_ozaki_crt_f32(h, c, a, b, static_cast<float *>(C), ldc,
                                  cOff, strideC, Cptrs, batch);
      if (mode & HCBLAS_SPLIT_BF16_MATH)
        return gemm_split_bf16(h, c, a, b, static_cast<float *>(C), ldc, cOff,
                               strideC, Cptrs, batch);
      return gemm_real<Elem::F32, float>(h, c, a, b, C, ldc, cOff, strideC,
                                         Cptrs, batch);
    }
    case HCBLAS_R_64F: {
      const double a = *static_cast<const double *>(alpha);
      const double b = *static_cast<const double *>(beta);
      if (mode & HCBLAS_OZAKI_CRT_MATH)
        return gemm_ozaki_crt(h, c, a, b, static_cast<double *>(C), ldc, cOff,
                              strideC, Cptrs, batch);
      if (mode & HCBLAS_OZAKI_INT8_MATH)
        return gemm_ozaki_int8(h, c, a, b, static_cast<double *>(C), ldc,
                               cOff, strideC, Cptrs, batch);
      return gemm_real<Elem::F64, double>(h, c, a, b, C, ldc, cOff, strideC,
                                          Cptrs, batch);
    }
    case HCBLAS_C_32F: {
      const float *a = static_cast<const float *>(alpha);
      const float *b = static_cast<const float *>(beta);
      return gemm_complex<float>(h, c, a[0], a[1], b[0], b[1], C, ldc, cOff,
                                 strideC, Cptrs, batch);
    }
    case HCBLAS_C_64F: {
      const double *a = static_cast<const double *>(alpha);
      const double *b = static_cast<const double *>(beta);
      return gemm_complex<double>(h, c, a[0], a[1], b[0], b[1], C, ldc, cOff,
                                  strideC, Cptrs, batch);
    }
    case HCBLAS_R_16F: {
      const float a = *static_cast<const float *>(alpha);
      const float b = *static_cast<const float *>(beta);
      return gemm_real<Elem::F16, uint16_t>(h, c, a, b, C, ldc, cOff, strideC,
                                            Cptrs, batch);
    }
    case HCBLAS_R_16BF: {
      const float a = *static_cast<const float *>(alpha);
      const float b = *static_cast<const float *>(beta);
      return gemm_real<Elem::BF16, float>(h, c, a, b, C, ldc, cOff, strideC,
                                          Cptrs, batch);
    }
  }
  return HCBLAS_STATUS_INVALID_VALUE;
}

bool valid_dtype(hcblasDatatype_t dt) {
  return dt >= HCBLAS_R_32F && dt <= HCBLAS_R_16BF;
}

hcblasStatus_t gemm_common(hcblasHandle_t h, hcblasDatatype_t dt,
                           hcblasOrder_t order, hcblasOperation_t ta,
                           hcblasOperation_t tb, int M, int N, int K,
                           const void *alpha, const void *A, int64_t lda,
                           int64_t aOff, int64_t sA, const void *const *Ap,
                           const void *B, int64_t ldb, int64_t bOff,
                           int64_t sB, const void *const *Bp, const void *beta,
                           void *C, int64_t ldc, int64_t cOff, int64_t sC,
                           void *const *Cp, int64_t batch,
                           int force_mode = 0) {
  if (!handle_ok(h)) return HCBLAS_STATUS_NOT_INITIALIZED;
  if (!valid_dtype(dt) || batch < 0) return HCBLAS_STATUS_INVALID_VALUE;
  Canon c{};
  hcblasStatus_t st = canonicalize(order, ta, tb, M, N, K, A, lda, aOff, sA,
                                   Ap, B, ldb, bOff, sB, Bp, ldc, c);
  if (st != HCBLAS_STATUS_SUCCESS) return st;
  if (M == 0 || N == 0 || batch == 0) return HCBLAS_STATUS_SUCCESS;
  if (alpha == nullptr || beta == nullptr) return HCBLAS_STATUS_INVALID_VALUE;
  if (hipSetDevice(h->device) != hipSuccess)
    return HCBLAS_STATUS_EXECUTION_FAILED;
  return dispatch(h, dt, c, alpha, beta, C, ldc, cOff, sC, Cp, batch,
                  force_mode ? force_mode : h->math_mode);
}

// C = alpha * op(P) * op(Q) (beta = 0) where P is FP32 and Q BF16
// (p_f32) or the reverse; output FP32. Used by the BF16 adjoints.
hcblasStatus_t mixed_gemm(hcblasHandle_t h, hcblasOrder_t order,
                          hcblasOperation_t tp, hcblasOperation_t tq, int M,
                          int N, int K, float alpha, const void *P,
                          int64_t ldp, int64_t sP, bool p_f32, const void *Q,
                          int64_t ldq, int64_t sQ, void *C, int64_t ldc,
                          int64_t sC, int64_t batch) {
  Canon c{};
  hcblasStatus_t st = canonicalize(order, tp, tq, M, N, K, P, ldp, 0, sP,
                                   nullptr, Q, ldq, 0, sQ, nullptr, ldc, c);
  if (st != HCBLAS_STATUS_SUCCESS) return st;
  if (C == nullptr || P == nullptr || Q == nullptr)
    return HCBLAS_STATUS_INVALID_VALUE;
  // the row-major form swaps the operands
  const bool a_f32 = order == HCBLAS_COL_MAJOR ? p_f32 : !p_f32;
  return gemm_mixed_bf16(h, c, a_f32, alpha, C, true, ldc, sC, batch);
}

}  // namespace

extern "C" {

hcblasStatus_t hcblasSgemm(hcblasHandle_t handle, hcblasOperation_t transa,
                           hcblasOperation_t transb, int m, int n, int k,
                           const float *alpha, const float *A, int lda,
                           const float *B, int ldb, const float *beta, float *C,
                           int ldc) {
  return gemm_common(handle, HCBLAS_R_32F, HCBLAS_COL_MAJOR, transa, transb, m,
                     n, k, alpha, A, lda, 0, 0, nullptr, B, ldb, 0, 0, nullptr,
                     beta, C, ldc, 0, 0, nullptr, 1);
}
hcblasStatus_t hcblasDgemm(hcblasHandle_t handle, hcblasOperation_t transa,
                           hcblasOperation_t transb, int m, int n, int k,
                           const double *alpha, const double *A, int lda,
                           const double *B, int ldb, const double *beta,
                           double *C, int ldc) {
  return gemm_common(handle, HCBLAS_R_64F, HCBLAS_COL_MAJOR, transa, transb, m,
                     n, k, alpha, A, lda, 0, 0, nullptr, B, ldb, 0, 0, nullptr,
                     beta, C, ldc, 0, 0, nullptr, 1);
}
hcblasStatus_t hcblasCgemm(hcblasHandle_t handle, hcblasOperation_t transa,
                           hcblasOperation_t transb, int m, int n, int k,
                           const hcComplex *alpha, const hcComplex *A, int lda,
                           const hcComplex *B, int ldb, const hcComplex *beta,
                           hcComplex *C, int ldc) {
  return gemm_common(handle, HCBLAS_C_32F, HCBLAS_COL_MAJOR, transa, transb, m,
                     n, k, alpha, A, lda, 0, 0, nullptr, B, ldb, 0, 0, nullptr,
                     beta, C, ldc, 0, 0, nullptr, 1);
}
hcblasStatus_t hcblasZgemm(hcblasHandle_t handle, hcblasOperation_t transa,
                           hcblasOperation_t transb, int m, int n, int k,
                           const hcDoubleComplex *alpha,
                           const hcDoubleComplex *A, int lda,
                           const hcDoubleComplex *B, int ldb,
                           const hcDoubleComplex *beta, hcDoubleComplex *C,
                           int ldc) {
  return gemm_common(handle, HCBLAS_C_64F, HCBLAS_COL_MAJOR, transa, transb, m,
                     n, k, alpha, A, lda, 0, 0, nullptr, B, ldb, 0, 0, nullptr,
                     beta, C, ldc, 0, 0, nullptr, 1);
}
hcblasStatus_t hcblasHgemm(hcblasHandle_t handle, hcblasOperation_t transa,
                           hcblasOperation_t transb, int m, int n, int k,
                           const hcblasHalf *alpha, const hcblasHalf *A,
                           int lda, const hcblasHalf *B, int ldb,
                           const hcblasHalf *beta, hcblasHalf *C, int ldc) {
  if (handle_ok(handle) && (alpha == nullptr || beta == nullptr))
    return HCBLAS_STATUS_INVALID_VALUE;
  float a = alpha ? half_to_float(*alpha) : 0.f;
  float b = beta ? half_to_float(*beta) : 0.f;
  return gemm_common(handle, HCBLAS_R_16F, HCBLAS_COL_MAJOR, transa, transb, m,
                     n, k, &a, A, lda, 0, 0, nullptr, B, ldb, 0, 0, nullptr,
                     &b, C, ldc, 0, 0, nullptr, 1);
}

#define HCB_BATCHED(NAME, DT, T, AT)                                          \
  hcblasStatus_t NAME(hcblasHandle_t handle, hcblasOperation_t transa,        \
                      hcblasOperation_t transb, int m, int n, int k,          \
                      const AT *alpha, const T *const Aarray[], int lda,      \
                      const T *const Barray[], int ldb, const AT *beta,       \
                      T *const Carray[], int ldc, int batchCount) {           \
    return gemm_common(handle, DT, HCBLAS_COL_MAJOR, transa, transb, m, n, k, \
                       alpha, nullptr, lda, 0, 0,                             \
                       reinterpret_cast<const void *const *>(Aarray),         \
                       nullptr, ldb, 0, 0,                                    \
                       reinterpret_cast<const void *const *>(Barray), beta,   \
                       nullptr, ldc, 0, 0,                                    \
                       reinterpret_cast<void *const *>(Carray), batchCount);  \
  }
HCB_BATCHED(hcblasSgemmBatched, HCBLAS_R_32F, float, float)
HCB_BATCHED(hcblasDgemmBatched, HCBLAS_R_64F, double, double)
HCB_BATCHED(hcblasCgemmBatched, HCBLAS_C_32F, hcComplex, hcComplex)
HCB_BATCHED(hcblasZgemmBatched, HCBLAS_C_64F, hcDoubleComplex,
            hcDoubleComplex)
#undef HCB_BATCHED

hcblasStatus_t hcblasHgemmBatched(hcblasHandle_t handle,
                                  hcblasOperation_t transa,
                                  hcblasOperation_t transb, int m, int n, int k,
                                  const hcblasHalf *alpha,
                                  const hcblasHalf *const Aarray[], int lda,
                                  const hcblasHalf *const Barray[], int ldb,
                                  const hcblasHalf *beta,
                                  hcblasHalf *const Carray[], int ldc,
                                  int batchCount) {
  if (handle_ok(handle) && (alpha == nullptr || beta == nullptr))
    return HCBLAS_STATUS_INVALID_VALUE;
  float a = alpha ? half_to_float(*alpha) : 0.f;
  float b = beta ? half_to_float(*beta) : 0.f;
  return gemm_common(handle, HCBLAS_R_16F, HCBLAS_COL_MAJOR, transa, transb, m,
                     n, k, &a, nullptr, lda, 0, 0,
                     reinterpret_cast<const void *const *>(Aarray), nullptr,
                     ldb, 0, 0, reinterpret_cast<const void *const *>(Barray),
                     &b, nullptr, ldc, 0, 0,
                     reinterpret_cast<void *const *>(Carray), batchCount);
}

hcblasStatus_t hcblas_sgemm(hcblasHandle_t handle, hcblasOrder_t order,
                            hcblasOperation_t typeA, hcblasOperation_t typeB,
                            int M, int N, int K, const float *alpha,
                            const float *A, int64_t lda, const float *B,
                            int64_t ldb, const float *beta, float *C,
                            int64_t ldc, int64_t aOffset, int64_t bOffset,
                            int64_t cOffset) {
  return gemm_common(handle, HCBLAS_R_32F, order, typeA, typeB, M, N, K, alpha,
                     A, lda, aOffset, 0, nullptr, B, ldb, bOffset, 0, nullptr,
                     beta, C, ldc, cOffset, 0, nullptr, 1);
}
hcblasStatus_t hcblas_dgemm(hcblasHandle_t handle, hcblasOrder_t order,
                            hcblasOperation_t typeA, hcblasOperation_t typeB,
                            int M, int N, int K, const double *alpha,
                            const double *A, int64_t lda, const double *B,
                            int64_t ldb, const double *beta, double *C,
                            int64_t ldc, int64_t aOffset, int64_t bOffset,
                            int64_t cOffset) {
  return gemm_common(handle, HCBLAS_R_64F, order, typeA, typeB, M, N, K, alpha,
                     A, lda, aOffset, 0, nullptr, B, ldb, bOffset, 0, nullptr,
                     beta, C, ldc, cOffset, 0, nullptr, 1);
}
hcblasStatus_t hcblas_hgemm(hcblasHandle_t handle, hcblasOrder_t order,
                            hcblasOperation_t typeA, hcblasOperation_t typeB,
                            int M, int N, int K, const hcblasHalf *alpha,
                            const hcblasHalf *A, int64_t lda,
                            const hcblasHalf *B, int64_t ldb,
                            const hcblasHalf *beta, hcblasHalf *C, int64_t ldc,
                            int64_t aOffset, int64_t bOffset,
                            int64_t cOffset) {
  if (handle_ok(handle) && (alpha == nullptr || beta == nullptr))
    return HCBLAS_STATUS_INVALID_VALUE;
  const float a = alpha ? half_to_float(*alpha) : 0.f;
  const float b = beta ? half_to_float(*beta) : 0.f;
  return gemm_common(handle, HCBLAS_R_16F, order, typeA, typeB, M, N, K, &a,
                     A, lda, aOffset, 0, nullptr, B, ldb, bOffset, 0, nullptr,
                     &b, C, ldc, cOffset, 0, nullptr, 1);
}
hcblasStatus_t hcblas_cgemm(hcblasHandle_t handle, hcblasOrder_t order,
                            hcblasOperation_t typeA, hcblasOperation_t typeB,
                            int M, int N, int K, const hcComplex *alpha,
                            const hcComplex *A, int64_t aOffset, int64_t lda,
                            const hcComplex *B, int64_t bOffset, int64_t ldb,
                            const hcComplex *beta, hcComplex *C,
                            int64_t cOffset, int64_t ldc) {
  return gemm_common(handle, HCBLAS_C_32F, order, typeA, typeB, M, N, K, alpha,
                     A, lda, aOffset, 0, nullptr, B, ldb, bOffset, 0, nullptr,
                     beta, C, ldc, cOffset, 0, nullptr, 1);
}
hcblasStatus_t hcblas_zgemm(hcblasHandle_t handle, hcblasOrder_t order,
                            hcblasOperation_t typeA, hcblasOperation_t typeB,
                            int M, int N, int K, const hcDoubleComplex *alpha,
                            const hcDoubleComplex *A, int64_t aOffset,
                            int64_t lda, const hcDoubleComplex *B,
                            int64_t bOffset, int64_t ldb,
                            const hcDoubleComplex *beta, hcDoubleComplex *C,
                            int64_t cOffset, int64_t ldc) {
  return gemm_common(handle, HCBLAS_C_64F, order, typeA, typeB, M, N, K, alpha,
                     A, lda, aOffset, 0, nullptr, B, ldb, bOffset, 0, nullptr,
                     beta, C, ldc, cOffset, 0, nullptr, 1);
}

hcblasStatus_t hcblas_gemm_strided_batched(
    hcblasHandle_t handle, hcblasDatatype_t dtype, hcblasOrder_t order,
    hcblasOperation_t typeA, hcblasOperation_t typeB, int M, int N, int K,
    const void *alpha, const void *A, int64_t lda, int64_t aOffset,
    int64_t strideA, const void *B, int64_t ldb, int64_t bOffset,
    int64_t strideB, const void *beta, void *C, int64_t ldc, int64_t cOffset,
    int64_t strideC, int batchCount) {
  return gemm_common(handle, dtype, order, typeA, typeB, M, N, K, alpha, A,
                     lda, aOffset, strideA, nullptr, B, ldb, bOffset, strideB,
                     nullptr, beta, C, ldc, cOffset, strideC, nullptr,
                     batchCount);
}

hcblasStatus_t hcblas_gemm_batched(
    hcblasHandle_t handle, hcblasDatatype_t dtype, hcblasOrder_t order,
    hcblasOperation_t typeA, hcblasOperation_t typeB, int M, int N, int K,
    const void *alpha, const void *const Aarray[], int64_t lda, int64_t aOffset,
    const void *const Barray[], int64_t ldb, int64_t bOffset, const void *beta,
    void *const Carray[], int64_t ldc, int64_t cOffset, int batchCount) {
  if (handle_ok(handle) && batchCount > 0 &&
      (Aarray == nullptr || Barray == nullptr || Carray == nullptr))
    return HCBLAS_STATUS_INVALID_VALUE;
  return gemm_common(handle, dtype, order, typeA, typeB, M, N, K, alpha,
                     nullptr, lda, aOffset, 0, Aarray, nullptr, ldb, bOffset, 0,
                     Barray, beta, nullptr, ldc, cOffset, 0, Carray,
                     batchCount);
}

hcblasStatus_t hcblas_sgemm_split_bf16(
    hcblasHandle_t handle, hcblasOrder_t order, hcblasOperation_t typeA,
    hcblasOperation_t typeB, int M, int N, int K, const float *alpha,
    const float *A, int64_t lda, const float *B, int64_t ldb, const float *beta,
    float *C, int64_t ldc, int64_t aOffset, int64_t bOffset, int64_t cOffset) {
  return gemm_common(handle, HCBLAS_R_32F, order, typeA, typeB, M, N, K, alpha,
                     A, lda, aOffset, 0, nullptr, B, ldb, bOffset, 0, nullptr,
                     beta, C, ldc, cOffset, 0, nullptr, 1,
                     HCBLAS_SPLIT_BF16_MATH);
}
hcblasStatus_t hcblas_sgemm_ozaki_crt(
    hcblasHandle_t handle, hcblasOrder_t order, hcblasOperation_t typeA,
    hcblasOperation_t typeB, int M, int N, int K, const float *alpha,
    const float *A, int64_t lda, const float *B, int64_t ldb, const float *beta,
    float *C, int64_t ldc, int64_t aOffset, int64_t bOffset, int64_t cOffset) {
  return gemm_common(handle, HCBLAS_R_32F, order, typeA, typeB, M, N, K, alpha,
                     A, lda, aOffset, 0, nullptr, B, ldb, bOffset, 0, nullptr,
                     beta, C, ldc, cOffset, 0, nullptr, 1,
                     HCBLAS_OZAKI_CRT_F32_MATH);
}
hcblasStatus_t hcblas_dgemm_ozaki_int8(
    hcblasHandle_t handle, hcblasOrder_t order, hcblasOperation_t typeA,
    hcblasOperation_t typeB, int M, int N, int K, const double *alpha,
    const double *A, int64_t lda, const double *B, int64_t ldb,
    const double *beta, double *C, int64_t ldc, int64_t aOffset,
    int64_t bOffset, int64_t cOffset) {
  return gemm_common(handle, HCBLAS_R_64F, order, typeA, typeB, M, N, K, alpha,
                     A, lda, aOffset, 0, nullptr, B, ldb, bOffset, 0, nullptr,
                     beta, C, ldc, cOffset, 0, nullptr, 1,
                     HCBLAS_OZAKI_INT8_MATH);
}
hcblasStatus_t hcblas_dgemm_ozaki_crt(
    hcblasHandle_t handle, hcblasOrder_t order, hcblasOperation_t typeA,
    hcblasOperation_t typeB, int M, int N, int K, const double *alpha,
    const double *A, int64_t lda, const double *B, int64_t ldb,
    const double *beta, double *C, int64_t ldc, int64_t aOffset,
    int64_t bOffset, int64_t cOffset) {
  return gemm_common(handle, HCBLAS_R_64F, order, typeA, typeB, M, N, K, alpha,
                     A, lda, aOffset, 0, nullptr, B, ldb, bOffset, 0, nullptr,
                     beta, C, ldc, cOffset, 0, nullptr, 1,
                     HCBLAS_OZAKI_CRT_MATH);
}

// Adjoints of the strided batched GEMM C = alpha*op(A)*op(B) + beta*C_in
// for a cotangent dC (no reference counterpart):
//   dA = alpha * dC * op(B)^T   (stored like A: alpha * op(B) * dC^T when
//                                op(A) = A^T),
//   dB = alpha * op(A)^T * dC   (alpha * dC^T * op(A) when op(B) = B^T),
//   dC_in = beta * dC.
// Each product runs on the MFMA GEMM kernels with the adjoint orientation.
// BF16: the mixed kernel reads A and B as they are and splits the FP32
// cotangent into three BF16 planes on the fly (exact), accumulates in FP32
// and rounds dA and dB once to BF16, the type of A and B.
hcblasStatus_t hcblas_gemm_strided_batched_backward(
    hcblasHandle_t handle, hcblasDatatype_t dtype, hcblasOrder_t order,
    hcblasOperation_t typeA, hcblasOperation_t typeB, int M, int N, int K,
    const void *alpha, const void *beta, const void *A, int64_t lda,
    int64_t strideA, const void *B, int64_t ldb, int64_t strideB,
    const void *dC, int64_t lddc, int64_t strideDC, void *dA, int64_t ldda,
    int64_t strideDA, void *dB, int64_t lddb, int64_t strideDB, void *dCin,
    int64_t lddcin, int64_t strideDCin, int batchCount) {
  if (!handle_ok(handle)) return HCBLAS_STATUS_NOT_INITIALIZED;
  if (dtype != HCBLAS_R_32F && dtype != HCBLAS_R_64F && dtype != HCBLAS_R_16BF)
    return HCBLAS_STATUS_INVALID_VALUE;
  if (order != HCBLAS_ROW_MAJOR && order != HCBLAS_COL_MAJOR)
    return HCBLAS_STATUS_INVALID_VALUE;
  if (canon_op(typeA) < 0 || canon_op(typeB) < 0)
    return HCBLAS_STATUS_INVALID_VALUE;
  typeA = static_cast<hcblasOperation_t>(canon_op(typeA));
  typeB = static_cast<hcblasOperation_t>(canon_op(typeB));
  if (M < 0 || N < 0 || K < 0 || batchCount < 0)
    return HCBLAS_STATUS_INVALID_VALUE;
  if (alpha == nullptr || beta == nullptr || dC == nullptr)
    return HCBLAS_STATUS_INVALID_VALUE;
  if (batchCount == 0) return HCBLAS_STATUS_SUCCESS;
  const bool ta = typeA != HCBLAS_OP_N, tb = typeB != HCBLAS_OP_N;
  const hcblasOperation_t OP_N = HCBLAS_OP_N, OP_T = HCBLAS_OP_T;
  const hcblasDatatype_t wt = dtype == HCBLAS_R_16BF ? HCBLAS_R_32F : dtype;
  const float zf = 0.f;
  const double zd = 0.0;
  const void *zero = wt == HCBLAS_R_64F ? (const void *)&zd : (const void *)&zf;
  if (dtype == HCBLAS_R_16BF) {
    // BF16 operands are read as they are by the mixed kernel (FP32 dC split
    // into three BF16 planes on the fly); same products as below.
    if ((dA && B == nullptr) || (dB && A == nullptr))
      return HCBLAS_STATUS_INVALID_VALUE;
    const float al = *static_cast<const float *>(alpha);
    hcblasStatus_t st = HCBLAS_STATUS_SUCCESS;
    if (dA) {
      st = !ta ? mixed_gemm(handle, order, OP_N, tb ? OP_N : OP_T, M, K, N, al,
                            dC, lddc, strideDC, true, B, ldb, strideB,
                            dA, ldda, strideDA, batchCount)
               : mixed_gemm(handle, order, tb ? OP_T : OP_N, OP_T, K, M, N, al,
                            B, ldb, strideB, false, dC, lddc, strideDC, dA,
                            ldda, strideDA, batchCount);
      if (st != HCBLAS_STATUS_SUCCESS) return st;
    }
    if (dB) {
      st = !tb ? mixed_gemm(handle, order, ta ? OP_N : OP_T, OP_N, K, N, M, al,
                            A, lda, strideA, false, dC, lddc, strideDC,
                            dB, lddb, strideDB, batchCount)
               : mixed_gemm(handle, order, OP_T, ta ? OP_T : OP_N, N, K, M, al,
                            dC, lddc, strideDC, true, A, lda, strideA, dB,
                            lddb, strideDB, batchCount);
      if (st != HCBLAS_STATUS_SUCCESS) return st;
    }
    dA = dB = nullptr;  // done; dC_in below
  }
  hcblasStatus_t st = HCBLAS_STATUS_SUCCESS;
  if (dA) {
    if (!ta)
      st = gemm_common(handle, wt, order, OP_N, tb ? OP_N : OP_T, M, K, N,
                       alpha, dC, lddc, 0, strideDC, nullptr, B, ldb, 0,
                       strideB, nullptr, zero, dA, ldda, 0, strideDA, nullptr,
                       batchCount);
    else
      st = gemm_common(handle, wt, order, tb ? OP_T : OP_N, OP_T, K, M, N,
                       alpha, B, ldb, 0, strideB, nullptr, dC, lddc, 0,
                       strideDC, nullptr, zero, dA, ldda, 0, strideDA, nullptr,
                       batchCount);
    if (st != HCBLAS_STATUS_SUCCESS) return st;
  }
  if (dB) {
    if (!tb)
      st = gemm_common(handle, wt, order, ta ? OP_N : OP_T, OP_N, K, N, M,
                       alpha, A, lda, 0, strideA, nullptr, dC, lddc, 0,
                       strideDC, nullptr, zero, dB, lddb, 0, strideDB, nullptr,
                       batchCount);
    else
      st = gemm_common(handle, wt, order, OP_T, ta ? OP_T : OP_N, N, K, M,
                       alpha, dC, lddc, 0, strideDC, nullptr, A, lda, 0,
                       strideA, nullptr, zero, dB, lddb, 0, strideDB, nullptr,
                       batchCount);
    if (st != HCBLAS_STATUS_SUCCESS) return st;
  }
  if (dCin && M > 0 && N > 0) {
    const bool col = order == HCBLAS_COL_MAJOR;
    const int64_t src_ld = lddc, dst_ld = lddcin;
    if (src_ld < (col ? M : N) || dst_ld < (col ? M : N))
      return HCBLAS_STATUS_INVALID_VALUE;
    if (hipSetDevice(handle->device) != hipSuccess)
      return HCBLAS_STATUS_EXECUTION_FAILED;
    for (int64_t b0 = 0; b0 < batchCount; b0 += 65535) {
      const unsigned nb = (unsigned)std::min<int64_t>(65535, batchCount - b0);
      dim3 g(grid_for((int64_t)M * N, handle->num_cus), 1, nb);
      if (wt == HCBLAS_R_64F)
        scaled_copy_kernel<double><<<g, 256, 0, handle->stream>>>(
            M, N, static_cast<const double *>(dC) + b0 * strideDC,
            col ? 1 : src_ld, col ? src_ld : 1, strideDC,
            static_cast<double *>(dCin) + b0 * strideDCin, col ? 1 : dst_ld,
            col ? dst_ld : 1, strideDCin, *static_cast<const double *>(beta));
      else
        scaled_copy_kernel<float><<<g, 256, 0, handle->stream>>>(
            M, N, static_cast<const float *>(dC) + b0 * strideDC,
            col ? 1 : src_ld, col ? src_ld : 1, strideDC,
            static_cast<float *>(dCin) + b0 * strideDCin, col ? 1 : dst_ld,
            col ? dst_ld : 1, strideDCin, *static_cast<const float *>(beta));
      st = launch_status();
      if (st != HCBLAS_STATUS_SUCCESS) return st;
    }
  }
  return HCBLAS_STATUS_SUCCESS;
}

}  // extern "C"
