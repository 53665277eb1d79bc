// Mixed BF16 / FP32 GEMM for the adjoints of the BF16 batched GEMM (no
// reference counterpart): C = alpha * A' * B' in FP32 where one operand is
// FP32 (the cotangent dC) and the other BF16 (the forward's A or B), read as
// they lie in memory -- no widened copy of the BF16 operand.
//
// The FP32 operand is split on its way into LDS into three BF16 planes,
// x = hi + mid + lo with each part the BF16 rounding of the remaining
// residual (exact for finite x: 3 x 8 bits hold the 24-bit significand), so
// hi*b + mid*b + lo*b equals x*b exactly for a BF16 b. Each K chunk issues
// the three products on v_mfma_f32_32x32x16_bf16 from one BF16 fragment of
// the other operand (smallest first) into one FP32 accumulator: the sums are
// those of an FP32 GEMM of the exact products, at three BF16 MFMAs per 16 k
// instead of eight FP32 MFMAs (v_mfma_f32_32x32x2f32 takes 2 k).
//
// Tile: 128 x 128 x 32, 4 waves of 64 x 64 (2 x 2 blocks of 32 x 32), LDS
// double-buffered like the plain kernel (the FP32 operand's tile is three
// planes), epilogue through LDS with the plain kernels' arithmetic; the
// result is written as FP32 or rounded once to BF16.
//
// Both operands reach the kernel k-contiguous: an operand stored the other
// way (the cotangent and the forward operand of dB = op(A)^T dC, whose
// inner index m is each one's row index) is first transposed into the
// handle's workspace by a 64 x 64 LDS-tiled copy. Staging such a tile
// directly would take one 2-byte LDS write per element and plane.
#include "gemm_common.hpp"

namespace hcb {
namespace {

__device__ __forceinline__ uint16_t bf16_rne(float x) {
  const uint32_t u = __float_as_uint(x);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_val(uint16_t b) {
  return __uint_as_float((uint32_t)b << 16);
}
// x = hi + mid + lo (BF16 bits); non-finite x keeps hi = bf16(x) (NaN stays
// NaN) and zero residual planes.
__device__ __forceinline__ void split3_bf16(float x, uint16_t &h, uint16_t &m,
                                            uint16_t &l) {
  const uint32_t u = __float_as_uint(x);
  if ((u & 0x7F800000u) == 0x7F800000u) {
    h = (uint16_t)((u >> 16) | ((u & 0x7FFFFFu) ? 0x40u : 0u));
    m = l = 0;
    return;
  }
  h = bf16_rne(x);
  const float r1 = x - bf16_val(h);
  m = bf16_rne(r1);
  l = bf16_rne(r1 - bf16_val(m));
}

// Global -> register loader of an FP32 tile (the plain kernel's loader) whose
// store writes three BF16 planes [ROWS][BK] (row stride BK*2 + 16 bytes,
// planes PLANE bytes apart) in the BF16 operand's LDS layout.
template <int ROWS, int BK, bool VLOAD, int NTH>
struct SplitPlaneLoader {
  using L = TileLoader<float, ROWS, BK, true, VLOAD, NTH>;
  static constexpr int ROWB = BK * 2 + 16;
  static constexpr int PLANE = ROWS * ROWB;
  L ld;
  __device__ __forceinline__ void load_step(const float *__restrict__ P,
                                            int64_t ld_, int r0, int k0,
                                            int R, int K, bool full) {
    ld.load_step(P, ld_, r0, k0, R, K, full);
  }
  __device__ __forceinline__ void store(char *lds) const {
#pragma unroll
    for (int i = 0; i < L::PER_THREAD; ++i) {
      const int v = threadIdx.x + i * NTH;
      if (L::NV % NTH != 0 && v >= L::NV) continue;
      int r, k;
      L::vec_pos(v, r, k);
      float t[4];
      __builtin_memcpy(t, &ld.reg[i], 16);
      uint16_t h[4], m[4], l[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) split3_bf16(t[e], h[e], m[e], l[e]);
      {  // four consecutive k of row r: 8 bytes per plane
        char *p = lds + r * ROWB + k * 2;
        uint2 hv = make_uint2(h[0] | ((uint32_t)h[1] << 16),
                              h[2] | ((uint32_t)h[3] << 16));
        uint2 mv = make_uint2(m[0] | ((uint32_t)m[1] << 16),
                              m[2] | ((uint32_t)m[3] << 16));
        uint2 lv = make_uint2(l[0] | ((uint32_t)l[1] << 16),
                              l[2] | ((uint32_t)l[3] << 16));
        *reinterpret_cast<uint2 *>(p) = hv;
        *reinterpret_cast<uint2 *>(p + PLANE) = mv;
        *reinterpret_cast<uint2 *>(p + 2 * PLANE) = lv;
      }
    }
  }
};

constexpr int kMBK = 32;
constexpr int kMRowB = kMBK * 2 + 16;
// LDS bytes of one K step: the FP32 operand's three planes and the BF16
// operand's tile; two of them double-buffer the main loop.
template <bool AF, int BM, int BN>
constexpr int mixed_buf_bytes() {
  return ((AF ? 3 : 1) * BM + (AF ? 1 : 3) * BN) * kMRowB;
}

// BF16 output element (bits), rounded to nearest even from FP32; NaN stays
// a quiet NaN.
struct Bf16Out {
  uint16_t bits;
};
}  // namespace
template <>
__device__ __forceinline__ Bf16Out to_out<Bf16Out, float>(float v) {
  const uint32_t u = __float_as_uint(v);
  if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x7FFFFFu))
    return Bf16Out{(uint16_t)((u >> 16) | 0x40u)};
  return Bf16Out{bf16_rne(v)};
}
template <>
__device__ __forceinline__ float from_out<float, Bf16Out>(Bf16Out v) {
  return bf16_val(v.bits);
}
namespace {

struct MixedParams {
  int M, N, K;
  const void *A;
  int64_t lda, strideA;
  const void *B;
  int64_t ldb, strideB;
  void *C;
  int64_t ldc, strideC;
  float alpha;
  int tiles_m, tiles_n;
};

// AF: A' is the FP32 operand (else B' is); both are k-contiguous. VL:
// 16-byte global loads legal for both. OutT: float or Bf16Out. BM x BN
// tile over WM x WN waves.
template <bool AF, bool VL, typename OutT, int BM, int BN, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN) void gemm_mixed_kernel(
    MixedParams p) {
  constexpr int NTH = 64 * WM * WN, MI = 32;
  using WT = WaveTile<Elem::BF16, BM, BN, WM, WN, MI>;
  constexpr int FM = WT::FM, FN = WT::FN, WTM = WT::WTM, WTN = WT::WTN;
  constexpr int PLANE_A = BM * kMRowB, PLANE_B = BN * kMRowB;
  constexpr int BUF = mixed_buf_bytes<AF, BM, BN>(), LDS = 2 * BUF;
  __shared__ __attribute__((aligned(16))) char lds[LDS];
  using FL_A = SplitPlaneLoader<BM, kMBK, VL, NTH>;
  using FL_B = SplitPlaneLoader<BN, kMBK, VL, NTH>;
  using HL_A = TileLoader<uint16_t, BM, kMBK, true, VL, NTH>;
  using HL_B = TileLoader<uint16_t, BN, kMBK, true, VL, NTH>;
  using LA = typename std::conditional<AF, FL_A, HL_A>::type;
  using LB = typename std::conditional<AF, HL_B, FL_B>::type;
  using TA = typename std::conditional<AF, float, uint16_t>::type;
  using TB = typename std::conditional<AF, uint16_t, float>::type;
  // LDS buffer: A region (3 planes if FP32, else 1) then B region
  constexpr int A_BYTES = (AF ? 3 : 1) * PLANE_A;

  const int z = blockIdx.z;
  const TA *A = static_cast<const TA *>(p.A) + (int64_t)z * p.strideA;
  const TB *B = static_cast<const TB *>(p.B) + (int64_t)z * p.strideB;
  OutT *C = static_cast<OutT *>(p.C) + (int64_t)z * p.strideC;
  int tile_m, tile_n;
  tile_coords(blockIdx.x, p.tiles_m, p.tiles_n, tile_m, tile_n);
  const int m0 = tile_m * BM, n0 = tile_n * BN;

  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wm = wave / WN, wn = wave % WN;
  const int lrow = lane % MI, lgrp = lane / MI;
  typename WT::AccV acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = typename WT::AccV{};

  LA la;
  LB lb;
  const int K = p.K;
  const int nk = (K + kMBK - 1) / kMBK, nfull = K / kMBK;
  auto buf = [&](int kt) { return lds + (kt & 1) * BUF; };
  auto load = [&](int k, bool full) {
    la.load_step(A, p.lda, m0, k * kMBK, p.M, K, full);
    lb.load_step(B, p.ldb, n0, k * kMBK, p.N, K, full);
  };
  auto store = [&](int k) {
    la.store(buf(k));
    lb.store(buf(k) + A_BYTES);
  };
  auto compute = [&](const char *cur) {
    const char *As = cur + (wm * WTM + lrow) * kMRowB;
    const char *Bs = cur + A_BYTES + (wn * WTN + lrow) * kMRowB;
#pragma unroll
    for (int c = 0; c < kMBK / 16; ++c) {
      const int kbyte = (c * 16 + lgrp * 8) * 2;
      vec4u af[AF ? 3 : 1][FM], bf[AF ? 1 : 3][FN];
#pragma unroll
      for (int q = 0; q < (AF ? 3 : 1); ++q)
#pragma unroll
        for (int i = 0; i < FM; ++i)
          af[q][i] = *reinterpret_cast<const vec4u *>(
              As + q * PLANE_A + i * MI * kMRowB + kbyte);
#pragma unroll
      for (int q = 0; q < (AF ? 1 : 3); ++q)
#pragma unroll
        for (int j = 0; j < FN; ++j)
          bf[q][j] = *reinterpret_cast<const vec4u *>(
              Bs + q * PLANE_B + j * MI * kMRowB + kbyte);
      // lo, mid, hi of the FP32 operand against the BF16 fragment
#pragma unroll
      for (int q = 2; q >= 0; --q)
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j)
            mfma_step<Elem::BF16, MI, 0>(af[AF ? q : 0][i], bf[AF ? 0 : q][j],
                                         acc[i][j]);
    }
  };
  if (nk > 0) {
    load(0, nfull > 0);
    store(0);
    if (nk > 1) load(1, 1 < nfull);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
      if (kt + 1 < nk) store(kt + 1);
      if (kt + 2 < nk) load(kt + 2, kt + 2 < nfull);
      __builtin_amdgcn_sched_barrier(0);
      compute(buf(kt));
      __syncthreads();
    }
  }
  const bool vec_c = reinterpret_cast<uintptr_t>(C) % 16 == 0 &&
                     p.ldc % (16 / (int)sizeof(OutT)) == 0;
  store_tile_lds<Elem::BF16, OutT, BM, BN, WM, WN, MI, LDS>(
      acc, C, p.ldc, p.M, p.N, m0, n0, (double)p.alpha, 0.0, lds, vec_c);
}

// Tiles: 256 x 256 over 8 waves of 64 x 128 (FP32 operand along the wave
// tile's short side: 10 fragment reads per 24 MFMAs of a 16-k chunk, against
// 8 per 12 for 64 x 64 wave tiles) when both output sides reach 256, else
// 128 x 128 over 4 waves of 64 x 64.
template <bool AF, typename OutT, int BM, int BN, int WM, int WN>
hcblasStatus_t launch_mixed_cfg(hipStream_t s, MixedParams p, bool vl,
                                int64_t batch) {
  p.tiles_m = (int)cdiv(p.M, BM);
  p.tiles_n = (int)cdiv(p.N, BN);
  const unsigned tiles = (unsigned)(p.tiles_m * p.tiles_n);
  for (int64_t b0 = 0; b0 < batch; b0 += 65535) {
    MixedParams q = p;
    const int64_t nb = batch - b0 < 65535 ? batch - b0 : 65535;
    q.A = static_cast<const char *>(q.A) +
          b0 * q.strideA * (AF ? sizeof(float) : sizeof(uint16_t));
    q.B = static_cast<const char *>(q.B) +
          b0 * q.strideB * (AF ? sizeof(uint16_t) : sizeof(float));
    q.C = static_cast<OutT *>(q.C) + b0 * q.strideC;
    dim3 grid(tiles, 1, (unsigned)nb);
    if (vl)
      gemm_mixed_kernel<AF, true, OutT, BM, BN, WM, WN>
          <<<grid, 64 * WM * WN, 0, s>>>(q);
    else
      gemm_mixed_kernel<AF, false, OutT, BM, BN, WM, WN>
          <<<grid, 64 * WM * WN, 0, s>>>(q);
    const hcblasStatus_t st = launch_status();
    if (st != HCBLAS_STATUS_SUCCESS) return st;
  }
  return HCBLAS_STATUS_SUCCESS;
}
template <bool AF, typename OutT>
hcblasStatus_t launch_mixed(hipStream_t s, MixedParams p, bool vl,
                            int64_t batch) {
  if (p.M >= 256 && p.N >= 256) {
    if constexpr (AF)
      return launch_mixed_cfg<AF, OutT, 256, 256, 4, 2>(s, p, vl, batch);
    else
      return launch_mixed_cfg<AF, OutT, 256, 256, 2, 4>(s, p, vl, batch);
  }
  return launch_mixed_cfg<AF, OutT, 128, 128, 2, 2>(s, p, vl, batch);
}

// out[z][c * ldo + r] = in[z][r * ldi + c] for r < rows, c < cols, z <
// batch: 64 x 64 tiles through LDS (row stride 65 elements), 64 consecutive
// elements of a line per wave on both sides.
template <typename T>
__global__ __launch_bounds__(256) void transpose_kernel(
    int rows, int cols, const T *__restrict__ in, int64_t ldi, int64_t si,
    T *__restrict__ out, int64_t ldo, int64_t so) {
  __shared__ T t[64][65];
  const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const T *src = in + (int64_t)blockIdx.z * si;
  T *dst = out + (int64_t)blockIdx.z * so;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int r = r0 + ty + 4 * i, c = c0 + tx;
    if (r < rows && c < cols) t[ty + 4 * i][tx] = src[(int64_t)r * ldi + c];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int c = c0 + ty + 4 * i, r = r0 + tx;
    if (r < rows && c < cols) dst[(int64_t)c * ldo + r] = t[tx][ty + 4 * i];
  }
}

// The same copy with 16-byte accesses on both sides, for whole tiles of 64
// rows x 256 bytes (rows % 64 == 0, cols a multiple of 256 bytes, every
// base, leading dimension and batch stride a whole number of vectors): a
// thread loads four 16-byte row pieces into LDS (row stride one element
// more than the tile width), then gathers four 16-byte output pieces down
// the LDS columns. One element per lane each way moved 4.4 TB/s (FP32) and
// 3.7 TB/s (BF16) on the differentiation configuration's copies.
template <typename T>
__global__ __launch_bounds__(256) void transpose_vec_kernel(
    const T *__restrict__ in, int64_t ldi, int64_t si, T *__restrict__ out,
    int64_t ldo, int64_t so) {
  constexpr int VE = 16 / (int)sizeof(T);  // elements per vector
  constexpr int TC = 256 / (int)sizeof(T);  // tile columns (input)
  constexpr int TR = 64;                    // tile rows (input)
  __shared__ T t[TR][TC + 1];
  const int r0 = blockIdx.y * TR, c0 = blockIdx.x * TC;
  const T *src = in + (int64_t)blockIdx.z * si;
  T *dst = out + (int64_t)blockIdx.z * so;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int v = threadIdx.x + 256 * i;
    const int r = v / (TC / VE), vc = (v % (TC / VE)) * VE;
    const vec4u w = *reinterpret_cast<const vec4u *>(
        src + (int64_t)(r0 + r) * ldi + c0 + vc);
    T e[VE];
    __builtin_memcpy(e, &w, 16);
#pragma unroll
    for (int j = 0; j < VE; ++j) t[r][vc + j] = e[j];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int v = threadIdx.x + 256 * i;
    const int c = v / (TR / VE), rv = (v % (TR / VE)) * VE;
    T e[VE];
#pragma unroll
    for (int j = 0; j < VE; ++j) e[j] = t[rv + j][c];
    vec4u w;
    __builtin_memcpy(&w, e, 16);
    *reinterpret_cast<vec4u *>(dst + (int64_t)(c0 + c) * ldo + r0 + rv) = w;
  }
}

template <typename T>
void launch_transpose(hipStream_t s, int rows, int cols, const T *in,
                      int64_t ldi, int64_t si, T *out, int64_t ldo, int64_t so,
                      int64_t batch) {
  constexpr int VE = 16 / (int)sizeof(T), TC = 256 / (int)sizeof(T);
  const bool vec = rows % 64 == 0 && cols % TC == 0 && ldi % VE == 0 &&
                   ldo % VE == 0 && si % VE == 0 && so % VE == 0 &&
                   reinterpret_cast<uintptr_t>(in) % 16 == 0 &&
                   reinterpret_cast<uintptr_t>(out) % 16 == 0;
  for (int64_t b0 = 0; b0 < batch; b0 += 65535) {
    const int64_t nb = batch - b0 < 65535 ? batch - b0 : 65535;
    if (vec) {
      dim3 g((unsigned)(cols / TC), (unsigned)(rows / 64), (unsigned)nb);
      transpose_vec_kernel<T><<<g, 256, 0, s>>>(in + b0 * si, ldi, si,
                                                out + b0 * so, ldo, so);
    } else {
      dim3 g((unsigned)cdiv(cols, 64), (unsigned)cdiv(rows, 64),
             (unsigned)nb);
      transpose_kernel<T><<<g, 256, 0, s>>>(rows, cols, in + b0 * si, ldi,
                                            si, out + b0 * so, ldo, so);
    }
  }
}

inline bool aligned_ptr(const void *p, size_t esize, int64_t off) {
  return ((reinterpret_cast<uintptr_t>(p) + (uintptr_t)(off * esize)) & 15u) ==
         0;
}

}  // namespace

hcblasStatus_t gemm_mixed_bf16(hcblasHandle_t h, const Canon &c, bool a_f32,
                               float alpha, void *C, bool c_bf16, int64_t ldc,
                               int64_t strideC, int64_t batch) {
  if (c.M == 0 || c.N == 0 || batch == 0) return HCBLAS_STATUS_SUCCESS;
  if (hipSetDevice(h->device) != hipSuccess)
    return HCBLAS_STATUS_EXECUTION_FAILED;
  MixedParams p{};
  p.M = c.M;
  p.N = c.N;
  p.K = c.K;
  const size_t ea = a_f32 ? 4 : 2, eb = a_f32 ? 2 : 4;
  const void *A = static_cast<const char *>(c.A) + c.aOff * ea;
  const void *B = static_cast<const char *>(c.B) + c.bOff * eb;
  int64_t lda = c.lda, ldb = c.ldb, sA = c.strideA, sB = c.strideB;
  // Operands stored m- / n-contiguous go k-contiguous through the workspace
  // (K rounded up to whole 16-byte vectors so the copies keep vector loads).
  const int64_t kpa = cdiv(c.K, 16 / (int64_t)ea) * (16 / (int64_t)ea);
  const int64_t kpb = cdiv(c.K, 16 / (int64_t)eb) * (16 / (int64_t)eb);
  const size_t ta_bytes = c.a_k ? 0 : (size_t)batch * c.M * kpa * ea;
  const size_t tb_bytes = c.b_k ? 0 : (size_t)batch * c.N * kpb * eb;
  if (ta_bytes + tb_bytes && c.K > 0) {  // K = 0 reads no operand
    const size_t off_b = (ta_bytes + 255) & ~size_t(255);
    if (ensure_workspace(h->adj_ws, off_b + tb_bytes) != hipSuccess)
      return HCBLAS_STATUS_ALLOC_FAILED;
    char *ws = static_cast<char *>(h->adj_ws.ptr);
    auto tr = [&](const void *X, int64_t ld, int64_t sx, size_t es, int R,
                  int64_t kp, void *dst) {
      // X(r, k) at X[r + k * ld]: rows = k, cols = r of the stored matrix
      if (es == 4)
        launch_transpose<float>(h->stream, c.K, R,
                                static_cast<const float *>(X), ld, sx,
                                static_cast<float *>(dst), kp, R * kp, batch);
      else
        launch_transpose<uint16_t>(h->stream, c.K, R,
                                   static_cast<const uint16_t *>(X), ld, sx,
                                   static_cast<uint16_t *>(dst), kp, R * kp,
                                   batch);
    };
    if (!c.a_k) {
      tr(A, lda, sA, ea, c.M, kpa, ws);
      A = ws;
      lda = kpa;
      sA = c.M * kpa;
    }
    if (!c.b_k) {
      tr(B, ldb, sB, eb, c.N, kpb, ws + off_b);
      B = ws + off_b;
      ldb = kpb;
      sB = c.N * kpb;
    }
    const hcblasStatus_t st = launch_status();
    if (st != HCBLAS_STATUS_SUCCESS) return st;
  }
  p.A = A;
  p.lda = lda;
  p.strideA = sA;
  p.B = B;
  p.ldb = ldb;
  p.strideB = sB;
  p.C = C;
  p.ldc = ldc;
  p.strideC = strideC;
  p.alpha = alpha;
  // 16-byte loads: every base, leading dimension and batch stride a whole
  // number of vectors of its type.
  const int va = 16 / (int)ea, vb = 16 / (int)eb;
  const bool vl = lda % va == 0 && ldb % vb == 0 && aligned_ptr(A, ea, 0) &&
                  aligned_ptr(B, eb, 0) &&
                  (batch == 1 || (sA % va == 0 && sB % vb == 0));
  if (c_bf16)
    return a_f32 ? launch_mixed<true, Bf16Out>(h->stream, p, vl, batch)
                 : launch_mixed<false, Bf16Out>(h->stream, p, vl, batch);
  return a_f32 ? launch_mixed<true, float>(h->stream, p, vl, batch)
               : launch_mixed<false, float>(h->stream, p, vl, batch);
}

}  // namespace hcb
