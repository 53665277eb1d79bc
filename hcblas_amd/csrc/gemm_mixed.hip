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
// planes), epilogue through LDS with the plain kernels' arithmetic.
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
template <int ROWS, int BK, bool KCONT, bool VLOAD, int NTH>
struct SplitPlaneLoader {
  using L = TileLoader<float, ROWS, BK, KCONT, VLOAD, NTH>;
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
      if (KCONT) {  // four consecutive k of row r: 8 bytes per plane
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
      } else {  // four consecutive rows at one k
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          char *p = lds + (r + e) * ROWB + k * 2;
          *reinterpret_cast<uint16_t *>(p) = h[e];
          *reinterpret_cast<uint16_t *>(p + PLANE) = m[e];
          *reinterpret_cast<uint16_t *>(p + 2 * PLANE) = l[e];
        }
      }
    }
  }
};

constexpr int kMBM = 128, kMBN = 128, kMBK = 32, kMWM = 2, kMWN = 2;
constexpr int kMRowB = kMBK * 2 + 16;
// one buffer: the FP32 operand's three planes and the BF16 operand's tile
constexpr int kMBuf = 4 * kMBM * kMRowB;
constexpr int kMLds = 2 * kMBuf;
static_assert(kMBM == kMBN, "the two operand tiles share one plane size");

struct MixedParams {
  int M, N, K;
  const void *A;
  int64_t lda, strideA;
  const void *B;
  int64_t ldb, strideB;
  float *C;
  int64_t ldc, strideC;
  float alpha;
  int tiles_m, tiles_n;
};

// AF: A' is the FP32 operand (else B' is). A_K / B_K: k-contiguous
// operands. VL: 16-byte global loads legal for both.
template <bool AF, bool A_K, bool B_K, bool VL>
__global__ __launch_bounds__(64 * kMWM * kMWN) void gemm_mixed_kernel(
    MixedParams p) {
  constexpr int NTH = 64 * kMWM * kMWN, MI = 32;
  using WT = WaveTile<Elem::BF16, kMBM, kMBN, kMWM, kMWN, MI>;
  constexpr int FM = WT::FM, FN = WT::FN, WTM = WT::WTM, WTN = WT::WTN;
  constexpr int PLANE = kMBM * kMRowB;
  __shared__ __attribute__((aligned(16))) char lds[kMLds];
  using FL_A = SplitPlaneLoader<kMBM, kMBK, A_K, VL, NTH>;
  using FL_B = SplitPlaneLoader<kMBN, kMBK, B_K, VL, NTH>;
  using HL_A = TileLoader<uint16_t, kMBM, kMBK, A_K, VL, NTH>;
  using HL_B = TileLoader<uint16_t, kMBN, kMBK, B_K, VL, NTH>;
  using LA = typename std::conditional<AF, FL_A, HL_A>::type;
  using LB = typename std::conditional<AF, HL_B, FL_B>::type;
  using TA = typename std::conditional<AF, float, uint16_t>::type;
  using TB = typename std::conditional<AF, uint16_t, float>::type;
  // LDS buffer: A region (3 planes if FP32, else 1) then B region
  constexpr int A_BYTES = (AF ? 3 : 1) * PLANE;

  const int z = blockIdx.z;
  const TA *A = static_cast<const TA *>(p.A) + (int64_t)z * p.strideA;
  const TB *B = static_cast<const TB *>(p.B) + (int64_t)z * p.strideB;
  float *C = p.C + (int64_t)z * p.strideC;
  int tile_m, tile_n;
  tile_coords(blockIdx.x, p.tiles_m, p.tiles_n, tile_m, tile_n);
  const int m0 = tile_m * kMBM, n0 = tile_n * kMBN;

  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wm = wave / kMWN, wn = wave % kMWN;
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
  auto buf = [&](int kt) { return lds + (kt & 1) * kMBuf; };
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
              As + q * PLANE + i * MI * kMRowB + kbyte);
#pragma unroll
      for (int q = 0; q < (AF ? 1 : 3); ++q)
#pragma unroll
        for (int j = 0; j < FN; ++j)
          bf[q][j] = *reinterpret_cast<const vec4u *>(
              Bs + q * PLANE + j * MI * kMRowB + kbyte);
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
  const bool vec_c = reinterpret_cast<uintptr_t>(C) % 16 == 0 && p.ldc % 4 == 0;
  store_tile_lds<Elem::BF16, float, kMBM, kMBN, kMWM, kMWN, MI, kMLds>(
      acc, C, p.ldc, p.M, p.N, m0, n0, (double)p.alpha, 0.0, lds, vec_c);
}

template <bool AF>
hcblasStatus_t launch_mixed(hipStream_t s, MixedParams p, bool a_k, bool b_k,
                            bool vl, int64_t batch) {
  p.tiles_m = (int)cdiv(p.M, kMBM);
  p.tiles_n = (int)cdiv(p.N, kMBN);
  const unsigned tiles = (unsigned)(p.tiles_m * p.tiles_n);
  for (int64_t b0 = 0; b0 < batch; b0 += 65535) {
    MixedParams q = p;
    const int64_t nb = batch - b0 < 65535 ? batch - b0 : 65535;
    q.A = static_cast<const char *>(q.A) +
          b0 * q.strideA * (AF ? sizeof(float) : sizeof(uint16_t));
    q.B = static_cast<const char *>(q.B) +
          b0 * q.strideB * (AF ? sizeof(uint16_t) : sizeof(float));
    q.C += b0 * q.strideC;
    dim3 grid(tiles, 1, (unsigned)nb);
#define HCB_MIXED_CASE(AK, BKC, VLD)                                 \
  if (a_k == AK && b_k == BKC && vl == VLD)                          \
    gemm_mixed_kernel<AF, AK, BKC, VLD><<<grid, 64 * kMWM * kMWN, 0, s>>>(q);
    HCB_MIXED_CASE(false, false, false)
    HCB_MIXED_CASE(false, true, false)
    HCB_MIXED_CASE(true, false, false)
    HCB_MIXED_CASE(true, true, false)
    HCB_MIXED_CASE(false, false, true)
    HCB_MIXED_CASE(false, true, true)
    HCB_MIXED_CASE(true, false, true)
    HCB_MIXED_CASE(true, true, true)
#undef HCB_MIXED_CASE
    const hcblasStatus_t st = launch_status();
    if (st != HCBLAS_STATUS_SUCCESS) return st;
  }
  return HCBLAS_STATUS_SUCCESS;
}

inline bool aligned_ptr(const void *p, size_t esize, int64_t off) {
  return ((reinterpret_cast<uintptr_t>(p) + (uintptr_t)(off * esize)) & 15u) ==
         0;
}

}  // namespace

hcblasStatus_t gemm_mixed_bf16(hcblasHandle_t h, const Canon &c, bool a_f32,
                               float alpha, float *C, int64_t ldc,
                               int64_t strideC, int64_t batch) {
  if (c.M == 0 || c.N == 0 || batch == 0) return HCBLAS_STATUS_SUCCESS;
  MixedParams p{};
  p.M = c.M;
  p.N = c.N;
  p.K = c.K;
  const size_t ea = a_f32 ? 4 : 2, eb = a_f32 ? 2 : 4;
  p.A = static_cast<const char *>(c.A) + c.aOff * ea;
  p.lda = c.lda;
  p.strideA = c.strideA;
  p.B = static_cast<const char *>(c.B) + c.bOff * eb;
  p.ldb = c.ldb;
  p.strideB = c.strideB;
  p.C = C;
  p.ldc = ldc;
  p.strideC = strideC;
  p.alpha = alpha;
  // 16-byte loads: every base, leading dimension and batch stride a whole
  // number of vectors of its type; r-contiguous operands also need whole
  // vectors of rows.
  const int va = 16 / (int)ea, vb = 16 / (int)eb;
  const bool vl = c.lda % va == 0 && c.ldb % vb == 0 &&
                  aligned_ptr(c.A, ea, c.aOff) && aligned_ptr(c.B, eb, c.bOff) &&
                  (batch == 1 || (c.strideA % va == 0 && c.strideB % vb == 0)) &&
                  (c.a_k || c.M % va == 0) && (c.b_k || c.N % vb == 0);
  if (hipSetDevice(h->device) != hipSuccess)
    return HCBLAS_STATUS_EXECUTION_FAILED;
  return a_f32 ? launch_mixed<true>(h->stream, p, c.a_k, c.b_k, vl, batch)
               : launch_mixed<false>(h->stream, p, c.a_k, c.b_k, vl, batch);
}

}  // namespace hcb
