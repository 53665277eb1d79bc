// Generic MFMA GEMM macro-tile kernel for gfx950.
//
// Replaces the VALU register-micro-tile kernels of the reference
// (e.g. sgemm_NN_kernels.cpp:1360-1469 MICRO_TS16XMTS2, :137
// Mini_Batch_M128_N128_K16_TS16XMTS4_MB2; dgemm/hgemm copies of the same).
//
// Problem (column-major storage, after the host has folded row-major order
// into an operand swap):  C[M x N] = alpha * A'[M x K] * B'[K x N] + beta * C
// where A' element (m,k) is A[k + m*lda] when A_K (k contiguous) and
// A[m + k*lda] otherwise; B' element (k,n) is B[k + n*ldb] when B_K and
// B[n + k*ldb] otherwise.
//
// Tiling:
//   - workgroup = 64*WM*WN threads (WM x WN waves), macro tile BM x BN,
//     K step BK; each wave owns a (BM/WM) x (BN/WN) sub-tile made of 16x16
//     MFMA blocks (v_mfma_*_16x16x{4,4,32,32,64}).
//   - LDS holds both operands k-contiguous ([rows][BK] + 16 bytes of padding
//     per row, which makes the 16-byte fragment reads of 16 rows
//     bank-conflict-free), double-buffered: global loads for tile k+1 are in
//     flight in registers while the MFMAs consume tile k, and one barrier per
//     K step separates the two buffers.
//   - a lane reads one 16-byte fragment per operand per chunk; the four lane
//     groups (lane/16) hold four disjoint k sub-ranges, and the same mapping is
//     used for A and B, so the MFMA sums exactly the products of the chunk.
//   - MFMA output layout (16x16, four values per lane): value r of lane l is
//     C(4*(l/16) + r, l%16) for the FP32/INT32-accumulating forms and
//     C(l/16 + 4*r, l%16) for v_mfma_f64_16x16x4f64 (out_row below).
//   - tile order: consecutive workgroup ids are spread over the 8 XCDs by the
//     dispatcher, so ids are remapped to give each XCD a contiguous slice of
//     a GROUP_M-grouped tile order (A and B panels stay in that XCD's L2).
#pragma once


#include "common.hpp"

namespace hcb {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));
typedef int intx4 __attribute__((ext_vector_type(4)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

enum class Elem { F32, F64, F16, BF16, I8 };

template <Elem E>
struct ElemT;
template <>
struct ElemT<Elem::F32> {
  using T = float;
  using Acc = float;
  using AccV = floatx4;
  static constexpr int VEC = 4;    // elements per 16-byte fragment
  static constexpr int STEPS = 4;  // MFMAs per fragment pair
};
template <>
struct ElemT<Elem::F64> {
  using T = double;
  using Acc = double;
  using AccV = doublex4;
  static constexpr int VEC = 2;
  static constexpr int STEPS = 2;
};
template <>
struct ElemT<Elem::F16> {
  using T = uint16_t;  // binary16 bits
  using Acc = float;
  using AccV = floatx4;
  static constexpr int VEC = 8;
  static constexpr int STEPS = 1;
};
template <>
struct ElemT<Elem::BF16> {
  using T = uint16_t;  // bfloat16 bits
  using Acc = float;
  using AccV = floatx4;
  static constexpr int VEC = 8;
  static constexpr int STEPS = 1;
};
template <>
struct ElemT<Elem::I8> {
  using T = int8_t;
  using Acc = int;
  using AccV = intx4;
  static constexpr int VEC = 16;
  static constexpr int STEPS = 1;
};

template <typename To, typename From>
__device__ __forceinline__ To bitcast(const From &f) {
  static_assert(sizeof(To) == sizeof(From), "bitcast size");
  To t;
  __builtin_memcpy(&t, &f, sizeof(To));
  return t;
}

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef int intx16 __attribute__((ext_vector_type(16)));

// Accumulator of one MI x MI MFMA block (MI = 16 or 32).
// 16-byte fragment / staging register as a native vector (the HIP uint4
// struct defeats scalar replacement of the staging arrays).
typedef unsigned int vec4u __attribute__((ext_vector_type(4)));
typedef unsigned int vec2u __attribute__((ext_vector_type(2)));

template <Elem E, int MI>
struct AccOf;
template <>
struct AccOf<Elem::F32, 16> { using type = floatx4; };
template <>
struct AccOf<Elem::F32, 32> { using type = floatx16; };
template <>
struct AccOf<Elem::F64, 16> { using type = doublex4; };
template <>
struct AccOf<Elem::F16, 16> { using type = floatx4; };
template <>
struct AccOf<Elem::F16, 32> { using type = floatx16; };
template <>
struct AccOf<Elem::BF16, 16> { using type = floatx4; };
template <>
struct AccOf<Elem::BF16, 32> { using type = floatx16; };
template <>
struct AccOf<Elem::I8, 16> { using type = intx4; };
template <>
struct AccOf<Elem::I8, 32> { using type = intx16; };

// MFMA step S of a fragment pair for MI x MI blocks:
//   FP32  v_mfma_f32_16x16x4f32 / v_mfma_f32_32x32x2f32 (one float per step)
//   FP64  v_mfma_f64_16x16x4f64                         (one double per step)
//   FP16  v_mfma_f32_16x16x32_f16 / v_mfma_f32_32x32x16_f16
//   BF16  v_mfma_f32_16x16x32_bf16 / v_mfma_f32_32x32x16_bf16
//   INT8  v_mfma_i32_16x16x64_i8 / v_mfma_i32_32x32x32_i8
template <Elem E, int MI, int S>
__device__ __forceinline__ void mfma_step(const vec4u &a, const vec4u &b,
                                          typename AccOf<E, MI>::type &c) {
  if constexpr (E == Elem::F32) {
    const uint32_t ua = S == 0 ? a.x : S == 1 ? a.y : S == 2 ? a.z : a.w;
    const uint32_t ub = S == 0 ? b.x : S == 1 ? b.y : S == 2 ? b.z : b.w;
    if constexpr (MI == 16)
      c = __builtin_amdgcn_mfma_f32_16x16x4f32(__uint_as_float(ua),
                                               __uint_as_float(ub), c, 0, 0, 0);
    else
      c = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(ua),
                                               __uint_as_float(ub), c, 0, 0, 0);
  } else if constexpr (E == Elem::F64) {
    static_assert(MI == 16, "FP64 MFMA is 16x16 only");
    const double av =
        bitcast<double>(S == 0 ? a.xy : a.zw);
    const double bv =
        bitcast<double>(S == 0 ? b.xy : b.zw);
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, c, 0, 0, 0);
  } else if constexpr (E == Elem::F16) {
    if constexpr (MI == 16)
      c = __builtin_amdgcn_mfma_f32_16x16x32_f16(bitcast<halfx8>(a),
                                                 bitcast<halfx8>(b), c, 0, 0, 0);
    else
      c = __builtin_amdgcn_mfma_f32_32x32x16_f16(bitcast<halfx8>(a),
                                                 bitcast<halfx8>(b), c, 0, 0, 0);
  } else if constexpr (E == Elem::BF16) {
    if constexpr (MI == 16)
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bitcast<bf16x8>(a),
                                                  bitcast<bf16x8>(b), c, 0, 0,
                                                  0);
    else
      c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bitcast<bf16x8>(a),
                                                  bitcast<bf16x8>(b), c, 0, 0,
                                                  0);
  } else {
    if constexpr (MI == 16)
      c = __builtin_amdgcn_mfma_i32_16x16x64_i8(bitcast<intx4>(a),
                                                bitcast<intx4>(b), c, 0, 0, 0);
    else
      c = __builtin_amdgcn_mfma_i32_32x32x32_i8(bitcast<intx4>(a),
                                                bitcast<intx4>(b), c, 0, 0, 0);
  }
}

// ---------------------------------------------------------------------------
// Global -> register -> LDS tile loader for one operand.
// ROWS = tile extent in m (or n); the LDS tile is [ROWS][BK] k-contiguous.
// KCONT: the global operand is k-contiguous (element (r,k) at P[k + r*ld]).
// VLOAD: 16-byte global loads are legal (base, ld aligned to the vector).
// ---------------------------------------------------------------------------
template <typename T, int ROWS, int BK, bool KCONT, bool VLOAD, int NTH = 256>
struct TileLoader {
  static constexpr int VEC = 16 / (int)sizeof(T);
  static constexpr int NV = ROWS * BK / VEC;  // 16-byte vectors per tile
  static constexpr int PER_THREAD = (NV + NTH - 1) / NTH;
  static constexpr int ROWB = BK * (int)sizeof(T) + 16;  // LDS row stride
  // Transposing (r-contiguous) tiles: each group of 64 vectors covers
  // 8 r-vectors x 8 k, so a wave reads 128-byte runs of global memory and its
  // element-wise LDS stores spread over the banks (2-way instead of 8-way
  // conflicts for 32 r-vectors of one k).
  static constexpr bool BLOCKED = !KCONT && (ROWS / VEC) % 8 == 0 &&
                                  BK % 8 == 0 && NV % 64 == 0;
  // FP32 transposing tiles with four vectors per thread: each thread loads
  // the four k-lines of one 4-row x 4-k block and writes it with four
  // 16-byte LDS stores along k (a 4 x 4 transpose in its registers) instead
  // of sixteen 4-byte ones. The 16 lanes of a group take 4 consecutive
  // r-vectors x 4 consecutive k-quads, whose rows (4j + e) x ROWB + 16 kq
  // fall on 16 distinct 4-bank groups for both row strides used (528 and
  // 144 bytes); the four groups of a wave continue along r, so a load
  // instruction reads four k-lines in 256-byte runs.
  static constexpr bool TRANS4 =
      !KCONT && VEC == 4 && NV == 4 * NTH && ROWS % 16 == 0 && BK % 16 == 0 &&
      (ROWS / 16) * (BK / 16) * 16 == NTH;
  // 16-bit transposing tiles with four vectors per thread and 64-deep K
  // steps: each thread loads the four k-lines of one 8-row vector and writes
  // its 8 rows x 4 k with eight 8-byte LDS stores along k (16-bit selects
  // pair the k values) instead of thirty-two 2-byte ones. The 32 lanes of a
  // half-wave take 2 consecutive r-vectors x the 16 k-quads, whose slots
  // (8 jr + e) x 144 + 8 kq are 32 distinct 8-byte slots of 256 bytes.
  static constexpr bool TRANS4H = !KCONT && VEC == 8 && BK == 64 &&
                                  NV == 4 * NTH && (ROWS / 8) * 16 == NTH;
  // The same with two vectors per thread (64-row tiles of 4 waves): a
  // thread loads two k-lines of one r-vector and writes its rows with
  // 8-byte (FP32, 32-deep steps) or 4-byte (16-bit, 64-deep steps) stores
  // along k. FP32: wave w takes r-vectors 4w..4w+3, its half-waves r-vectors
  // {0, 2} and {1, 3} of those x the 16 k-pairs (row offsets 0 / 128 and
  // 64 / 192 of 256 bytes); 16-bit: 2 r-vectors x 32 k-pairs per wave.
  static constexpr bool TRANS2 = !KCONT && VEC == 4 && BK == 32 &&
                                 NV == 2 * NTH && ROWS / 4 == NTH / 16;
  static constexpr bool TRANS2H = !KCONT && VEC == 8 && BK == 64 &&
                                  NV == 2 * NTH && (ROWS / 8) * 32 == NTH;
  vec4u reg[PER_THREAD];

  // (r, k) of vector v of the tile (r, k are element indices).
  __device__ __forceinline__ static void vec_pos(int v, int &r, int &k) {
    if (KCONT) {
      r = v / (BK / VEC);
      k = (v % (BK / VEC)) * VEC;
    } else if (TRANS4) {
      const int t = v % NTH, i = v / NTH;
      const int g = t >> 4, jj = t & 3, kk = (t >> 2) & 3;
      constexpr int GJ = ROWS / 16;  // groups of 4 r-vectors along r
      r = (4 * (g % GJ) + jj) * 4;
      k = (4 * (g / GJ) + kk) * 4 + i;
    } else if (TRANS2) {
      const int t = v % NTH, i = v / NTH, l = t & 63;
      r = (4 * (t >> 6) + ((l >> 4) & 1) * 2 + (l >> 5)) * 4;
      k = (l & 15) * 2 + i;
    } else if (TRANS2H) {
      const int t = v % NTH, i = v / NTH;
      r = (t >> 5) * 8;
      k = (t & 31) * 2 + i;
    } else if (TRANS4H) {
      const int t = v % NTH, i = v / NTH;
      r = (t >> 4) * 8;
      k = (t & 15) * 4 + i;
    } else if (BLOCKED) {
      constexpr int NRB = ROWS / VEC / 8;
      const int w = v >> 6, l = v & 63;
      r = ((w % NRB) * 8 + (l & 7)) * VEC;
      k = (w / NRB) * 8 + (l >> 3);
    } else {
      k = v / (ROWS / VEC);
      r = (v % (ROWS / VEC)) * VEC;
    }
  }

  __device__ __forceinline__ void load(const T *__restrict__ P, int64_t ld,
                                       int r0, int k0, int R, int K) {
#pragma unroll
    for (int i = 0; i < PER_THREAD; ++i) {
      const int v = threadIdx.x + i * NTH;
      if (NV % NTH != 0 && v >= NV) continue;
      int r, k;
      vec_pos(v, r, k);
      const int gr = r0 + r, gk = k0 + k;
      if (KCONT) {
        const T *src = P + (int64_t)gr * ld + gk;
        if (VLOAD && gr < R && gk + VEC <= K) {
          reg[i] = *reinterpret_cast<const vec4u *>(src);
        } else {
          T t[VEC];
#pragma unroll
          for (int e = 0; e < VEC; ++e)
            t[e] = (gr < R && gk + e < K) ? src[e] : T(0);
          __builtin_memcpy(&reg[i], t, 16);
        }
      } else {
        const T *src = P + (int64_t)gk * ld + gr;
        if (VLOAD && gk < K && gr + VEC <= R) {
          reg[i] = *reinterpret_cast<const vec4u *>(src);
        } else {
          T t[VEC];
#pragma unroll
          for (int e = 0; e < VEC; ++e)
            t[e] = (gk < K && gr + e < R) ? src[e] : T(0);
          __builtin_memcpy(&reg[i], t, 16);
        }
      }
    }
  }

  // A K step that lies entirely inside [0, K): no K checks, and rows past R
  // load row r0 instead (always inside the operand). Those rows only feed
  // output rows / columns >= M / N, which the epilogue never stores, so they
  // need no zeroing, and the loaded registers are first consumed by the LDS
  // store after the MFMAs: the global latency stays hidden behind them.
  // Requires VLOAD and, for r-contiguous operands, R % VEC == 0 (a vector
  // is then either wholly inside or wholly outside [0, R)).
  __device__ __forceinline__ void load_full(const T *__restrict__ P,
                                            int64_t ld, int r0, int k0,
                                            int R) {
    static_assert(VLOAD, "vector loads only");
#pragma unroll
    for (int i = 0; i < PER_THREAD; ++i) {
      const int v = threadIdx.x + i * NTH;
      if (NV % NTH != 0 && v >= NV) continue;
      int r, k;
      vec_pos(v, r, k);
      const int gr = r0 + r, gk = k0 + k;
      const int rr = gr < R ? gr : r0;
      const T *src = KCONT ? P + (int64_t)rr * ld + gk : P + (int64_t)gk * ld + rr;
      reg[i] = *reinterpret_cast<const vec4u *>(src);
    }
  }

  // Full steps take the branch-free path; the K tail the checked one.
  __device__ __forceinline__ void load_step(const T *__restrict__ P,
                                            int64_t ld, int r0, int k0, int R,
                                            int K, bool full) {
    if constexpr (VLOAD) {
      if (full) {
        load_full(P, ld, r0, k0, R);
        return;
      }
    }
    load(P, ld, r0, k0, R, K);
  }

  __device__ __forceinline__ void store(char *lds) const {
    if constexpr (TRANS4) {
      int r, k;
      vec_pos(threadIdx.x, r, k);  // reg[i] holds rows r..r+3 at k + i
#pragma unroll
      for (int e = 0; e < 4; ++e)
        *reinterpret_cast<vec4u *>(lds + (r + e) * ROWB + k * 4) =
            vec4u{reg[0][e], reg[1][e], reg[2][e], reg[3][e]};
      return;
    }
    if constexpr (TRANS2) {
      int r, k;
      vec_pos(threadIdx.x, r, k);  // reg[i] holds rows r..r+3 at k + i
#pragma unroll
      for (int e = 0; e < 4; ++e)
        *reinterpret_cast<vec2u *>(lds + (r + e) * ROWB + k * 4) =
            vec2u{reg[0][e], reg[1][e]};
      return;
    }
    if constexpr (TRANS2H) {
      int r, k;
      vec_pos(threadIdx.x, r, k);  // reg[i] holds rows r..r+7 at k + i
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int d = e >> 1;
        *reinterpret_cast<uint32_t *>(lds + (r + e) * ROWB + k * 2) =
            (e & 1) ? (reg[0][d] >> 16) | (reg[1][d] & 0xffff0000u)
                    : (reg[0][d] & 0xffffu) | (reg[1][d] << 16);
      }
      return;
    }
    if constexpr (TRANS4H) {
      int r, k;
      vec_pos(threadIdx.x, r, k);  // reg[i] holds rows r..r+7 at k + i
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        // Row r + e is the low (e even) or high half of dword e / 2.
        const int d = e >> 1;
        const uint32_t lo =
            (e & 1) ? (reg[0][d] >> 16) | (reg[1][d] & 0xffff0000u)
                    : (reg[0][d] & 0xffffu) | (reg[1][d] << 16);
        const uint32_t hi =
            (e & 1) ? (reg[2][d] >> 16) | (reg[3][d] & 0xffff0000u)
                    : (reg[2][d] & 0xffffu) | (reg[3][d] << 16);
        *reinterpret_cast<vec2u *>(lds + (r + e) * ROWB + k * 2) =
            vec2u{lo, hi};
      }
      return;
    }
#pragma unroll
    for (int i = 0; i < PER_THREAD; ++i) {
      const int v = threadIdx.x + i * NTH;
      if (NV % NTH != 0 && v >= NV) continue;
      int r, k;
      vec_pos(v, r, k);
      if (KCONT) {
        *reinterpret_cast<vec4u *>(lds + r * ROWB + k * (int)sizeof(T)) =
            reg[i];
      } else {
        T t[VEC];
        __builtin_memcpy(t, &reg[i], 16);
#pragma unroll
        for (int e = 0; e < VEC; ++e)
          *reinterpret_cast<T *>(lds + (r + e) * ROWB + k * (int)sizeof(T)) =
              t[e];
      }
    }
  }
};

// Tile id -> (tile_m, tile_n): XCD-contiguous slices of a GROUP_M-grouped order.
__device__ __forceinline__ void tile_coords(int pid, int tm, int tn,
                                            int &tile_m, int &tile_n) {
  constexpr int kXcd = 8;
  constexpr int GROUP_M = 8;
  const int nt = tm * tn;
  if (nt % kXcd == 0) pid = (pid % kXcd) * (nt / kXcd) + pid / kXcd;
  const int width = GROUP_M * tn;
  const int group = pid / width;
  const int first_m = group * GROUP_M;
  const int gsize = (tm - first_m) < GROUP_M ? (tm - first_m) : GROUP_M;
  const int in_group = pid % width;
  tile_m = first_m + in_group % gsize;
  tile_n = in_group / gsize;
}

// Row (within an MI x MI MFMA block) of accumulator value r of lane group g
// (g = lane / MI; the column is lane % MI):
//   16x16 FP32/INT32 accumulators: 4g + r;   16x16 FP64: g + 4r;
//   32x32 (16 values per lane):    8*(r/4) + 4g + r%4.
template <Elem E, int MI>
__device__ __forceinline__ int out_row(int g, int r) {
  if constexpr (MI == 32)
    return 8 * (r >> 2) + 4 * g + (r & 3);
  else if constexpr (E == Elem::F64)
    return g + 4 * r;
  else
    return 4 * g + r;
}

// Output conversion.
template <typename O, typename Acc>
__device__ __forceinline__ O to_out(Acc v) {
  return (O)v;
}
template <>
__device__ __forceinline__ uint16_t to_out<uint16_t, float>(float v) {
  return bitcast<uint16_t>((_Float16)v);
}
template <typename Acc, typename O>
__device__ __forceinline__ Acc from_out(O v) {
  return (Acc)v;
}
template <>
__device__ __forceinline__ float from_out<float, uint16_t>(uint16_t v) {
  return (float)bitcast<_Float16>(v);
}

__device__ __forceinline__ float fma_acc(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}
__device__ __forceinline__ double fma_acc(double a, double b, double c) {
  return __builtin_fma(a, b, c);
}

struct GemmParams {
  int M, N, K;
  const void *A;
  int64_t lda, strideA;
  const void *B;
  int64_t ldb, strideB;
  void *C;
  int64_t ldc, strideC;
  // Pointer-array batches (device arrays); nullptr selects the strided form.
  const void *const *Aptrs;
  const void *const *Bptrs;
  void *const *Cptrs;
  int64_t aOff, bOff, cOff;
  double alpha, beta;
  int tiles_m, tiles_n;
};

// LDS bytes used by one macro tile (two buffers of A and B).
template <Elem E, int BM, int BN, int BK>
constexpr int tile_lds_bytes() {
  return 2 * (BM + BN) * (BK * (int)sizeof(typename ElemT<E>::T) + 16);
}

// Wave sub-tile geometry of a BM x BN macro tile split over WM x WN waves.
template <Elem E, int BM, int BN, int WM, int WN, int MI>
struct WaveTile {
  static constexpr int WTM = BM / WM, WTN = BN / WN;
  static constexpr int FM = WTM / MI, FN = WTN / MI;
  static constexpr int NVAL = MI * MI / 64;  // accumulator values per block
  static_assert(FM * MI == WTM && FN * MI == WTN, "wave tile");
  using AccV = typename AccOf<E, MI>::type;
};

// Main loop of one BM x BN tile at (m0, n0): acc = A'(m0.., :) * B'(:, n0..)
// over the whole K range. Shared by the plain / batched kernel, the grouped
// kernel and the residue kernel, so all of them produce bit-identical sums
// for the same K step.
template <Elem E, bool A_K, bool B_K, int BM, int BN, int BK, int WM, int WN,
          bool VA, bool VB, int MI, bool KFULL = false>
__device__ __forceinline__ void gemm_mainloop(
    const typename ElemT<E>::T *__restrict__ A, int64_t lda,
    const typename ElemT<E>::T *__restrict__ B, int64_t ldb, int M, int N,
    int K, int m0, int n0, char *lds,
    typename WaveTile<E, BM, BN, WM, WN, MI>::AccV (
        &acc)[WaveTile<E, BM, BN, WM, WN, MI>::FM]
             [WaveTile<E, BM, BN, WM, WN, MI>::FN]) {
  using Tr = ElemT<E>;
  using T = typename Tr::T;
  using WT = WaveTile<E, BM, BN, WM, WN, MI>;
  using AccV = typename WT::AccV;
  constexpr int VEC = Tr::VEC;
  constexpr int NG = 64 / MI;       // lane groups along k
  constexpr int CK = NG * VEC;      // k per fragment chunk
  static_assert(BK % CK == 0, "BK must be a multiple of the chunk");
  constexpr int NTH = 64 * WM * WN;  // workgroup size
  constexpr int WTM = WT::WTM, WTN = WT::WTN, FM = WT::FM, FN = WT::FN;

  using LA = TileLoader<T, BM, BK, A_K, VA, NTH>;
  using LB = TileLoader<T, BN, BK, B_K, VB, NTH>;
  constexpr int ROWB = LA::ROWB;
  constexpr int A_BYTES = BM * ROWB, B_BYTES = BN * ROWB;

  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wm = wave / WN, wn = wave % WN;
  const int lrow = lane % MI, lgrp = lane / MI;

#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = AccV{};

  LA la;
  LB lb;
  const int nk = (K + BK - 1) / BK;
  // Steps with no K tail (KFULL: the caller guarantees K % BK == 0).
  const int nfull = KFULL ? nk : K / BK;

  // MFMAs over one LDS buffer. Small wave tiles (the 64 x 64 tiles of 8
  // waves, at most 2 fragments of each operand, and the 4-wave 64 x 64 FP64
  // tiles, 2 x 2 fragments) read chunk c + 1's fragments into a second
  // register set before chunk c's MFMAs, so the LDS latency is not paid
  // between chunks (1024^3 DGEMM 71.5 against 75.3 us); the larger tiles
  // have no registers to spare for it, and the 4-wave 128 x 128 16-bit tile
  // measured no faster with it.
  constexpr bool PIPE = (WM * WN == 8 && FM * FN <= 2) ||
                        (E == Elem::F64 && WM * WN == 4 && FM * FN == 4);
  // The large 8-wave tiles (256 x 256 and 256 x 128, two waves per SIMD)
  // issue the step's LDS writes and global loads among the first chunk's
  // MFMAs, behind that chunk's fragment reads, instead of before them: both
  // waves of a SIMD leave the barrier together, and the writes would
  // otherwise hold the MFMA pipe idle while they drain and the fragment
  // reads queue behind them (8192^3 SGEMM 3-7% faster over the four
  // transpose forms, DGEMM 6-11%, 3072^3 SGEMM on 256 x 128 tiles 10%).
  // The 4-wave 128 x 128 FP32 tile (64 x 64 per wave on 16 x 16 blocks, two
  // workgroups per CU) gains the same way (2048^3 SGEMM 166 against 187 us),
  // and so does the 8-wave 128 x 64 FP64 tile (2048^3 DGEMM 294 against
  // 344 us); the 4-wave 128 x 128 16-bit tile, with 4 MFMAs per chunk, was
  // 5-18% slower with it.
  constexpr bool ILV = !PIPE && ((WM * WN == 8 && FM * FN >= 4) ||
                                 (WM * WN == 4 && MI == 16 && FM * FN >= 8));
  // 16-bit tiles have 8 MFMAs per chunk, too few to carry a whole step's
  // writes and loads: their four chunks take A's LDS writes, B's, A's loads
  // and B's loads in turn (8192^3 HGEMM 2-4% faster than all in chunk 0).
  constexpr bool SPREAD = ILV && BK / CK == 4 && FM * FN * Tr::STEPS < 16;
  // hook(c): the step's LDS writes and loads, or the part of them that
  // chunk c carries.
  auto compute = [&](const char *cur, auto hook) {
    const char *As = cur;
    const char *Bs = cur + A_BYTES;
    vec4u afp[PIPE ? 2 : 1][FM], bfp[PIPE ? 2 : 1][FN];
    auto frag = [&](int c, int set) {
      const int kbyte = (c * CK + lgrp * VEC) * (int)sizeof(T);
#pragma unroll
      for (int i = 0; i < FM; ++i)
        afp[set][i] = *reinterpret_cast<const vec4u *>(
            As + (wm * WTM + i * MI + lrow) * ROWB + kbyte);
#pragma unroll
      for (int j = 0; j < FN; ++j)
        bfp[set][j] = *reinterpret_cast<const vec4u *>(
            Bs + (wn * WTN + j * MI + lrow) * ROWB + kbyte);
    };
    if constexpr (PIPE) frag(0, 0);
#pragma unroll
    for (int c = 0; c < BK / CK; ++c) {
      const int set = PIPE ? (c & 1) : 0;
      if constexpr (PIPE) {
        if (c + 1 < BK / CK) frag(c + 1, (c + 1) & 1);
      } else {
        frag(c, 0);
      }
      if constexpr (ILV) {
        if (SPREAD || c == 0) hook(c);
      }
      const vec4u(&af)[FM] = afp[set];
      const vec4u(&bf)[FN] = bfp[set];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
          mfma_step<E, MI, 0>(af[i], bf[j], acc[i][j]);
      if constexpr (Tr::STEPS > 1) {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j)
            mfma_step<E, MI, 1>(af[i], bf[j], acc[i][j]);
      }
      if constexpr (Tr::STEPS > 2) {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j)
            mfma_step<E, MI, 2>(af[i], bf[j], acc[i][j]);
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j)
            mfma_step<E, MI, 3>(af[i], bf[j], acc[i][j]);
      }
      if constexpr (PIPE) {
        // The next chunk's LDS reads go out behind the first MFMAs.
        constexpr int NMF = FM * FN * Tr::STEPS, NRD = FM + FN;
#pragma unroll
        for (int i = 0; i < NMF; ++i) {
          __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
          if (i < NRD) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      if constexpr (SPREAD) {
        constexpr int NMF = FM * FN * Tr::STEPS, NRD = FM + FN;
        __builtin_amdgcn_sched_group_barrier(0x100, NRD, 0);  // DS read
#pragma unroll
        for (int i = 0; i < NMF; ++i) {
          __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
          if (c < 2)
            __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);  // DS write
          else
            __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // VMEM read
        }
        __builtin_amdgcn_sched_barrier(0);
      } else if constexpr (ILV) {
        if (c == 0) {
          constexpr int NMF = FM * FN * Tr::STEPS, NRD = FM + FN;
          __builtin_amdgcn_sched_group_barrier(0x100, NRD, 0);  // DS read
          // One LDS write, then one global load, every SP MFMAs (4, or
          // fewer where the chunk is short).
          constexpr int NWR = LA::PER_THREAD + LB::PER_THREAD;
          constexpr int SP = NMF >= 4 * NWR ? 4 : (NMF / NWR >= 2 ? 2 : 1);
#pragma unroll
          for (int i = 0; i < NMF; ++i) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
            if (i % SP == 0)
              __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);  // DS write
            if (i % SP == SP / 2)
              __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // VMEM read
          }
          __builtin_amdgcn_sched_barrier(0);
        }
      }
    }
  };
  auto none = [](int) {};
  auto buf = [&](int kt) { return lds + (kt & 1) * (A_BYTES + B_BYTES); };

  if (nk == 0) return;
  // Pipeline: the registers hold step kt+1 while step kt computes. At the
  // top of step kt they are written to the other LDS buffer (free: every
  // wave passed the previous barrier) and refilled with step kt+2, so the
  // LDS writes overlap this step's MFMAs and one barrier ends the step.
  auto load = [&](int k, bool full) {
    la.load_step(A, lda, m0, k * BK, M, K, full);
    lb.load_step(B, ldb, n0, k * BK, N, K, full);
  };
  auto store = [&](int k) {
    la.store(buf(k));
    lb.store(buf(k) + A_BYTES);
  };
  load(0, nfull > 0);
  store(0);
  if (nk > 1) load(1, 1 < nfull);
  __syncthreads();
  int kt = 0;
  for (; kt + 2 < nfull; ++kt) {
    auto step = [&](int c) {
      if constexpr (SPREAD) {
        if (c == 0) la.store(buf(kt + 1));
        if (c == 1) lb.store(buf(kt + 1) + A_BYTES);
        if (c == 2) la.load_step(A, lda, m0, (kt + 2) * BK, M, K, true);
        if (c == 3) lb.load_step(B, ldb, n0, (kt + 2) * BK, N, K, true);
      } else {
        store(kt + 1);
        load(kt + 2, true);
      }
    };
    if constexpr (ILV) {
      compute(buf(kt), step);
    } else {
      step(0);
      __builtin_amdgcn_sched_barrier(0);
      compute(buf(kt), none);
    }
    __syncthreads();
  }
  for (; kt < nk; ++kt) {
    if (kt + 1 < nk) store(kt + 1);
    if (kt + 2 < nk) load(kt + 2, kt + 2 < nfull);
    compute(buf(kt), none);
    __syncthreads();
  }
}

// Epilogue of a BM x BN tile through LDS. Stored straight from the
// accumulator layout, one store instruction writes one value per lane into
// 16 or 32 different columns, so the tile leaves in BM*BN/64 scattered store
// instructions; for short K loops (batched and grouped problems) that weighs
// as much as the main loop. Here the accumulators go to LDS as [n][m] (m
// contiguous, row stride BM plus 16 bytes), NC columns at a time (as many as
// the main loop's LDS holds), and each thread then blends and stores VE
// consecutive rows of one column with one 16-byte access, a column's BM rows
// by BM / VE consecutive lanes. The arithmetic is the direct form's
// (alpha * acc, then fma(beta, C, .)), so results are bit-identical.
// Accumulator layouts (value r of lane l, block row / column):
//   32x32 FP32/INT32: row 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32;
//   16x16 FP32: row 4 (l / 16) + r, column l % 16;
//   16x16 FP64: row l / 16 + 4 r, column l % 16.
// vec_c: C and ldc allow 16-byte accesses. Expects the LDS free on entry
// (the main loop ends with a barrier) and ends with a barrier.
template <Elem E, typename OutT, int BM, int BN, int WM, int WN, int MI,
          int LDS_BYTES, typename AccArr>
__device__ __forceinline__ void store_tile_lds(const AccArr &acc, OutT *C,
                                               int64_t ldc, int M, int N,
                                               int m0, int n0, double alpha_d,
                                               double beta_d, char *lds,
                                               bool vec_c) {
  using WT = WaveTile<E, BM, BN, WM, WN, MI>;
  using Acc = typename ElemT<E>::Acc;
  constexpr int LROW = BM + 16 / (int)sizeof(Acc);  // elements
  constexpr int NC_FIT = LDS_BYTES / (LROW * (int)sizeof(Acc));
  constexpr int NC = NC_FIT >= BN ? BN
                     : NC_FIT >= 128 ? 128
                     : NC_FIT >= 64  ? 64
                     : NC_FIT >= 32  ? 32
                                     : 16;
  static_assert(NC >= MI && BN % NC == 0 && NC * LROW * (int)sizeof(Acc) <= LDS_BYTES,
                "epilogue column chunk");
  constexpr int VE = 16 / (int)sizeof(OutT);
  constexpr int NTH = 64 * WM * WN;
  constexpr int TPC = BM / VE;  // threads per column
  static_assert(BM % VE == 0 && NTH % TPC == 0, "epilogue shape");
  Acc *t = reinterpret_cast<Acc *>(lds);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wm = wave / WN, wn = wave % WN;
  const int lrow = lane % MI, lgrp = lane / MI;
  const Acc alpha = (Acc)alpha_d, beta = (Acc)beta_d;
#pragma unroll
  for (int c0 = 0; c0 < BN; c0 += NC) {
#pragma unroll
    for (int j = 0; j < WT::FN; ++j) {
      const int cb = wn * WT::WTN + j * MI;  // wave-uniform
      if (cb < c0 || cb >= c0 + NC) continue;
      const int n = cb - c0 + lrow;
#pragma unroll
      for (int i = 0; i < WT::FM; ++i) {
        const int mb = wm * WT::WTM + i * MI;
        if constexpr (MI == 32) {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const auto &a = acc[i][j];
            floatx4 v = {(float)a[4 * q], (float)a[4 * q + 1],
                         (float)a[4 * q + 2], (float)a[4 * q + 3]};
            *reinterpret_cast<floatx4 *>(t + n * LROW + mb + 8 * q +
                                         4 * lgrp) = v;
          }
        } else if constexpr (sizeof(Acc) == 8) {
#pragma unroll
          for (int r = 0; r < 4; ++r)
            t[n * LROW + mb + lgrp + 4 * r] = acc[i][j][r];
        } else {
          *reinterpret_cast<floatx4 *>(t + n * LROW + mb + 4 * lgrp) =
              acc[i][j];
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int cc = 0; cc < NC; cc += NTH / TPC) {
      const int nl = cc + (int)threadIdx.x / TPC;
      const int ml = ((int)threadIdx.x % TPC) * VE;
      const int n = n0 + c0 + nl, m = m0 + ml;
      if (nl < NC && n < N && m < M) {
        OutT *dst = C + (int64_t)n * ldc + m;
        const Acc *src = t + nl * LROW + ml;
        if (vec_c && m + VE <= M) {
          OutT cin[VE], o[VE];
          if (beta != Acc(0)) {
            const vec4u w = *reinterpret_cast<const vec4u *>(dst);
            __builtin_memcpy(cin, &w, 16);
          }
#pragma unroll
          for (int e = 0; e < VE; ++e) {
            Acc v = alpha * src[e];
            if (beta != Acc(0)) v = fma_acc(beta, from_out<Acc>(cin[e]), v);
            o[e] = to_out<OutT>(v);
          }
          vec4u w;
          __builtin_memcpy(&w, o, 16);
          *reinterpret_cast<vec4u *>(dst) = w;
        } else {
          const int ne = M - m < VE ? M - m : VE;
          for (int e = 0; e < ne; ++e) {
            Acc v = alpha * src[e];
            if (beta != Acc(0)) v = fma_acc(beta, from_out<Acc>(dst[e]), v);
            dst[e] = to_out<OutT>(v);
          }
        }
      }
    }
    __syncthreads();
  }
}

// One BM x BN output tile at (m0, n0) of C = alpha*A'*B' + beta*C, the whole
// K range; beta == 0 never reads C.
template <Elem E, typename OutT, bool A_K, bool B_K, int BM, int BN, int BK,
          int WM, int WN, bool VA, bool VB, int MI = 16>
__device__ __forceinline__ void gemm_tile(
    const typename ElemT<E>::T *__restrict__ A, int64_t lda,
    const typename ElemT<E>::T *__restrict__ B, int64_t ldb, OutT *C,
    int64_t ldc, int M, int N, int K, double alpha_d, double beta_d, int m0,
    int n0, char *lds) {
  using WT = WaveTile<E, BM, BN, WM, WN, MI>;
  typename WT::AccV acc[WT::FM][WT::FN];
  gemm_mainloop<E, A_K, B_K, BM, BN, BK, WM, WN, VA, VB, MI>(
      A, lda, B, ldb, M, N, K, m0, n0, lds, acc);
  constexpr int VE = 16 / (int)sizeof(OutT);
  const bool vec_c = reinterpret_cast<uintptr_t>(C) % 16 == 0 && ldc % VE == 0;
  store_tile_lds<E, OutT, BM, BN, WM, WN, MI, tile_lds_bytes<E, BM, BN, BK>()>(
      acc, C, ldc, M, N, m0, n0, alpha_d, beta_d, lds, vec_c);
}

// CHK (pointer-array batches with 16-byte loads): the matrices of a batch
// come from a device array of pointers that the host never sees, so each
// workgroup checks its problem's A and B bases (one problem per blockIdx.z,
// so the branch is uniform) and runs the element-wise loaders for a base
// that is not 16-byte aligned. Offsets and leading dimensions were checked
// on the host.
template <Elem E, typename OutT, bool A_K, bool B_K, int BM, int BN, int BK,
          int WM, int WN, bool VA, bool VB, int MI = 16, bool CHK = false>
__global__ __launch_bounds__(64 * WM * WN) void gemm_mfma_kernel(
    GemmParams p) {
  using T = typename ElemT<E>::T;
  __shared__ __attribute__((aligned(16))) char lds[tile_lds_bytes<E, BM, BN, BK>()];
  const int z = blockIdx.z;
  const T *A;
  const T *B;
  OutT *C;
  if (p.Aptrs)
    A = static_cast<const T *>(p.Aptrs[z]) + p.aOff;
  else
    A = static_cast<const T *>(p.A) + p.aOff + (int64_t)z * p.strideA;
  if (p.Bptrs)
    B = static_cast<const T *>(p.Bptrs[z]) + p.bOff;
  else
    B = static_cast<const T *>(p.B) + p.bOff + (int64_t)z * p.strideB;
  if (p.Cptrs)
    C = static_cast<OutT *>(p.Cptrs[z]) + p.cOff;
  else
    C = static_cast<OutT *>(p.C) + p.cOff + (int64_t)z * p.strideC;
  int tile_m, tile_n;
  tile_coords(blockIdx.x, p.tiles_m, p.tiles_n, tile_m, tile_n);
  if constexpr (CHK) {
    if (((reinterpret_cast<uintptr_t>(A) | reinterpret_cast<uintptr_t>(B)) &
         15u) != 0) {
      gemm_tile<E, OutT, A_K, B_K, BM, BN, BK, WM, WN, false, false, MI>(
          A, p.lda, B, p.ldb, C, p.ldc, p.M, p.N, p.K, p.alpha, p.beta,
          tile_m * BM, tile_n * BN, lds);
      return;
    }
  }
  gemm_tile<E, OutT, A_K, B_K, BM, BN, BK, WM, WN, VA, VB, MI>(
      A, p.lda, B, p.ldb, C, p.ldc, p.M, p.N, p.K, p.alpha, p.beta,
      tile_m * BM, tile_n * BN, lds);
}

// Split-K partial tile: the K range [z ks, min(K, (z + 1) ks)) of one
// tile of a single (unbatched) problem, stored without alpha / beta into
// W + z M N (column-major, ld M). For problems with about as many tiles as
// CUs and a long K, splitting K gives every CU several workgroups to switch
// between; splitk_reduce_kernel then adds the partials in split order.
template <Elem E, bool A_K, bool B_K, int BM, int BN, int BK, int WM, int WN,
          bool VA, bool VB, int MI>
__global__ __launch_bounds__(64 * WM * WN) void gemm_splitk_kernel(
    GemmParams p, int ks, typename ElemT<E>::Acc *W) {
  using T = typename ElemT<E>::T;
  using Acc = typename ElemT<E>::Acc;
  constexpr int LDS_BYTES = tile_lds_bytes<E, BM, BN, BK>();
  __shared__ __attribute__((aligned(16))) char lds[LDS_BYTES];
  const int z = blockIdx.z;
  const int k0 = z * ks;
  const int kl = p.K - k0 < ks ? p.K - k0 : ks;
  const T *A = static_cast<const T *>(p.A) + p.aOff +
               (A_K ? (int64_t)k0 : (int64_t)k0 * p.lda);
  const T *B = static_cast<const T *>(p.B) + p.bOff +
               (B_K ? (int64_t)k0 : (int64_t)k0 * p.ldb);
  int tile_m, tile_n;
  tile_coords(blockIdx.x, p.tiles_m, p.tiles_n, tile_m, tile_n);
  using WT = WaveTile<E, BM, BN, WM, WN, MI>;
  typename WT::AccV acc[WT::FM][WT::FN];
  gemm_mainloop<E, A_K, B_K, BM, BN, BK, WM, WN, VA, VB, MI>(
      A, p.lda, B, p.ldb, p.M, p.N, kl, tile_m * BM, tile_n * BN, lds, acc);
  constexpr int VE = 16 / (int)sizeof(Acc);
  store_tile_lds<E, Acc, BM, BN, WM, WN, MI, LDS_BYTES>(
      acc, W + (int64_t)z * p.M * p.N, p.M, p.M, p.N, tile_m * BM,
      tile_n * BN, 1.0, 0.0, lds, p.M % VE == 0);
}

// C = alpha * sum_z W[z] + beta * C, the partials added in split order.
template <typename Acc, typename OutT>
__global__ __launch_bounds__(256) void splitk_reduce_kernel(
    int S, int M, int N, const Acc *__restrict__ W, OutT *C, int64_t ldc,
    Acc alpha, Acc beta) {
  const int64_t MN = (int64_t)M * N;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < MN;
       i += (int64_t)gridDim.x * 256) {
    Acc s = W[i];
    for (int z = 1; z < S; ++z) s += W[(int64_t)z * MN + i];
    const int64_t n = i / M, m = i - n * M;
    OutT *dst = C + n * ldc + m;
    Acc v = alpha * s;
    if (beta != Acc(0)) v = fma_acc(beta, from_out<Acc>(*dst), v);
    *dst = to_out<OutT>(v);
  }
}

}  // namespace hcb
