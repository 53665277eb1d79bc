// Split-precision GEMM (no counterpart in the reference):
//   - SGEMM / CGEMM as three-way BF16 splits on BF16 MFMA with FP32
//     accumulation (six products: hi*hi, hi*mid, mid*hi, hi*lo, mid*mid,
//     lo*hi);
//   - DGEMM / ZGEMM as Ozaki exact INT8 slices on INT8 MFMA.
// The BF16 split packs both operands into blocked planes and runs the fused
// kernel below (all six products from one set of fragments per K step); the
// INT8 slice form packs k-contiguous planes for the plane-group kernel of
// gemm_planes.hpp.
//
// BF16 split: x = hi + mid + lo + d with hi = bf16(x), mid = bf16(x - hi),
// lo = bf16(x - hi - mid) (round-to-nearest-even; the residuals are exact in
// FP32), so |d| <= 2^-24 |x|; the dropped products (mid*lo, lo*mid, lo*lo and
// d terms) are below 2^-24 relative.
//
// Ozaki slicing (scheme I with power-of-two row / column scaling):
//   e_m = exponent of max_k |A'(m,k)| (|A'(m,k)| < 2^e_m), v = A' * 2^(6-e_m)
//   a_1 = rint(v), a_t = rint(128 * (previous residual)), t = 1..s,
// every a_t in [-64, 64]; A' = 2^(e_m-6) * sum_t a_t 2^(-7(t-1)) + O(2^-7s).
// Same for B' by columns. Products with t + u = d share the weight
// 2^(-7(d-2)); their INT32 sums are exact while s*K*64*64 < 2^31, and K is
// processed in chunks when it is larger. Only d <= s+1 is kept.
// FP64 is used only for the exponent scan, the slicing and the recombination.
//
// Ozaki scheme II (CRT): A' = rint(A * 2^(b - e_m)) and B' likewise are
// integers with |A'|, |B'| <= 2^b. For pairwise coprime moduli m_1..m_n
// (all <= 256, product P) the symmetric residues of A' and B' are INT8, one
// INT8 GEMM per modulus gives A'B' mod m_t exactly (INT32 while
// K * 128 * 128 < 2^31), and the explicit CRT recovers A'B' exactly whenever
// K * 2^(2b) < P/2, which fixes b. n = 16 moduli give b = 55 bits per operand at K = 8192 with 16
// INT8 products, against 36 for eight slices of the scheme above.
#include <climits>
#include <cmath>
#include <vector>

#include "gemm_common.hpp"
#include "gemm_planes.hpp"

namespace hcb {
namespace {

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float x) {
  uint32_t u = __float_as_uint(x);
  uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(r >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t b) {
  return __uint_as_float((uint32_t)b << 16);
}

// Reads element (r, k) of an R x K operand view with the given orientation
// (kcont: element at X[k + r*ld], else X[r + k*ld]).
template <typename T>
__device__ __forceinline__ T read_rk(const T *X, int64_t ld, bool kcont, int r,
                                     int k) {
  return kcont ? X[(int64_t)r * ld + k] : X[(int64_t)k * ld + r];
}

// 32 x 32 (r, k) tile staged through LDS: reads along the source's
// contiguous dimension, hands each thread (r, k) pairs along k.
template <typename T>
struct TileStage {
  T v[32][33];
  __device__ void load(const T *X, int64_t ld, bool kcont, int R, int K,
                       int r0, int k0) {
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int i = ty; i < 32; i += 8) {
      const int r = kcont ? r0 + i : r0 + tx;
      const int k = kcont ? k0 + tx : k0 + i;
      T x = (r < R && k < K) ? read_rk(X, ld, kcont, r, k) : T(0);
      if (kcont)
        v[i][tx] = x;
      else
        v[tx][i] = x;
    }
    __syncthreads();
  }
};

// RT (r) x KT (k) tile of an R x K operand view staged through LDS (row
// stride KT + 1 elements), read along the source's contiguous dimension.
// With vec (the base pointer and ld are multiples of the 16-byte vector)
// every lane reads whole 16-byte vectors; partial vectors at the edges and
// unaligned operands are read element by element, with zeros outside the
// view.
template <typename T, int RT, int KT>
struct TileRK {
  T v[RT][KT + 1];
  template <bool KC>
  __device__ __forceinline__ void load_as(const T *__restrict__ X, int64_t ld,
                                          bool vec, int R, int K, int r0,
                                          int k0) {
    constexpr int VE = 16 / (int)sizeof(T);  // elements per vector
    constexpr int LEN = KC ? KT : RT;        // elements per source line
    constexpr int NL = KC ? RT : KT;         // source lines in the tile
    constexpr int LPL = LEN / VE;            // lanes per line
    constexpr int LPP = 256 / LPL;           // lines per pass
    static_assert(LPL * VE == LEN && LPP * LPL == 256 && NL % LPP == 0,
                  "tile shape");
    const int line0 = threadIdx.x / LPL, c = (threadIdx.x % LPL) * VE;
#pragma unroll
    for (int p = 0; p < NL / LPP; ++p) {
      const int line = line0 + p * LPP;
      const int r = KC ? r0 + line : r0 + c;
      const int k = KC ? k0 + c : k0 + line;
      const T *src = KC ? X + (int64_t)r * ld + k : X + (int64_t)k * ld + r;
      T t[VE];
      const bool full = KC ? (r < R && k + VE <= K) : (k < K && r + VE <= R);
      if (vec && full) {
        // The operand is read once: non-temporal loads.
        const vec4u w =
            __builtin_nontemporal_load(reinterpret_cast<const vec4u *>(src));
        __builtin_memcpy(t, &w, 16);
      } else {
#pragma unroll
        for (int e = 0; e < VE; ++e) {
          const bool in = KC ? (r < R && k + e < K) : (k < K && r + e < R);
          t[e] = in ? src[e] : T(0);
        }
      }
#pragma unroll
      for (int e = 0; e < VE; ++e) {
        if (KC)
          v[line][c + e] = t[e];
        else
          v[c + e][line] = t[e];
      }
    }
  }
  __device__ __forceinline__ void load(const T *__restrict__ X, int64_t ld,
                                       bool kcont, bool vec, int R, int K,
                                       int r0, int k0) {
    if (kcont)
      load_as<true>(X, ld, vec, R, K, r0, k0);
    else
      load_as<false>(X, ld, vec, R, K, r0, k0);
    __syncthreads();
  }
};
template <typename T>
using Tile64 = TileRK<T, 64, 64>;

// True when 16-byte vector reads of an operand with leading dimension ld are
// aligned (the base pointer and every line start).
template <typename T>
inline bool vec_aligned(const T *X, int64_t ld) {
  return (reinterpret_cast<uintptr_t>(X) % 16 == 0) &&
         ((ld * (int64_t)sizeof(T)) % 16 == 0);
}

// ---------------------------------------------------------------------------
// Fused BF16x3 split GEMM.
//
// Planes are packed per 16-wide K block (one 32x32x16 MFMA chunk):
//   P[kb][r][q][0..16)   q = 0 (hi), 1 (mid), 2 (lo),  kb = k / 16,
// so the three planes of one K block of a tile of rows are one contiguous run
// of rows * 96 bytes. K is padded with zeros to a whole number of blocks.
constexpr int kS3K = 16;             // k per block
constexpr int kS3Row = 3 * kS3K;     // elements per row per block

// Splits x into hi + mid + lo (each the BF16 rounding of the remaining
// residual). Non-finite x keeps hi = bf16(x) (NaN stays NaN) and zero
// residual planes, so Inf and NaN propagate as in FP32.
__device__ __forceinline__ void split3(float x, uint16_t &h, uint16_t &m,
                                       uint16_t &l) {
  const uint32_t u = __float_as_uint(x);
  if ((u & 0x7F800000u) == 0x7F800000u) {
    h = (uint16_t)((u >> 16) | ((u & 0x7FFFFFu) ? 0x40u : 0u));
    m = l = 0;
    return;
  }
  h = f32_to_bf16_rne(x);
  const float r1 = x - bf16_to_f32(h);
  m = f32_to_bf16_rne(r1);
  const float r2 = r1 - bf16_to_f32(m);
  l = f32_to_bf16_rne(r2);
}

// Packs an R x K operand view (KC: element (r,k) at X[k + r*ld], else at
// X[r + k*ld]) into the blocked planes above. A workgroup stages a tile
// through LDS with 16-byte reads along the source's contiguous dimension
// (TileRK): 64 rows x 32 k for a k-contiguous operand, 128 rows x 16 k for an
// r-contiguous one (512-byte column pieces). Each thread splits 8
// consecutive k of one row into the LDS image of the tile's output runs (one
// contiguous run of rows x 96 bytes per K block), which then leave as whole
// 16-byte slots in order, 1 KiB per wave instruction. (Each thread writing
// its three 16-byte pieces straight to the planes, 32 of every 96 bytes per
// instruction, and 4-byte reads: 182-211 us per 8192^2 operand.)
template <bool KC>
__global__ __launch_bounds__(256) void split3_pack_kernel(
    int R, int K, const float *__restrict__ X, int64_t ld, bool vec,
    uint16_t *__restrict__ P, int nkb) {
  constexpr int RT = KC ? 64 : 128, KT = KC ? 32 : 16;
  constexpr int GPR = KT / 8;  // threads per row
  __shared__ TileRK<float, RT, KT> st;
  __shared__ vec4u ob[KT / 16][RT][6];  // [K block][row][16-byte slot]
  const int r0 = blockIdx.y * RT, k0 = blockIdx.x * KT;
  const int tid = threadIdx.x;
  st.template load_as<KC>(X, ld, vec, R, K, r0, k0);
  __syncthreads();
  const int row = tid / GPR, kg = tid % GPR;
  uint32_t hv[4], mv[4], lv[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    uint16_t h0, m0, l0, h1, m1, l1;
    split3(st.v[row][kg * 8 + 2 * e], h0, m0, l0);
    split3(st.v[row][kg * 8 + 2 * e + 1], h1, m1, l1);
    hv[e] = (uint32_t)h0 | ((uint32_t)h1 << 16);
    mv[e] = (uint32_t)m0 | ((uint32_t)m1 << 16);
    lv[e] = (uint32_t)l0 | ((uint32_t)l1 << 16);
  }
  // slot 2 q + h of a row: plane q (hi, mid, lo), k half h of the block
  vec4u *o = ob[kg >> 1][row] + (kg & 1);
  o[0] = vec4u{hv[0], hv[1], hv[2], hv[3]};
  o[2] = vec4u{mv[0], mv[1], mv[2], mv[3]};
  o[4] = vec4u{lv[0], lv[1], lv[2], lv[3]};
  __syncthreads();
  const int rows = R - r0 < RT ? R - r0 : RT;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int v = tid + 256 * j;  // KT / 16 blocks x RT rows x 6 = 768
    const int kbl = v / (RT * 6), w = v % (RT * 6);
    const int kb = (k0 >> 4) + kbl;
    if (w / 6 >= rows || kb >= nkb) continue;
    uint16_t *dst = P + ((int64_t)kb * R + r0) * kS3Row + w * 8;
    __builtin_nontemporal_store(ob[kbl][w / 6][w % 6],
                                reinterpret_cast<vec4u *>(dst));
  }
}
void launch_split3_pack(hipStream_t s, int R, int K, const float *X,
                        int64_t ld, bool kcont, uint16_t *P, int nkb) {
  if (kcont)
    split3_pack_kernel<true>
        <<<dim3((unsigned)cdiv(nkb * kS3K, 32), (unsigned)cdiv(R, 64)), 256,
           0, s>>>(R, K, X, ld, vec_aligned(X, ld), P, nkb);
  else
    split3_pack_kernel<false>
        <<<dim3((unsigned)nkb, (unsigned)cdiv(R, 128)), 256, 0, s>>>(
            R, K, X, ld, vec_aligned(X, ld), P, nkb);
}

struct Split3Params {
  int M, N, nkb;
  const uint16_t *Ap;  // [nkb][M][3][16]
  const uint16_t *Bp;  // [nkb][N][3][16]
  float *C;            // column-major M x N
  int64_t ldc;
  float alpha, beta;
  int tiles_m, tiles_n;
};

// K steps (of 16 k) per block accumulator of the fused BF16x3 kernel.
// Measured at 8192^3 on wide-exponent data (sign * 2^u, u ~ U[-20, 20]):
// one chain over all K gives a max error of 6.7e-7 relative to |A||B|,
// above the FP32 kernel's 6.4e-7; folding every 8 steps gives 1.5e-7 at
// 3% more time (6.18 against 6.01 ms).
constexpr int kSplit3Fold = 8;

// Global -> register -> LDS loader of the SKB K blocks of one step of a
// ROWS-row tile. LDS tile: [ROWS][SKB][3][16] plus 16 bytes per row (the row
// stride of 112 or 208 bytes spreads the 16-byte fragment reads of 8
// consecutive rows over distinct banks). Rows past R read row r0 instead;
// they only feed outputs that are never stored.
template <int ROWS, int SKB, int NTH>
struct Split3Loader {
  static constexpr int VPB = ROWS * 6;       // 16-byte vectors per block
  static constexpr int NV = VPB * SKB;
  static constexpr int PER = (NV + NTH - 1) / NTH;
  static constexpr int ROWB = SKB * 96 + 16;
  vec4u reg[PER];

  __device__ __forceinline__ void load(const uint16_t *__restrict__ P, int R,
                                       int r0, int kb) {
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int v = threadIdx.x + i * NTH;
      if (NV % NTH != 0 && v >= NV) continue;
      const int b = v / VPB, w = v % VPB;
      const int row = w / 6, c = w % 6;
      const int gr = r0 + row < R ? r0 + row : r0;
      reg[i] = *reinterpret_cast<const vec4u *>(
          P + ((int64_t)(kb + b) * R + gr) * kS3Row + c * 8);
    }
  }
  __device__ __forceinline__ void store(char *lds) const {
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int v = threadIdx.x + i * NTH;
      if (NV % NTH != 0 && v >= NV) continue;
      const int b = v / VPB, w = v % VPB;
      const int row = w / 6, c = w % 6;
      *reinterpret_cast<vec4u *>(lds + row * ROWB + b * 96 + c * 16) = reg[i];
    }
  }
};

// C = alpha * A'B' + beta * C from the packed planes, all six products of a
// K step issued from one set of fragments (the five smaller products hi*lo,
// mid*mid, lo*hi, hi*mid, mid*hi first, then hi*hi) into a block
// accumulator that restarts from zero every kSplit3Fold K steps and is then
// added into a running sum with one VALU add (round to nearest). A
// v_mfma_f32_32x32x16_bf16 chain loses more than round-to-nearest when a
// long sum keeps adding terms far below the accumulator's magnitude (on
// BF16-exact wide-exponent data, K = 8192, one chain: mean error 16.5 ulp of
// C against 18.3 for the FP32 MFMA chain and 7.1 with a VALU sum every 16
// instructions); short block chains keep the accumulator close to the
// terms it adds.
// Each wave owns a WTM x WTN sub-tile of 32x32 blocks computed transposed
// (MFMA rows = n, columns = m) so that the 32 lanes of a block store 32
// consecutive m of a column-major C column.
template <int BM, int BN, int WM, int WN, int SKB>
__global__ __launch_bounds__(64 * WM * WN) void gemm_split3_kernel(
    Split3Params p) {
  constexpr int NTH = 64 * WM * WN;
  constexpr int WTM = BM / WM, WTN = BN / WN;
  constexpr int FM = WTM / 32, FN = WTN / 32;
  static_assert(FM * 32 == WTM && FN * 32 == WTN, "wave tile");
  using LA = Split3Loader<BM, SKB, NTH>;
  using LB = Split3Loader<BN, SKB, NTH>;
  constexpr int ROWB = LA::ROWB;
  constexpr int A_BYTES = BM * ROWB, B_BYTES = BN * ROWB;
  __shared__ __attribute__((aligned(16))) char lds[2 * (A_BYTES + B_BYTES)];

  int tile_m, tile_n;
  tile_coords(blockIdx.x, p.tiles_m, p.tiles_n, tile_m, tile_n);
  const int m0 = tile_m * BM, n0 = tile_n * BN;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int wm = wave / WN, wn = wave % WN;
  const int l32 = lane & 31, half = lane >> 5;

  floatx16 blk[FM][FN], tot[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) blk[i][j] = tot[i][j] = floatx16{};

  const int nsteps = p.nkb / SKB;  // nkb is a multiple of SKB
  // Two register sets: step j is staged in set j % 2.
  LA la[2];
  LB lb[2];
  auto load = [&](int set, int s) {
    la[set].load(p.Ap, p.M, m0, s * SKB);
    lb[set].load(p.Bp, p.N, n0, s * SKB);
  };
  auto buf = [&](int s) { return lds + (s & 1) * (A_BYTES + B_BYTES); };
  auto store = [&](int set, int s) {
    la[set].store(buf(s));
    lb[set].store(buf(s) + A_BYTES);
  };
  auto compute = [&](const char *cur) {
    const char *As = cur + (wm * WTM + l32) * ROWB + half * 16;
    const char *Bs = cur + A_BYTES + (wn * WTN + l32) * ROWB + half * 16;
#pragma unroll
    for (int b = 0; b < SKB; ++b) {
      vec4u af[3][FM], bf[3][FN];
#pragma unroll
      for (int q = 0; q < 3; ++q) {
#pragma unroll
        for (int i = 0; i < FM; ++i)
          af[q][i] = *reinterpret_cast<const vec4u *>(As + i * 32 * ROWB +
                                                      b * 96 + q * 32);
#pragma unroll
        for (int j = 0; j < FN; ++j)
          bf[q][j] = *reinterpret_cast<const vec4u *>(Bs + j * 32 * ROWB +
                                                      b * 96 + q * 32);
      }
      // (A plane, B plane) of the small products, smallest first.
      constexpr int qa[5] = {0, 1, 2, 0, 1};
      constexpr int qb[5] = {2, 1, 0, 1, 0};
#pragma unroll
      for (int t = 0; t < 5; ++t)
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j)
            blk[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                bitcast<bf16x8>(bf[qb[t]][j]), bitcast<bf16x8>(af[qa[t]][i]),
                blk[i][j], 0, 0, 0);
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
          blk[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
              bitcast<bf16x8>(bf[0][j]), bitcast<bf16x8>(af[0][i]), blk[i][j],
              0, 0, 0);
    }
  };
  auto fold = [&]() {
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        tot[i][j] += blk[i][j];
        blk[i][j] = floatx16{};
      }
  };

  // While step s computes, the registers hold steps s + 1 and s + 2; step
  // s + 1 is written to the other LDS buffer at the top of the step and its
  // set is refilled with step s + 3, which has two steps to arrive. Loads
  // past the last step reload it and stores past it go to the buffer no
  // later step reads, so every step issues the same loads: the wait counts
  // the compiler derives are then the same on every path into the loop.
  const int last = nsteps - 1;
  auto clampk = [&](int s) { return s < last ? s : last; };
  load(0, 0);
  __builtin_amdgcn_sched_barrier(0);
  load(1, clampk(1));
  __builtin_amdgcn_sched_barrier(0);
  store(0, 0);
  __builtin_amdgcn_sched_barrier(0);
  load(0, clampk(2));
  __syncthreads();
  auto step = [&](int s, int set) {  // set = (s + 1) % 2
    store(set, s + 1);
    load(set, clampk(s + 3));
    __builtin_amdgcn_sched_barrier(0);
    compute(buf(s));
    if ((s + 1) % kSplit3Fold == 0) fold();
    __syncthreads();
  };
  int s = 0;
  for (; s + 1 < nsteps; s += 2) {
    step(s, 1);
    step(s + 1, 0);
  }
  if (s < nsteps) step(s, 1);
  fold();

#pragma unroll
  for (int i = 0; i < FM; ++i) {
    const int row = m0 + wm * WTM + i * 32 + l32;
    if (row >= p.M) continue;
#pragma unroll
    for (int j = 0; j < FN; ++j) {
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int col = n0 + wn * WTN + j * 32 + out_row<Elem::BF16, 32>(half, r);
        if (col >= p.N) continue;
        float *dst = p.C + (int64_t)col * p.ldc + row;
        float o = p.alpha * tot[i][j][r];
        if (p.beta != 0.f) o = __builtin_fmaf(p.beta, *dst, o);
        *dst = o;
      }
    }
  }
}

template <int BM, int BN, int WM, int WN, int SKB>
hcblasStatus_t launch_split3(hipStream_t stream, Split3Params p) {
  p.tiles_m = (int)cdiv(p.M, BM);
  p.tiles_n = (int)cdiv(p.N, BN);
  gemm_split3_kernel<BM, BN, WM, WN, SKB>
      <<<p.tiles_m * p.tiles_n, 64 * WM * WN, 0, stream>>>(p);
  return launch_status();
}

// Row markers shared by the exponent / norm scans and their consumers:
// kNoScale for a row without a nonzero (it contributes zeros), kNonFinite for
// a row holding an Inf or NaN (every output it feeds becomes NaN).
constexpr int kNoScale = INT_MIN;
constexpr int kNonFinite = INT_MAX;

// Smallest e with |x| < 2^e for finite nonzero x (frexp gives x = f*2^e,
// 0.5 <= |f| < 1); kNonFinite for Inf / NaN.
__device__ __forceinline__ int exp_bound(double x) {
  if (!__builtin_isfinite(x)) return kNonFinite;
  int e;
  (void)frexp(x, &e);
  return e;
}

__global__ void fill_int_kernel(int *p, int n, int v) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

// Row exponents: exps[r] = max over nonzero x of ilogb(x) + 1, so that
// |x| < 2^exps[r]; rows without a nonzero keep kNoScale, rows with an Inf or
// NaN get kNonFinite (the largest int, so the max keeps it). A workgroup
// scans a 64 x 64 tile; four threads share a row (16 k each).
__global__ __launch_bounds__(256) void row_exp_kernel(
    int R, int K, const double *__restrict__ X, int64_t ld, bool kcont,
    bool vec, int *exps) {
  __shared__ Tile64<double> st;
  const int r0 = blockIdx.y * 64, k0 = blockIdx.x * 64;
  st.load(X, ld, kcont, vec, R, K, r0, k0);
  const int row = threadIdx.x >> 2, kq = (threadIdx.x & 3) * 16;
  int e = kNoScale;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const double x = st.v[row][kq + j];
    if (x != 0.0) e = max(e, exp_bound(x));
  }
  e = max(e, __shfl_xor(e, 1, 64));
  e = max(e, __shfl_xor(e, 2, 64));
  if ((threadIdx.x & 3) == 0 && r0 + row < R && e != kNoScale)
    atomicMax(&exps[r0 + row], e);
}

// Row sums of squares in FP64, in two deterministic stages: partial sums
// part[b][r] over 64-wide k blocks b (a 64 x 64 tile per workgroup), then
// row_norm_shift_kernel adds them in block order. FP32 input: squares are
// exact in FP64 and cannot overflow.
__global__ __launch_bounds__(256) void row_sumsq_kernel(
    int R, int K, const float *__restrict__ X, int64_t ld, bool kcont,
    bool vec, double *part) {
  __shared__ Tile64<float> st;
  const int r0 = blockIdx.y * 64, k0 = blockIdx.x * 64;
  st.load(X, ld, kcont, vec, R, K, r0, k0);
  const int row = threadIdx.x >> 2, kq = (threadIdx.x & 3) * 16;
  double s = 0.0;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const double x = (double)st.v[row][kq + j];
    s = __builtin_fma(x, x, s);
  }
  s += __shfl_xor(s, 1, 64);
  s += __shfl_xor(s, 2, 64);
  if ((threadIdx.x & 3) == 0 && r0 + row < R)
    part[(int64_t)blockIdx.x * R + r0 + row] = s;
}
// Per-row scale for the norm-bounded CRT operands: with s = ||x_r||_2 the
// row is scaled by 2^sigma, sigma the largest integer with
// 2^sigma s + sqrt(kc)/2 <= 2^beta (so ||rint(x_r 2^sigma)||_2 <= 2^beta);
// stored as exps[r] = -sigma, the form the residue and reconstruction
// kernels take with bits = 0 and shift = 0. Zero rows get kNoScale, rows
// with a non-finite sum kNonFinite.
__global__ __launch_bounds__(256) void row_norm_shift_kernel(
    int R, int nblk, const double *part, int beta, int kc, int *exps) {
  // One wave per row: lane l adds blocks l, l + 64, ... in order, then a
  // fixed shuffle tree combines the lanes (the same sum on every run).
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (r >= R) return;
  double s = 0.0;
  for (int b = lane; b < nblk; b += 64) s += part[(int64_t)b * R + r];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane != 0) return;
  if (s == 0.0) {
    exps[r] = kNoScale;
    return;
  }
  if (!__builtin_isfinite(s)) {
    exps[r] = kNonFinite;
    return;
  }
  // Slack of 2^-30 relative covers the rounding of the sum, the square root
  // and the quotient.
  const double room =
      __builtin_ldexp(1.0, beta) - 0.5 * __builtin_sqrt((double)kc) - 1.0;
  const double t = room / (__builtin_sqrt(s) * (1.0 + 0x1p-30));
  int e;
  (void)frexp(t, &e);  // t = f 2^e, 0.5 <= f < 1: floor(log2 t) = e - 1
  exps[r] = -(e - 1);
}

void launch_row_exp(hipStream_t s, int R, int K, const double *X, int64_t ld,
                    bool kcont, int *exps) {
  row_exp_kernel<<<dim3((unsigned)cdiv(K, 64), (unsigned)cdiv(R, 64)), 256, 0,
                   s>>>(R, K, X, ld, kcont, vec_aligned(X, ld), exps);
}

// INT8 slices into planes P[r][q][0..Kp) (q = t, or s-1-t when reverse).
__global__ __launch_bounds__(256) void ozaki_slice_kernel(
    int R, int K, const double *X, int64_t ld, bool kcont, const int *exps,
    int s, int8_t *P, int64_t Kp, bool reverse) {
  __shared__ TileStage<double> st;
  const int r0 = blockIdx.y * 32, k0 = blockIdx.x * 32;
  st.load(X, ld, kcont, R, K, r0, k0);
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int i = ty; i < 32; i += 8) {
    const int r = r0 + i, k = k0 + tx;
    if (r >= R || k >= Kp) continue;
    const int e = exps[r];
    double v = (e == kNoScale || e == kNonFinite)
                   ? 0.0
                   : __builtin_ldexp(st.v[i][tx], 6 - e);
    int8_t *row = P + (int64_t)r * s * Kp + k;
    for (int t = 0; t < s; ++t) {
      const double a = __builtin_rint(v);
      v = (v - a) * 128.0;
      row[(int64_t)(reverse ? s - 1 - t : t) * Kp] = (int8_t)(int)a;
    }
  }
}

template <Elem E, typename OutT, int BM, int BN, int BK, int WM = 2,
          int WN = 2, int MI = 16>
hcblasStatus_t launch_planes(hipStream_t stream, PlaneParams p) {
  p.tiles_m = (int)cdiv(p.M, BM);
  p.tiles_n = (int)cdiv(p.N, BN);
  gemm_planes_kernel<E, OutT, BM, BN, BK, WM, WN, MI>
      <<<p.tiles_m * p.tiles_n, 64 * WM * WN, 0, stream>>>(p);
  return launch_status();
}

// Host view of the batch entries (pointer arrays are read back once).
struct BatchView {
  std::vector<const void *> A, B;
  std::vector<void *> C;
};

hcblasStatus_t batch_pointers(hcblasHandle_t h, const Canon &c, void *C,
                              int64_t strideC, void *const *Cptrs,
                              int64_t batch, size_t esize, BatchView &v) {
  v.A.resize(batch);
  v.B.resize(batch);
  v.C.resize(batch);
  auto fetch = [&](const void *const *dev, std::vector<const void *> &out) {
    return hipMemcpy(out.data(), dev, sizeof(void *) * batch,
                     hipMemcpyDeviceToHost) == hipSuccess;
  };
  if (c.Aptrs) {
    if (!fetch(c.Aptrs, v.A)) return HCBLAS_STATUS_EXECUTION_FAILED;
  }
  if (c.Bptrs) {
    if (!fetch(c.Bptrs, v.B)) return HCBLAS_STATUS_EXECUTION_FAILED;
  }
  if (Cptrs) {
    std::vector<const void *> tmp(batch);
    if (!fetch(reinterpret_cast<const void *const *>(Cptrs), tmp))
      return HCBLAS_STATUS_EXECUTION_FAILED;
    for (int64_t i = 0; i < batch; ++i) v.C[i] = const_cast<void *>(tmp[i]);
  }
  for (int64_t i = 0; i < batch; ++i) {
    if (!c.Aptrs)
      v.A[i] = static_cast<const char *>(c.A) + i * c.strideA * esize;
    if (!c.Bptrs)
      v.B[i] = static_cast<const char *>(c.B) + i * c.strideB * esize;
    if (!Cptrs) v.C[i] = static_cast<char *>(C) + i * strideC * esize;
  }
  (void)h;
  return HCBLAS_STATUS_SUCCESS;
}

// ---------------------------------------------------------------------------
// CRT (Ozaki scheme II) kernels.
// ---------------------------------------------------------------------------
constexpr int kMaxModuli = 20;
// Pairwise coprime, descending: 2^8, 3*5*17, 11*23, 251, 13*19, primes,
// 7*31, primes.
constexpr int kModuli[kMaxModuli] = {256, 255, 253, 251, 247, 241, 239,
                                     233, 229, 227, 223, 217, 211, 199,
                                     197, 193, 191, 181, 179, 173};
__constant__ int d_moduli[kMaxModuli] = {256, 255, 253, 251, 247, 241, 239,
                                         233, 229, 227, 223, 217, 211, 199,
                                         197, 193, 191, 181, 179, 173};

// a^-1 mod m (gcd(a, m) = 1), extended Euclid.
constexpr int inv_mod(int a, int m) {
  int t = 0, nt = 1, r = m, nr = a % m;
  while (nr != 0) {
    const int q = r / nr;
    const int t2 = t - q * nt;
    t = nt;
    nt = t2;
    const int r2 = r - q * nr;
    r = nr;
    nr = r2;
  }
  return t < 0 ? t + m : t;
}

template <int T>
__device__ __forceinline__ constexpr int modulus() {
  constexpr int m[kMaxModuli] = {256, 255, 253, 251, 247, 241, 239,
                                 233, 229, 227, 223, 217, 211, 199,
                                 197, 193, 191, 181, 179, 173};
  return m[T];
}

// Low bytes of four ints packed into one dword.
__device__ __forceinline__ uint32_t pack_bytes(int b0, int b1, int b2, int b3) {
  const uint32_t lo = __builtin_amdgcn_perm((uint32_t)b1, (uint32_t)b0,
                                            0x0c0c0400u);
  const uint32_t hi = __builtin_amdgcn_perm((uint32_t)b3, (uint32_t)b2,
                                            0x0c0c0400u);
  return __builtin_amdgcn_perm(hi, lo, 0x05040100u);
}

// Residue inputs of 16 integers v_j, as unsigned words whose byte expansion
// is congruent to v_j up to a known offset:
//   FP32 (|v| <= 2^30):  lo = v + 2^30                      (v = lo - 2^30)
//   FP64 (|v| <= 2^62):  lo = v mod 2^32, hi = floor(v / 2^32) + 2^30
//                                            (v = lo + hi 2^32 - 2^62)
template <bool WIDE>
struct ResidueIn {
  uint32_t lo[16], hi[WIDE ? 16 : 1];
};

// Byte weights and offset of modulus M: the dot product of the input bytes
// with (2^(8i) mod M) plus k, where k = -offset mod M, is a nonnegative
// integer below 2^19 congruent to v modulo M.
template <int M, bool WIDE>
struct ResidueWeights {
  static constexpr uint32_t w(int shift) {
    uint32_t r = 0;
    for (int i = 0; i < 4; ++i)
      r |= (uint32_t)((1ull << (shift + 8 * i)) % M) << (8 * i);
    return r;
  }
  static constexpr uint32_t clo = w(0), chi = w(32);
  static constexpr uint32_t k =
      (uint32_t)((M - (1ull << (WIDE ? 62 : 30)) % M) % M);
};

// Residue planes of the 16 integers of one row segment: one 16-byte store
// per modulus plane (plane t's vector at row + t * pstride). Each stored byte is
// the symmetric residue of v modulo m_t, in [-(m-1)/2, (m-1)/2] for odd m
// and the two's-complement low byte of v for m = 256. For odd m, s (the
// weighted byte sum, s < 2^19) is exact in float, rint(s * (1/m)) is the
// nearest integer to s/m (the product's error, below 2^-4/m, is smaller than
// the distance 1/(2m) of s/m from a half-integer), and s - q*m is exact.
template <int NMOD, bool WIDE, int T = 0>
__device__ __forceinline__ void write_residues16(const ResidueIn<WIDE> &in,
                                                 int8_t *row, int64_t pstride) {
  if constexpr (T < NMOD) {
    constexpr int M = modulus<T>();
    int q[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      if constexpr (M == 256) {
        q[j] = (int)in.lo[j];  // v and lo agree modulo 2^8
      } else {
        using W = ResidueWeights<M, WIDE>;
        // The accumulator starts at the bits of 2^23, so it ends as the
        // float 2^23 + s (s < 2^20) without a conversion.
        uint32_t acc = W::k + 0x4B000000u;
        if constexpr (WIDE) acc = __builtin_amdgcn_udot4(in.hi[j], W::chi, acc, false);
        acc = __builtin_amdgcn_udot4(in.lo[j], W::clo, acc, false);
        constexpr float inv = 1.0f / (float)M;
        constexpr float c23 = 8388608.0f * inv;  // 2^23 / m as rounded: exact
        // (2^23 + s) * inv - 2^23 * inv is s * inv exactly before the one
        // rounding, so the quotient is rint(fl(s / m)) as above.
        const float qf = __builtin_rintf(
            __builtin_fmaf(__uint_as_float(acc), inv, -c23));
        // 1.5 * 2^23 + s - q m = 1.5 * 2^23 + r is an exact float whose low
        // mantissa byte is r's two's-complement byte (|r| < 2^22).
        const float rf = __builtin_fmaf(-qf, (float)M,
                                        __uint_as_float(acc | 0x400000u));
        q[j] = (int)__float_as_uint(rf);
      }
    }
    vec4u w;
    w.x = pack_bytes(q[0], q[1], q[2], q[3]);
    w.y = pack_bytes(q[4], q[5], q[6], q[7]);
    w.z = pack_bytes(q[8], q[9], q[10], q[11]);
    w.w = pack_bytes(q[12], q[13], q[14], q[15]);
    __builtin_nontemporal_store(w, reinterpret_cast<vec4u *>(row + (int64_t)T * pstride));
    write_residues16<NMOD, WIDE, T + 1>(in, row, pstride);
  }
}

// Residue planes of v = rint(X(r, k) * 2^(bits - e_r)) over k < Kp (Kp a
// multiple of 128; zeros past K), in MFMA fragment order: plane t starts at
// P + t * pstride (pstride = Rpad * Kp, Rpad a multiple of 256), and the
// 32 x 32 block (r / 32, k / 32) is one 1 KiB run holding the 16 bytes of
// (r, 16-k group g) at lane (r % 32) + 32 * (g % 2) -- the operand layout of
// v_mfma_i32_32x32x32_i8, so a wave loads a fragment with one coalesced
// 16-byte load per lane. Rows from R to Rpad are not written. A workgroup
// stages a tile through LDS, reading along the source's contiguous
// dimension: 32 rows x 128 k for a k-contiguous operand (KC, 512-byte row
// pieces), 128 rows x 32 k for an r-contiguous one (512-byte column pieces;
// a 32 x 128 tile read them in 128-byte pieces, 25% slower at 8192^2).
// Each thread then converts 16 consecutive k of one row and writes one
// 16-byte vector per plane, so a wave writes whole 1 KiB runs of each
// plane.
// FP32 input: bits - e_r keeps |v| <= 2^30, and the scaling and rounding are
// exact in float (values below 2^-126 round to 0 either way); FP64 input:
// |v| <= 2^62, split exactly into 32-bit halves in double.
template <typename T, int NMOD, bool KC>
__global__ __launch_bounds__(256) void crt_residue_kernel(
    int R, int K, const T *__restrict__ X, int64_t ld, bool vec,
    const int *__restrict__ exps, int bits, int8_t *__restrict__ P,
    int64_t Kp, int64_t pstride) {
  constexpr bool WIDE = sizeof(T) == 8;
  constexpr int RT = KC ? 32 : 128, KT = KC ? 128 : 32;
  __shared__ TileRK<T, RT, KT> st;
  const int r0 = blockIdx.y * RT, k0 = blockIdx.x * KT;
  const int tid = threadIdx.x;
  st.template load_as<KC>(X, ld, vec, R, K, r0, k0);
  __syncthreads();
  const int row = tid / (KT / 16), kq = (tid % (KT / 16)) * 16;
  const int r = r0 + row;
  if (r >= R) return;
  const int e = exps[r];
  const bool zero = e == kNoScale || e == kNonFinite;
  ResidueIn<WIDE> in;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    if constexpr (WIDE) {
      const double v =
          zero ? 0.0
               : __builtin_rint(__builtin_ldexp(st.v[row][kq + j], bits - e));
      const double h = __builtin_floor(v * 0x1p-32);
      in.lo[j] = (uint32_t)__builtin_fma(-h, 0x1p32, v);
      in.hi[j] = (uint32_t)(int)h + (1u << 30);
    } else {
      const float v =
          zero ? 0.0f
               : __builtin_rintf(__builtin_ldexpf(st.v[row][kq + j], bits - e));
      in.lo[j] = (uint32_t)(int)v + (1u << 30);
    }
  }
  const int k = k0 + kq;
  const int64_t blk = (int64_t)(r >> 5) * (Kp >> 5) + (k >> 5);
  const int lane = (r & 31) + 32 * ((k >> 4) & 1);
  write_residues16<NMOD, WIDE>(in, P + blk * 1024 + lane * 16, pstride);
}

struct CrtParams {
  int M, N;
  int64_t Kp;
  int nmod;
  const int8_t *Ap, *Bp;  // residue planes in fragment order (plane t at
                          // Ap + t * astride, Bp + t * bstride)
  int64_t astride, bstride;
  uint8_t *R;  // residues of A'B' mod m_t in [0, m_t), in 32 x 32 blocks
               // (crt_store_blocks); plane t at R + t * Mb * Nb * 1024
  int Mb, Nb;  // blocks along m and n
  int tiles_m, tiles_n;
};

// Residues modulo `mod` of the 32 x 32 blocks a wave holds, written in the
// accumulator's own layout: block (bm, bn) of a plane is the 1 KiB run at
// (bn * Mb + bm) * 1024, lane l's 16 bytes being rows 8 (r / 4) + 4 (l / 32)
// + r % 4 (r = 0..15) of column l % 32 of the block -- the output layout of
// v_mfma_i32_32x32x32_i8, so each block leaves as one coalesced 16-byte
// store per lane, with no LDS pass and no barrier. Blocks past the matrix
// (bm >= Mb or bn >= Nb) are not written.
// a mod m: |a| <= K * 128^2 <= 2^30, so q = floor(float(a) / m) is within
// one of the true quotient and |q| < 2^23 (24-bit multiply).
template <int FM, int FN, typename AccArr>
__device__ __forceinline__ void crt_store_blocks(const AccArr &acc, int mod,
                                                 uint8_t *Rt, int Mb, int Nb,
                                                 int bm0, int bn0) {
  const int lane = threadIdx.x & 63;
  const float inv = 1.0f / (float)mod;
  auto red = [&](int a) {
    const int q = (int)__builtin_floorf((float)a * inv);
    int u = a - __mul24(q, mod);
    u += u < 0 ? mod : 0;
    u -= u >= mod ? mod : 0;
    return u;
  };
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int bm = bm0 + i, bn = bn0 + j;
      if (bm >= Mb || bn >= Nb) continue;
      vec4u w;
      uint32_t *wd = reinterpret_cast<uint32_t *>(&w);
#pragma unroll
      for (int g = 0; g < 4; ++g)
        wd[g] = pack_bytes(red(acc[i][j][4 * g]), red(acc[i][j][4 * g + 1]),
                           red(acc[i][j][4 * g + 2]),
                           red(acc[i][j][4 * g + 3]));
      __builtin_nontemporal_store(
          w, reinterpret_cast<vec4u *>(Rt + ((int64_t)bn * Mb + bm) * 1024 +
                                       lane * 16));
    }
}

// Main loop of the residue GEMM: a 256 x 256 tile of one modulus' product
// over INT8 planes whose length Kp is a multiple of 256 (zero padded), 4
// waves, one per SIMD, each with 16 blocks of v_mfma_i32_32x32x32_i8 (256
// accumulator registers). Both operands come from fragment-ordered planes
// (crt_residue_kernel).
//
// A goes through LDS; each wave loads its B fragments itself, one 1 KiB
// coalesced load per fragment. The waves split the tile 1 x 4: each takes
// all 256 rows and 64 columns, so every B fragment is loaded by one wave
// only, and a 128-deep K step moves 64 KiB through the vector memory path
// (32 KiB of A staging, 32 KiB of B) and reads 128 KiB from LDS. A 2 x 2
// split (128 x 128 per wave) halves the LDS reads but loads every B
// fragment twice (96 KiB per step); measured at 8192^3 it is 3% slower,
// with the texture data path busy 85% of the time. Latency is hidden
// inside the wave:
//   - the A fragments of chunk c + 1 are read while the MFMAs of chunk c
//     run (two A fragment sets);
//   - the B fragments of chunk c of step kt + 1 are loaded right after the
//     MFMAs of chunk c of step kt, a whole step ahead of their use;
//   - A tiles are staged in registers two steps ahead (two staging sets):
//     the tile of step kt + 1 is written to the other LDS buffer during
//     chunk 0 of step kt and its set is refilled with step kt + 3 during
//     chunk 1 (2-3% faster than one set, whose loads had three chunks to
//     arrive);
//   - the one barrier of a step comes before the MFMAs of its last chunk,
//     followed by the A reads of the next step's first chunk.
// Rows past M or N are inside the padded planes (never written, so any
// bytes); they only feed outputs the epilogue never stores.
#ifndef HCB_CRT_WN
#define HCB_CRT_WN 4  // waves along n (4: 1 x 4, 2: 2 x 2)
#endif
constexpr int kCrtBM = 256, kCrtBK = 128, kCrtRowB = kCrtBK + 16;
constexpr int kCrtBuf = kCrtBM * kCrtRowB;  // the A tile of one step
constexpr int kCrtLds = 2 * kCrtBuf;
// Waves of the tile: kCrtWM x kCrtWN, each with FM x FN 32 x 32 blocks.
constexpr int kCrtWN = HCB_CRT_WN, kCrtWM = 4 / kCrtWN;
constexpr int kCrtFM = 8 / kCrtWM, kCrtFN = 8 / kCrtWN;

using CrtAcc = intx16[kCrtFM][kCrtFN];

__device__ __forceinline__ void crt_mainloop(const int8_t *__restrict__ Af,
                                             const int8_t *__restrict__ Bf,
                                             int Kp, int m0, int n0,
                                             char *lds, CrtAcc &acc) {
  constexpr int NTH = 256, PT = kCrtBM * (kCrtBK / 16) / NTH;  // 8
  constexpr int FM = kCrtFM, FN = kCrtFN;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int wm = wave / kCrtWN, wn = wave % kCrtWN;
  const int lrow = lane & 31, lgrp = lane >> 5;
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = intx16{};
  // A loader: a K step of the tile is 8 row blocks x 4 chunks of 1 KiB
  // fragments; load i of wave w is fragment (row block i, chunk w), lane l
  // holding row 32 i + l % 32, bytes 32 w + 16 (l / 32) of the step.
  const int64_t fstep = (int64_t)Kp * 32;  // bytes per row block of a plane
  const int8_t *ga =
      Af + (int64_t)(m0 >> 5) * fstep + wave * 1024 + lane * 16;
  const int soff = lrow * kCrtRowB + wave * 32 + lgrp * 16;
  // Two staging sets: tile j of A goes through set j % 2 and is loaded
  // seven chunks before the chunk that stores it.
  vec4u sa[2][PT];
  // h: half of the tile's loads / stores (rows 0-127 or 128-255), or -1
  // for all of them.
  auto gload = [&](vec4u(&st)[PT], int k0, int h) {  // k0: step's byte offset
#pragma unroll
    for (int i = 0; i < PT; ++i)
      if (h < 0 || i / (PT / 2) == h)
        st[i] = *reinterpret_cast<const vec4u *>(ga + i * fstep + k0 * 32);
  };
  auto store_a = [&](const vec4u(&st)[PT], char *b, int h) {
#pragma unroll
    for (int i = 0; i < PT; ++i)
      if (h < 0 || i / (PT / 2) == h)
        *reinterpret_cast<vec4u *>(b + soff + 32 * i * kCrtRowB) = st[i];
  };
  auto buf = [&](int kt) { return lds + (kt & 1) * kCrtBuf; };
  const int aoff = (wm * FM * 32 + lrow) * kCrtRowB + lgrp * 16;
  vec4u fa[2][FM];
  auto frags = [&](vec4u(&a)[FM], const char *cur, int c) {
#pragma unroll
    for (int i = 0; i < FM; ++i)
      a[i] = *reinterpret_cast<const vec4u *>(cur + aoff + i * 32 * kCrtRowB +
                                              c * 32);
  };
  // B fragment (block j, chunk q) at bbase + j * jstride + q * 1024. The
  // fragments of chunk c of step kt + 1 replace those of step kt right
  // after the chunk's MFMAs.
  const int64_t jstride = (int64_t)Kp * 32;
  const int8_t *bbase =
      Bf + (int64_t)((n0 + wn * FN * 32) >> 5) * jstride + lane * 16;
  vec4u fbX[4][FN];
  auto bload = [&](vec4u(&fb)[4][FN], int kt, int c) {
#pragma unroll
    for (int j = 0; j < FN; ++j)
      fb[c][j] = *reinterpret_cast<const vec4u *>(
          bbase + j * jstride + (int64_t)(4 * kt + c) * 1024);
  };
  auto mfmas = [&](const vec4u(&a)[FM], const vec4u(&b)[FN]) {
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(
            bitcast<intx4>(a[i]), bitcast<intx4>(b[j]), acc[i][j], 0, 0, 0);
  };
  const int nk = Kp / kCrtBK;
  if (nk == 0) return;
  // The prologue issues its loads in the order the loop does (tile 1, B
  // chunk 0, first half of tile 2, B chunk 1, second half, B chunks 2-3),
  // unconditionally (tiles past the planes' length reload tile 0 and are
  // never stored), so at the loop head every register has at least as many
  // younger loads outstanding on entry as on the back edge. The wait counts
  // the compiler derives for the loop from the merge of the two paths are
  // then the loop's own; a skipped load on either path would make it wait
  // for loads issued a chunk earlier.
  const int k2 = nk > 2 ? 2 * kCrtBK : 0;
  gload(sa[0], 0, -1);
  __builtin_amdgcn_sched_barrier(0);
  gload(sa[1], nk > 1 ? kCrtBK : 0, -1);
  __builtin_amdgcn_sched_barrier(0);
  bload(fbX, 0, 0);
  __builtin_amdgcn_sched_barrier(0);
  store_a(sa[0], buf(0), -1);
  __builtin_amdgcn_sched_barrier(0);
  gload(sa[0], k2, 0);
  __builtin_amdgcn_sched_barrier(0);
  bload(fbX, 0, 1);
  __builtin_amdgcn_sched_barrier(0);
  gload(sa[0], k2, 1);
  __builtin_amdgcn_sched_barrier(0);
  bload(fbX, 0, 2);
  __builtin_amdgcn_sched_barrier(0);
  bload(fbX, 0, 3);
  __builtin_amdgcn_sched_barrier(0);
  __syncthreads();
  frags(fa[0], buf(0), 0);
  // Issue order inside a chunk of FM x FN = 16 MFMAs: the LDS reads after
  // the first MFMAs, the LDS writes and global loads after the last ones,
  // so the MFMA pipe stays busy while they are sent. The closing
  // sched_barrier keeps each chunk's instructions in its own region, so
  // loads are issued in the order they are needed.
  auto interleave = [](int reads, int writes, int vmem) {
    for (int i = 0; i < FM * FN; ++i) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
      if (i < reads) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
      if (i >= FM * FN - writes)
        __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);
      if (i >= FM * FN - vmem)
        __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
    }
    __builtin_amdgcn_sched_barrier(0);
  };
  // Step kt with B fragments fb and A staging set sn (holding tile kt + 1,
  // refilled with tile kt + 3). Every step runs this schedule: loads of
  // steps past the last one reload the last step's data, which is never
  // used (the tile they stage goes to the LDS buffer no later step reads),
  // so no condition splits a chunk's scheduling region.
  const int last = nk - 1;
  auto steady = [&](int kt, vec4u(&fb)[4][FN], vec4u(&sn)[PT]) {
    const char *cur = buf(kt);
    char *nxt = buf(kt + 1);
    const int kb = kt + 1 < last ? kt + 1 : last;
    const int ka = kt + 3 < last ? kt + 3 : last;
    // The next tile's LDS writes and its set's refill are spread over
    // chunks 0-2, half a tile at a time, so no chunk issues more than 12
    // LDS instructions.
    frags(fa[1], cur, 1);
    store_a(sn, nxt, 0);
    mfmas(fa[0], fb[0]);
    bload(fb, kb, 0);
    interleave(FM, PT / 2, FN);
    frags(fa[0], cur, 2);
    store_a(sn, nxt, 1);
    gload(sn, ka * kCrtBK, 0);
    mfmas(fa[1], fb[1]);
    bload(fb, kb, 1);
    interleave(FM, PT / 2, PT / 2 + FN);
    frags(fa[1], cur, 3);
    gload(sn, ka * kCrtBK, 1);
    mfmas(fa[0], fb[2]);
    bload(fb, kb, 2);
    interleave(FM, 0, PT / 2 + FN);
    __syncthreads();
    frags(fa[0], nxt, 0);
    mfmas(fa[1], fb[3]);
    bload(fb, kb, 3);
    interleave(FM, 0, FN);
  };
  int kt = 0;
  do {  // nk is even and positive
    steady(kt, fbX, sa[1]);
    steady(kt + 1, fbX, sa[0]);
    kt += 2;
  } while (kt < nk);
}

// One INT8 GEMM per modulus (blockIdx.z = t) with the 4-wave main loop.
__global__ __launch_bounds__(256) void crt_gemm4_kernel(CrtParams p) {
  __shared__ __attribute__((aligned(16))) char lds[kCrtLds];
  const int t = blockIdx.z;
  const int mod = d_moduli[t];
  int tm, tn;
  tile_coords(blockIdx.x, p.tiles_m, p.tiles_n, tm, tn);
  CrtAcc acc;
  crt_mainloop(p.Ap + t * p.astride, p.Bp + t * p.bstride, (int)p.Kp,
               tm * kCrtBM, tn * kCrtBM, lds, acc);
  const int wave = threadIdx.x >> 6;
  crt_store_blocks<kCrtFM, kCrtFN>(
      acc, mod, p.R + (int64_t)t * p.Mb * p.Nb * 1024, p.Mb, p.Nb,
      (tm * kCrtBM >> 5) + (wave / kCrtWN) * kCrtFM,
      (tn * kCrtBM >> 5) + (wave % kCrtWN) * kCrtFN);
}

// Reconstruction by the explicit CRT. With P = prod m_t, M_t = P / m_t and
// w_t = M_t * (M_t^-1 mod m_t) (< P), the integer C with C = u_t (mod m_t)
// and |C| < P/2 is C = S - q P, S = sum_t u_t w_t, q = rint(S / P) (S < 20 *
// 256 * P, so q < 2^13). Every w_t and P are split into 32-bit parts
// (x = sum_j x_j 2^(32 j)); then
//   S_j = sum_t u_t w_tj < 2^46   and   D_j = S_j - q P_j   (|D_j| < 2^46)
// are exact in FP64, q comes from sum_j S_j (2^(32 j) / P) (absolute error
// below 2^-38, while |C| / P <= 0.4965 by the choice of the operand bits),
// and C = sum_j D_j 2^(32 j) is summed by Horner's rule in double-double, so
// the returned double is C correctly rounded except within 2^-105 of a
// rounding boundary, and exactly representable values come out exact.
struct CrtBig {
  uint32_t d[6];  // little-endian 32-bit limbs (P < 2^160 for 20 moduli)
};
constexpr CrtBig crt_product(int n) {
  CrtBig p{};
  p.d[0] = 1;
  for (int t = 0; t < n; ++t) {
    uint64_t carry = 0;
    for (int i = 0; i < 6; ++i) {
      const uint64_t x = (uint64_t)p.d[i] * (uint64_t)kModuli[t] + carry;
      p.d[i] = (uint32_t)x;
      carry = x >> 32;
    }
  }
  return p;
}
constexpr CrtBig big_div_small(CrtBig a, uint32_t m) {
  uint64_t rem = 0;
  for (int i = 5; i >= 0; --i) {
    const uint64_t cur = (rem << 32) | a.d[i];
    a.d[i] = (uint32_t)(cur / m);
    rem = cur % m;
  }
  return a;
}
constexpr uint32_t big_mod_small(const CrtBig &a, uint32_t m) {
  uint64_t rem = 0;
  for (int i = 5; i >= 0; --i) rem = ((rem << 32) | a.d[i]) % m;
  return (uint32_t)rem;
}
constexpr CrtBig big_mul_small(CrtBig a, uint32_t m) {
  uint64_t carry = 0;
  for (int i = 0; i < 6; ++i) {
    const uint64_t x = (uint64_t)a.d[i] * m + carry;
    a.d[i] = (uint32_t)x;
    carry = x >> 32;
  }
  return a;
}
constexpr int crt_parts(int n) {
  const CrtBig p = crt_product(n);
  int top = 5;
  while (top > 0 && p.d[top] == 0) --top;
  return top + 1;
}
template <int NMOD>
struct CrtRecon {
  static constexpr int NP = crt_parts(NMOD);
  double w[NMOD][NP];  // 32-bit parts of w_t
  double p[NP];        // 32-bit parts of P
  double f[NP];        // 2^(32 j) / P, rounded
};
template <int NMOD>
constexpr CrtRecon<NMOD> crt_recon() {
  CrtRecon<NMOD> c{};
  constexpr int NP = CrtRecon<NMOD>::NP;
  const CrtBig P = crt_product(NMOD);
  double pd = 0.0;
  for (int i = 5; i >= 0; --i) pd = pd * 4294967296.0 + (double)P.d[i];
  double scale = 1.0;
  for (int j = 0; j < NP; ++j) {
    c.p[j] = (double)P.d[j];
    c.f[j] = scale / pd;
    scale *= 4294967296.0;
  }
  for (int t = 0; t < NMOD; ++t) {
    const uint32_t m = (uint32_t)kModuli[t];
    const CrtBig Mt = big_div_small(P, m);
    const uint32_t y = (uint32_t)inv_mod((int)big_mod_small(Mt, m), (int)m);
    const CrtBig w = big_mul_small(Mt, y);
    for (int j = 0; j < NP; ++j) c.w[t][j] = (double)w.d[j];
  }
  return c;
}

// C from its residues u_t in [0, m_t).
template <int NMOD>
__device__ __forceinline__ double crt_value(const int *u) {
#pragma clang fp contract(off)
  constexpr CrtRecon<NMOD> c = crt_recon<NMOD>();
  constexpr int NP = CrtRecon<NMOD>::NP;
  double S[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) S[j] = 0.0;
#pragma unroll
  for (int t = 0; t < NMOD; ++t) {
    const double ut = (double)u[t];
#pragma unroll
    for (int j = 0; j < NP; ++j) S[j] = __builtin_fma(ut, c.w[t][j], S[j]);
  }
  double qf = 0.0;
#pragma unroll
  for (int j = 0; j < NP; ++j) qf = __builtin_fma(S[j], c.f[j], qf);
  const double q = __builtin_rint(qf);
  if constexpr (NP == 2) {
    // D_1 2^32 + D_0 with exact D_j: one rounding gives C correctly rounded.
    return __builtin_fma(__builtin_fma(-q, c.p[1], S[1]), 4294967296.0,
                         __builtin_fma(-q, c.p[0], S[0]));
  }
  double hi = __builtin_fma(-q, c.p[NP - 1], S[NP - 1]), lo = 0.0;
#pragma unroll
  for (int j = NP - 2; j >= 0; --j) {
    const double dj = __builtin_fma(-q, c.p[j], S[j]);
    hi *= 4294967296.0;
    lo *= 4294967296.0;
    const double sm = hi + dj;  // two-sum: hi + dj = sm + e
    const double bb = sm - hi;
    const double e = (hi - (sm - bb)) + (dj - bb);
    const double l = lo + e;
    hi = sm + l;  // renormalise
    lo = l - (hi - sm);
  }
  return hi + lo;
}

// alpha * 2^(re + ce - shift) * x, or NaN for a row / column marked
// non-finite, or 0 for an all-zero row / column.
__device__ __forceinline__ double crt_scaled(double x, int re, int ce,
                                             int shift, double alpha) {
  if (re == kNonFinite || ce == kNonFinite) return __builtin_nan("");
  if (re == kNoScale || ce == kNoScale) return 0.0;
  return alpha * __builtin_ldexp(x, re + ce - shift);
}
// o = v + beta * c in the output precision (FP32: v rounded once, then the
// beta term added in FP32).
template <typename OutT>
__device__ __forceinline__ OutT crt_out(double v, double beta, OutT c) {
  if constexpr (sizeof(OutT) == 8) {
    return beta != 0.0 ? __builtin_fma(beta, (double)c, v) : v;
  } else {
    const float o = (float)v;
    return beta != 0.0 ? __builtin_fmaf((float)beta, (float)c, o) : o;
  }
}

// C(m, n) = alpha * 2^(e_m + f_n - shift) * CRT(u(m, n)) + beta * C(m, n),
// the residues u read from the 32 x 32 blocks crt_store_blocks writes.
// Every block's outputs are staged through an LDS tile [column][row] (row
// stride 32 + VE elements, VE = elements per 16 bytes) and leave with
// 16-byte stores covering whole 128-byte runs of columns. (Written straight
// from the accumulator layout, each store touched 32 columns with 16-byte
// pieces: 2.3x slower.)
//   - Up to 12 moduli (crt_combine_waves): one wave per block; lane l loads
//     its 16 bytes of the block from every plane (one coalesced 16-byte
//     load per plane) and reconstructs rows 8 q + 4 (l / 32) + e (q, e =
//     0..3) of column l % 32. 8 moduli at 8192^2: 150-158 us, against
//     171-184 for the former [n][m] layout's four rows per thread.
//   - More moduli (crt_combine_groups): one workgroup per block, thread t
//     taking dword t % 4 of lane t / 4's 16 bytes (each plane one contiguous
//     1 KiB run), which keeps the registers of the 16-modulus FP64 form low
//     (whole 16-byte slots per lane cost it occupancy: 443-465 us against
//     420-485 this way).
template <int NMOD, typename OutT>
__device__ __forceinline__ void crt_combine_rows(
    const uint32_t *w, int m, int n, int M, const int *__restrict__ ea, int ce,
    int shift, double alpha, double beta, const OutT *C, int64_t ldc,
    OutT *trow) {
  constexpr int VE = 16 / (int)sizeof(OutT);
  int er[4];
  if (m + 4 <= M) {
    const int4 e4 = *reinterpret_cast<const int4 *>(ea + m);
    er[0] = e4.x, er[1] = e4.y, er[2] = e4.z, er[3] = e4.w;
  } else {
#pragma unroll
    for (int e = 0; e < 4; ++e) er[e] = m + e < M ? ea[m + e] : kNoScale;
  }
  OutT o[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    int u[NMOD];
#pragma unroll
    for (int i = 0; i < NMOD; ++i) u[i] = (w[i] >> (8 * e)) & 0xff;
    const double v = crt_scaled(crt_value<NMOD>(u), er[e], ce, shift, alpha);
    o[e] = crt_out<OutT>(
        v, beta, beta != 0.0 && m + e < M ? C[(int64_t)n * ldc + m + e] : OutT(0));
  }
#pragma unroll
  for (int h = 0; h < 4 / VE; ++h) {
    vec4u cv;
    __builtin_memcpy(&cv, o + h * VE, 16);
    *reinterpret_cast<vec4u *>(trow + h * VE) = cv;
  }
}
// The LDS tile of block (bm, bn) to C, threads t0, t0 + step, ... .
template <typename OutT>
__device__ __forceinline__ void crt_tile_out(const OutT *tile, int bm, int bn,
                                             int M, int N, OutT *C,
                                             int64_t ldc, bool vec_c, int t0,
                                             int step) {
  constexpr int VE = 16 / (int)sizeof(OutT), LS = 32 + VE;
  constexpr int LPC = 32 / VE;  // threads per column run
  for (int v = t0; v < 32 * LPC; v += step) {
    const int cc = v / LPC, rv = (v % LPC) * VE;
    const int nn = bn * 32 + cc, mm = bm * 32 + rv;
    if (nn >= N || mm >= M) continue;
    OutT *dst = C + (int64_t)nn * ldc + mm;
    const OutT *src = tile + cc * LS + rv;
    if (vec_c && mm + VE <= M) {
      __builtin_nontemporal_store(*reinterpret_cast<const vec4u *>(src),
                                  reinterpret_cast<vec4u *>(dst));
    } else {
      for (int e = 0; e < VE && mm + e < M; ++e) dst[e] = src[e];
    }
  }
}
template <int NMOD, typename OutT>
__global__ __launch_bounds__(256) void crt_combine_waves(
    int M, int N, int Mb, int nblk, const uint8_t *__restrict__ R,
    const int *__restrict__ ea, const int *__restrict__ eb, int shift,
    double alpha, double beta, OutT *C, int64_t ldc, bool vec_c) {
  constexpr int LS = 32 + 16 / (int)sizeof(OutT);
  __shared__ __attribute__((aligned(16))) OutT tile[4][32 * LS];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int blk = blockIdx.x * 4 + wave;
  if (blk >= nblk) return;  // wave-uniform; no barrier follows
  const int bm = blk % Mb, bn = blk / Mb;
  const int c = lane & 31, n = bn * 32 + c;
  const int64_t plane = (int64_t)nblk * 1024;
  vec4u w[NMOD];
#pragma unroll
  for (int i = 0; i < NMOD; ++i)
    w[i] = __builtin_nontemporal_load(reinterpret_cast<const vec4u *>(
        R + i * plane + (int64_t)blk * 1024 + lane * 16));
  OutT *tw = tile[wave];
  if (n < N) {
    const int ce = eb[n];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r0 = 8 * q + 4 * (lane >> 5), m = bm * 32 + r0;
      if (m >= M) break;
      uint32_t wq[NMOD];
#pragma unroll
      for (int i = 0; i < NMOD; ++i)
        wq[i] = reinterpret_cast<const uint32_t *>(&w[i])[q];
      crt_combine_rows<NMOD, OutT>(wq, m, n, M, ea, ce, shift, alpha, beta, C,
                                   ldc, tw + c * LS + r0);
    }
  }
  // Within the wave only: the LDS writes above and the reads below.
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  crt_tile_out<OutT>(tw, bm, bn, M, N, C, ldc, vec_c, lane, 64);
}
template <int NMOD, typename OutT>
__global__ __launch_bounds__(256) void crt_combine_groups(
    int M, int N, int Mb, int nblk, const uint8_t *__restrict__ R,
    const int *__restrict__ ea, const int *__restrict__ eb, int shift,
    double alpha, double beta, OutT *C, int64_t ldc, bool vec_c) {
  constexpr int LS = 32 + 16 / (int)sizeof(OutT);
  __shared__ __attribute__((aligned(16))) OutT tile[32 * LS];
  const int blk = blockIdx.x;
  const int bm = blk % Mb, bn = blk / Mb;
  const int t = threadIdx.x, q = t & 3, lr = t >> 2;
  const int c = lr & 31, r0 = 8 * q + 4 * (lr >> 5);
  const int n = bn * 32 + c, m = bm * 32 + r0;
  const int64_t plane = (int64_t)nblk * 1024;
  uint32_t w[NMOD];
#pragma unroll
  for (int i = 0; i < NMOD; ++i)
    w[i] = __builtin_nontemporal_load(reinterpret_cast<const uint32_t *>(
        R + i * plane + (int64_t)blk * 1024 + t * 4));
  if (n < N && m < M)
    crt_combine_rows<NMOD, OutT>(w, m, n, M, ea, eb[n], shift, alpha, beta, C,
                                 ldc, tile + c * LS + r0);
  __syncthreads();
  crt_tile_out<OutT>(tile, bm, bn, M, N, C, ldc, vec_c, t, 256);
}

template <int NMOD, typename T>
void launch_crt_residues(hipStream_t s, int R, int K, const T *X, int64_t ld,
                         bool kcont, const int *exps, int bits, int8_t *P,
                         int64_t Kp, int64_t pstride) {
  // Kp is a multiple of 128 (the INT8 GEMM's K step).
  if (kcont)
    crt_residue_kernel<T, NMOD, true>
        <<<dim3((unsigned)(Kp / 128), (unsigned)cdiv(R, 32)), 256, 0, s>>>(
            R, K, X, ld, vec_aligned(X, ld), exps, bits, P, Kp, pstride);
  else
    crt_residue_kernel<T, NMOD, false>
        <<<dim3((unsigned)(Kp / 32), (unsigned)cdiv(R, 128)), 256, 0, s>>>(
            R, K, X, ld, vec_aligned(X, ld), exps, bits, P, Kp, pstride);
}
template <int NMOD, typename OutT>
void launch_crt_combine(hipStream_t s, int M, int N, const uint8_t *R,
                        const int *ea, const int *eb, int shift, double alpha,
                        double beta, OutT *C, int64_t ldc) {
  const int Mb = (int)cdiv(M, 32), nblk = Mb * (int)cdiv(N, 32);
  if constexpr (NMOD <= 12)
    crt_combine_waves<NMOD, OutT><<<(unsigned)cdiv(nblk, 4), 256, 0, s>>>(
        M, N, Mb, nblk, R, ea, eb, shift, alpha, beta, C, ldc,
        vec_aligned(C, ldc));
  else
    crt_combine_groups<NMOD, OutT><<<(unsigned)nblk, 256, 0, s>>>(
        M, N, Mb, nblk, R, ea, eb, shift, alpha, beta, C, ldc,
        vec_aligned(C, ldc));
}

// Runtime modulus count -> template instance (6..20).
template <template <int> class F, typename... Args>
bool with_nmod(int n, Args &&...args) {
  switch (n) {
#define HCB_NMOD(k) \
  case k:           \
    F<k>::run(args...); \
    return true;
    HCB_NMOD(6) HCB_NMOD(7) HCB_NMOD(8) HCB_NMOD(9) HCB_NMOD(10) HCB_NMOD(11) HCB_NMOD(12)
    HCB_NMOD(13) HCB_NMOD(14) HCB_NMOD(15) HCB_NMOD(16) HCB_NMOD(17)
    HCB_NMOD(18) HCB_NMOD(19) HCB_NMOD(20)
#undef HCB_NMOD
  }
  return false;
}
template <int NMOD>
struct ResidueLaunch {
  template <typename... A>
  static void run(A... a) { launch_crt_residues<NMOD>(a...); }
};
template <int NMOD>
struct CombineLaunch {
  template <typename... A>
  static void run(A... a) { launch_crt_combine<NMOD>(a...); }
};

}  // namespace

// ---------------------------------------------------------------------------
hcblasStatus_t gemm_split_bf16(hcblasHandle_t h, const Canon &c, float alpha,
                               float beta, float *C, int64_t ldc, int64_t cOff,
                               int64_t strideC, void *const *Cptrs,
                               int64_t batch) {
  if (c.M == 0 || c.N == 0 || batch == 0) return HCBLAS_STATUS_SUCCESS;
  if (alpha == 0.f || c.K == 0)
    return gemm_real<Elem::F32, float>(h, c, alpha, beta, C, ldc, cOff,
                                       strideC, Cptrs, batch);
  // One K block per step (two blocks per step need 256 VGPRs and spill at
  // 8 waves); K is padded to a whole number of blocks.
  constexpr int SKB = 1;
  const int nkb = (int)(cdiv(c.K, kS3K * SKB) * SKB);
  const size_t a_bytes = sizeof(uint16_t) * kS3Row * (size_t)nkb * c.M;
  const size_t b_bytes = sizeof(uint16_t) * kS3Row * (size_t)nkb * c.N;
  if (ensure_workspace(h->split_ws, a_bytes + b_bytes) != hipSuccess)
    return HCBLAS_STATUS_ALLOC_FAILED;
  uint16_t *Ap = static_cast<uint16_t *>(h->split_ws.ptr);
  uint16_t *Bp = Ap + a_bytes / sizeof(uint16_t);
  BatchView bv;
  hcblasStatus_t st =
      batch_pointers(h, c, C, strideC, Cptrs, batch, sizeof(float), bv);
  if (st != HCBLAS_STATUS_SUCCESS) return st;
  for (int64_t z = 0; z < batch; ++z) {
    const float *A = static_cast<const float *>(bv.A[z]) + c.aOff;
    const float *B = static_cast<const float *>(bv.B[z]) + c.bOff;
    float *Cz = static_cast<float *>(bv.C[z]) + cOff;
    launch_split3_pack(h->stream, c.M, c.K, A, c.lda, c.a_k, Ap, nkb);
    launch_split3_pack(h->stream, c.N, c.K, B, c.ldb, c.b_k, Bp, nkb);
    st = launch_status();
    if (st != HCBLAS_STATUS_SUCCESS) return st;
    Split3Params p{};
    p.M = c.M;
    p.N = c.N;
    p.nkb = nkb;
    p.Ap = Ap;
    p.Bp = Bp;
    p.C = Cz;
    p.ldc = ldc;
    p.alpha = alpha;
    p.beta = beta;
    // 256x128 tiles, 8 waves of 64x64 (two accumulators each); problems
    // that cannot fill the CUs take 128x128 with 4 waves.
    if (cdiv(c.M, 256) * cdiv(c.N, 128) >= h->num_cus)
      st = launch_split3<256, 128, 4, 2, SKB>(h->stream, p);
    else
      st = launch_split3<128, 128, 2, 2, SKB>(h->stream, p);
    if (st != HCBLAS_STATUS_SUCCESS) return st;
  }
  return HCBLAS_STATUS_SUCCESS;
}

// ---------------------------------------------------------------------------
hcblasStatus_t gemm_ozaki_int8(hcblasHandle_t h, const Canon &c, double alpha,
                               double beta, double *C, int64_t ldc,
                               int64_t cOff, int64_t strideC,
                               void *const *Cptrs, int64_t batch) {
  if (c.M == 0 || c.N == 0 || batch == 0) return HCBLAS_STATUS_SUCCESS;
  if (alpha == 0.0 || c.K == 0)
    return gemm_real<Elem::F64, double>(h, c, alpha, beta, C, ldc, cOff,
                                        strideC, Cptrs, batch);
  const int s = h->ozaki_slices;
  constexpr int BK = 128;
  // INT32 exactness: s * Kc * 64 * 64 < 2^31.
  const int64_t kmax = ((int64_t)1 << 31) / (4096 * (int64_t)s) / BK * BK;
  const int64_t kchunk = std::min<int64_t>(cdiv(c.K, BK) * BK, kmax);
  const int64_t Kp = kchunk;
  const size_t a_bytes = (size_t)s * c.M * Kp;
  const size_t b_bytes = (size_t)s * c.N * Kp;
  const size_t e_bytes = sizeof(int) * ((size_t)c.M + c.N);
  if (ensure_workspace(h->split_ws, a_bytes + b_bytes + e_bytes + 256) !=
      hipSuccess)
    return HCBLAS_STATUS_ALLOC_FAILED;
  int8_t *Ap = static_cast<int8_t *>(h->split_ws.ptr);
  int8_t *Bp = Ap + a_bytes;
  int *exps = reinterpret_cast<int *>(
      (reinterpret_cast<uintptr_t>(Bp + b_bytes) + 255) & ~uintptr_t(255));
  int *ea = exps, *eb = exps + c.M;
  BatchView bv;
  hcblasStatus_t st =
      batch_pointers(h, c, C, strideC, Cptrs, batch, sizeof(double), bv);
  if (st != HCBLAS_STATUS_SUCCESS) return st;
  for (int64_t z = 0; z < batch; ++z) {
    double *Cz = static_cast<double *>(bv.C[z]) + cOff;
    for (int64_t k0 = 0; k0 < c.K; k0 += kchunk) {
      const int kc = (int)std::min<int64_t>(kchunk, c.K - k0);
      // Operand views of this K chunk.
      const double *A = static_cast<const double *>(bv.A[z]) + c.aOff +
                        (c.a_k ? k0 : k0 * c.lda);
      const double *B = static_cast<const double *>(bv.B[z]) + c.bOff +
                        (c.b_k ? k0 : k0 * c.ldb);
      fill_int_kernel<<<(unsigned)cdiv(c.M + c.N, 256), 256, 0, h->stream>>>(
          exps, c.M + c.N, kNoScale);
      launch_row_exp(h->stream, c.M, kc, A, c.lda, c.a_k, ea);
      launch_row_exp(h->stream, c.N, kc, B, c.ldb, c.b_k, eb);
      dim3 ga((unsigned)cdiv(Kp, 32), (unsigned)cdiv(c.M, 32));
      ozaki_slice_kernel<<<ga, 256, 0, h->stream>>>(c.M, kc, A, c.lda, c.a_k,
                                                    ea, s, Ap, Kp, false);
      dim3 gb((unsigned)cdiv(Kp, 32), (unsigned)cdiv(c.N, 32));
      ozaki_slice_kernel<<<gb, 256, 0, h->stream>>>(c.N, kc, B, c.ldb, c.b_k,
                                                    eb, s, Bp, Kp, true);
      st = launch_status();
      if (st != HCBLAS_STATUS_SUCCESS) return st;
      PlaneParams p{};
      p.M = c.M;
      p.N = c.N;
      p.Kp = Kp;
      p.S = s;
      p.Ap = Ap;
      p.Bp = Bp;
      // group d = t + u (smallest weight first): A[0..d-1) x B[s-d+1..s)
      p.ngroups = s;
      for (int g = 0; g < s; ++g) {
        const int d = s + 1 - g;
        p.a0[g] = 0;
        p.b0[g] = s - d + 1;
        p.np[g] = d - 1;
        p.weight[g] = __builtin_ldexp(1.0, -7 * (d - 2));
      }
      p.row_exp = ea;
      p.col_exp = eb;
      p.exp_bias = -12;
      p.C = Cz;
      p.ldc = ldc;
      p.alpha = alpha;
      p.beta = k0 == 0 ? beta : 1.0;
      st = launch_planes<Elem::I8, double, 128, 128, BK, 2, 2, 32>(h->stream, p);
      if (st != HCBLAS_STATUS_SUCCESS) return st;
    }
  }
  return HCBLAS_STATUS_SUCCESS;
}

// ---------------------------------------------------------------------------
// Bits per operand for n moduli and a K chunk of kc: the largest b with
// kc * 2^(2b) < P / 2^1.01 (P = product of the moduli), at most 62.
static int crt_bits(int nmod, int64_t kc) {
  double log2p = 0.0;
  for (int t = 0; t < nmod; ++t) log2p += std::log2((double)kModuli[t]);
  // 0.01 bit of margin keeps |C| / P <= 0.4965, which the reconstruction's
  // rounding of S / P relies on.
  const double room = log2p - 1.0 - std::log2((double)kc) - 0.01;
  int b = (int)std::floor(room / 2.0);
  return b > 62 ? 62 : b;
}

// Default modulus count: the fewest moduli whose operand precision matches
// the FP64 kernel's error at this K. Measured on MI355X at K = 4096: for
// uniform[-1,1] inputs the FP64 kernel's mean relative error is about
// 1.2 sqrt(K) 2^-53 and the CRT product's about 6 * 2^-b; for the
// wide-exponent inputs sign * 2^u, u ~ U[-20, 20], the per-row / per-column
// scaling costs about three more bits (52 bits: 1.02e-14 against 9.2e-15).
// b >= 59 - log2(K)/2 keeps both at or below the FP64 kernel's error
// (16 moduli, 55-56 bits, at K = 4096..8192).
static int crt_auto_moduli(int64_t kc) {
  const int need = (int)std::ceil(59.0 - 0.5 * std::log2((double)std::max<int64_t>(kc, 1)));
  for (int n = 8; n <= kMaxModuli; ++n)
    if (crt_bits(n, kc) >= need) return n;
  return kMaxModuli;
}

hcblasStatus_t gemm_ozaki_crt(hcblasHandle_t h, const Canon &c, double alpha,
                              double beta, double *C, int64_t ldc,
                              int64_t cOff, int64_t strideC,
                              void *const *Cptrs, int64_t batch) {
  if (c.M == 0 || c.N == 0 || batch == 0) return HCBLAS_STATUS_SUCCESS;
  if (alpha == 0.0 || c.K == 0)
    return gemm_real<Elem::F64, double>(h, c, alpha, beta, C, ldc, cOff,
                                        strideC, Cptrs, batch);
  constexpr int BK = 2 * kCrtBK;  // the main loop takes K steps in pairs
  // INT32 exactness of one modulus product: kc * 128 * 128 < 2^31.
  constexpr int64_t kmax = 65536;
  const int64_t kchunk = std::min<int64_t>(cdiv(c.K, BK) * BK, kmax);
  const int64_t Kp = kchunk;
  // Workspace for the largest modulus count any chunk uses.
  const int nmax = h->ozaki_moduli
                       ? h->ozaki_moduli
                       : std::max(crt_auto_moduli((int)kchunk),
                                  crt_auto_moduli((int)(c.K % kchunk ? c.K % kchunk : kchunk)));
  // Planes in fragment order over M and N rounded up to the 256-wide tile.
  const int64_t Mpad = cdiv(c.M, kCrtBM) * kCrtBM;
  const int64_t Npad = cdiv(c.N, kCrtBM) * kCrtBM;
  const size_t a_bytes = (size_t)nmax * Mpad * Kp;
  const size_t b_bytes = (size_t)nmax * Npad * Kp;
  const int Mb = (int)cdiv(c.M, 32), Nb = (int)cdiv(c.N, 32);
  const size_t r_bytes = (size_t)nmax * Mb * Nb * 1024;
  const size_t e_bytes = sizeof(int) * ((size_t)c.M + c.N);
  auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
  if (ensure_workspace(h->split_ws, up(a_bytes) + up(b_bytes) + up(r_bytes) +
                                        e_bytes) != hipSuccess)
    return HCBLAS_STATUS_ALLOC_FAILED;
  char *ws = static_cast<char *>(h->split_ws.ptr);
  int8_t *Ap = reinterpret_cast<int8_t *>(ws);
  int8_t *Bp = reinterpret_cast<int8_t *>(ws + up(a_bytes));
  uint8_t *R = reinterpret_cast<uint8_t *>(ws + up(a_bytes) + up(b_bytes));
  int *exps = reinterpret_cast<int *>(ws + up(a_bytes) + up(b_bytes) +
                                      up(r_bytes));
  int *ea = exps, *eb = exps + c.M;
  BatchView bv;
  hcblasStatus_t st =
      batch_pointers(h, c, C, strideC, Cptrs, batch, sizeof(double), bv);
  if (st != HCBLAS_STATUS_SUCCESS) return st;
  for (int64_t z = 0; z < batch; ++z) {
    double *Cz = static_cast<double *>(bv.C[z]) + cOff;
    for (int64_t k0 = 0; k0 < c.K; k0 += kchunk) {
      const int kc = (int)std::min<int64_t>(kchunk, c.K - k0);
      const int nmod = h->ozaki_moduli ? h->ozaki_moduli : crt_auto_moduli(kc);
      const int bits = crt_bits(nmod, kc);
      if (bits < 1) return HCBLAS_STATUS_INVALID_VALUE;
      const double *A = static_cast<const double *>(bv.A[z]) + c.aOff +
                        (c.a_k ? k0 : k0 * c.lda);
      const double *B = static_cast<const double *>(bv.B[z]) + c.bOff +
                        (c.b_k ? k0 : k0 * c.ldb);
      fill_int_kernel<<<(unsigned)cdiv(c.M + c.N, 256), 256, 0, h->stream>>>(
          exps, c.M + c.N, kNoScale);
      launch_row_exp(h->stream, c.M, kc, A, c.lda, c.a_k, ea);
      launch_row_exp(h->stream, c.N, kc, B, c.ldb, c.b_k, eb);
      if (!with_nmod<ResidueLaunch>(nmod, h->stream, c.M, kc, A, c.lda,
                                    c.a_k, (const int *)ea, bits, Ap, Kp,
                                    Mpad * Kp) ||
          !with_nmod<ResidueLaunch>(nmod, h->stream, c.N, kc, B, c.ldb,
                                    c.b_k, (const int *)eb, bits, Bp, Kp,
                                    Npad * Kp))
        return HCBLAS_STATUS_INVALID_VALUE;
      st = launch_status();
      if (st != HCBLAS_STATUS_SUCCESS) return st;
      CrtParams p{};
      p.M = c.M;
      p.N = c.N;
      p.Kp = Kp;
      p.nmod = nmod;
      p.Ap = Ap;
      p.Bp = Bp;
      p.astride = Mpad * Kp;
      p.bstride = Npad * Kp;
      p.R = R;
      p.Mb = Mb;
      p.Nb = Nb;
      p.tiles_m = (int)cdiv(c.M, kCrtBM);
      p.tiles_n = (int)cdiv(c.N, kCrtBM);
      crt_gemm4_kernel<<<dim3((unsigned)(p.tiles_m * p.tiles_n), 1,
                              (unsigned)nmod),
                         256, 0, h->stream>>>(p);
      st = launch_status();
      if (st != HCBLAS_STATUS_SUCCESS) return st;
      with_nmod<CombineLaunch>(nmod, h->stream, c.M, c.N, (const uint8_t *)R,
                               (const int *)ea, (const int *)eb, 2 * bits,
                               alpha, k0 == 0 ? beta : 1.0, Cz, ldc);
      st = launch_status();
      if (st != HCBLAS_STATUS_SUCCESS) return st;
    }
  }
  return HCBLAS_STATUS_SUCCESS;
}

// ---------------------------------------------------------------------------
// Ozaki scheme II for FP32 (S / C GEMM). Operands are scaled per row (A) and
// per column (B) by their 2-norms instead of their largest element: with
// ||A'_m||_2, ||B'_n||_2 <= 2^beta, Cauchy-Schwarz bounds every |(A'B')(m,n)|
// by 2^(2 beta) < P / 2 for any K, and the operand precision of a row is its
// norm's, which for rows with a wide exponent spread is several bits more
// than the max-element scaling gives. beta = floor((log2 P - 1.01) / 2), at
// most 30 so that the scaled operands fit one 32-bit limb. The quantisation
// error of an output is about 0.58 * 2^-beta * ||A_m|| ||B_n||. Default 8
// moduli (beta = 30): measured at 8192^3 the max and mean errors are 1-3
// orders of magnitude below the FP32 kernel's for uniform and wide-exponent
// inputs. 7 moduli (beta = 27) stay below it for uniform inputs but not for
// wide-exponent ones (mean 1.6e-6 against 1.2e-6 at K = 4096), whose norms
// are dominated by a few large elements.
static int crt_beta_f32(int nmod) {
  double log2p = 0.0;
  for (int t = 0; t < nmod; ++t) log2p += std::log2((double)kModuli[t]);
  const int b = (int)std::floor((log2p - 1.01) / 2.0);
  return b > 30 ? 30 : b;
}

hcblasStatus_t gemm_ozaki_crt_f32(hcblasHandle_t h, const Canon &c,
                                  float alpha, float beta, float *C,
                                  int64_t ldc, int64_t cOff, int64_t strideC,
                                  void *const *Cptrs, int64_t batch) {
  if (c.M == 0 || c.N == 0 || batch == 0) return HCBLAS_STATUS_SUCCESS;
  if (alpha == 0.f || c.K == 0)
    return gemm_real<Elem::F32, float>(h, c, alpha, beta, C, ldc, cOff,
                                       strideC, Cptrs, batch);
  constexpr int BK = 2 * kCrtBK;  // the main loop takes K steps in pairs
  constexpr int64_t kmax = 65536;  // INT32 exactness: kc * 128 * 128 < 2^31
  const int64_t kchunk = std::min<int64_t>(cdiv(c.K, BK) * BK, kmax);
  const int64_t Kp = kchunk;
  const int nmod = h->ozaki_moduli ? h->ozaki_moduli : 8;
  const int bb = crt_beta_f32(nmod);
  const int pa = (int)cdiv(kchunk, 64), pb = pa;  // 64-wide k blocks
  // Planes in fragment order over M and N rounded up to the 256-wide tile.
  const int64_t Mpad = cdiv(c.M, kCrtBM) * kCrtBM;
  const int64_t Npad = cdiv(c.N, kCrtBM) * kCrtBM;
  const size_t a_bytes = (size_t)nmod * Mpad * Kp;
  const size_t b_bytes = (size_t)nmod * Npad * Kp;
  const int Mb = (int)cdiv(c.M, 32), Nb = (int)cdiv(c.N, 32);
  const size_t r_bytes = (size_t)nmod * Mb * Nb * 1024;
  const size_t e_bytes = sizeof(int) * ((size_t)c.M + c.N);
  const size_t p_bytes = sizeof(double) * ((size_t)pa * c.M + (size_t)pb * c.N);
  auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
  if (ensure_workspace(h->split_ws, up(a_bytes) + up(b_bytes) + up(r_bytes) +
                                        up(e_bytes) + p_bytes) != hipSuccess)
    return HCBLAS_STATUS_ALLOC_FAILED;
  char *ws = static_cast<char *>(h->split_ws.ptr);
  int8_t *Ap = reinterpret_cast<int8_t *>(ws);
  int8_t *Bp = reinterpret_cast<int8_t *>(ws + up(a_bytes));
  uint8_t *R = reinterpret_cast<uint8_t *>(ws + up(a_bytes) + up(b_bytes));
  int *exps = reinterpret_cast<int *>(ws + up(a_bytes) + up(b_bytes) +
                                      up(r_bytes));
  double *part = reinterpret_cast<double *>(ws + up(a_bytes) + up(b_bytes) +
                                            up(r_bytes) + up(e_bytes));
  int *ea = exps, *eb = exps + c.M;
  double *parta = part, *partb = part + (size_t)pa * c.M;
  BatchView bv;
  hcblasStatus_t st =
      batch_pointers(h, c, C, strideC, Cptrs, batch, sizeof(float), bv);
  if (st != HCBLAS_STATUS_SUCCESS) return st;
  auto norms = [&](int R, int kc, const float *X, int64_t ld, bool kcont,
                   double *pp, int *ex) {
    const int nb = (int)cdiv(kc, 64);
    row_sumsq_kernel<<<dim3((unsigned)nb, (unsigned)cdiv(R, 64)), 256, 0,
                       h->stream>>>(R, kc, X, ld, kcont, vec_aligned(X, ld),
                                    pp);
    row_norm_shift_kernel<<<(unsigned)cdiv(R, 4), 256, 0, h->stream>>>(
        R, nb, pp, bb, kc, ex);
  };
  for (int64_t z = 0; z < batch; ++z) {
    float *Cz = static_cast<float *>(bv.C[z]) + cOff;
    for (int64_t k0 = 0; k0 < c.K; k0 += kchunk) {
      const int kc = (int)std::min<int64_t>(kchunk, c.K - k0);
      const float *A = static_cast<const float *>(bv.A[z]) + c.aOff +
                       (c.a_k ? k0 : k0 * c.lda);
      const float *B = static_cast<const float *>(bv.B[z]) + c.bOff +
                       (c.b_k ? k0 : k0 * c.ldb);
      norms(c.M, kc, A, c.lda, c.a_k, parta, ea);
      norms(c.N, kc, B, c.ldb, c.b_k, partb, eb);
      if (!with_nmod<ResidueLaunch>(nmod, h->stream, c.M, kc, A, c.lda,
                                    c.a_k, (const int *)ea, 0, Ap, Kp,
                                    Mpad * Kp) ||
          !with_nmod<ResidueLaunch>(nmod, h->stream, c.N, kc, B, c.ldb,
                                    c.b_k, (const int *)eb, 0, Bp, Kp,
                                    Npad * Kp))
        return HCBLAS_STATUS_INVALID_VALUE;
      st = launch_status();
      if (st != HCBLAS_STATUS_SUCCESS) return st;
      CrtParams p{};
      p.M = c.M;
      p.N = c.N;
      p.Kp = Kp;
      p.nmod = nmod;
      p.Ap = Ap;
      p.Bp = Bp;
      p.astride = Mpad * Kp;
      p.bstride = Npad * Kp;
      p.R = R;
      p.Mb = Mb;
      p.Nb = Nb;
      p.tiles_m = (int)cdiv(c.M, kCrtBM);
      p.tiles_n = (int)cdiv(c.N, kCrtBM);
      crt_gemm4_kernel<<<dim3((unsigned)(p.tiles_m * p.tiles_n), 1,
                              (unsigned)nmod),
                         256, 0, h->stream>>>(p);
      st = launch_status();
      if (st != HCBLAS_STATUS_SUCCESS) return st;
      with_nmod<CombineLaunch>(nmod, h->stream, c.M, c.N, (const uint8_t *)R,
                               (const int *)ea, (const int *)eb, 0,
                               (double)alpha, k0 == 0 ? (double)beta : 1.0,
                               Cz, ldc);
      st = launch_status();
      if (st != HCBLAS_STATUS_SUCCESS) return st;
    }
  }
  return HCBLAS_STATUS_SUCCESS;
}

}  // namespace hcb
