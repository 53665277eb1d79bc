// Plane-group MFMA GEMM: the common core of the split-precision paths.
//
// Operands arrive pre-packed, k-contiguous, as S planes per row:
//   Ap[m][q][0..Kp)  (q = 0..S-1),   Bp[n][q][0..Kp)
// A "group" g multiplies the window of np_g consecutive A planes starting at
// plane a0_g with the window of np_g consecutive B planes starting at b0_g:
//     G_g(m,n) = sum_{j<np_g} sum_k Ap[m][a0_g+j][k] * Bp[n][b0_g+j][k]
// which is one long-K GEMM (K = np_g*Kp) because the windows are contiguous.
// Groups collect the products of equal magnitude. Ozaki INT8 DGEMM: A planes = slices 1..s, B planes = slices s..1; group
//     d (t+u = d) is A[0..d-1) x B[s-d+1..s), summed exactly in INT32 and
//     added to an FP64 accumulator with weight 2^(-7(d-2)); the epilogue
//     applies the per-row / per-column power-of-two scales.
#pragma once

#include <climits>

#include "gemm_kernel.hpp"

namespace hcb {

constexpr int kMaxGroups = 16;

struct PlaneParams {
  int M, N;
  int64_t Kp;       // padded K per plane (multiple of BK)
  int S;            // planes per row
  const void *Ap;   // M rows of S*Kp elements
  const void *Bp;   // N rows of S*Kp elements
  int ngroups;
  int a0[kMaxGroups], b0[kMaxGroups], np[kMaxGroups];
  double weight[kMaxGroups];  // INT8 only: scale of group g
  const int *row_exp;         // INT8 only: per-row exponent of A (M)
  const int *col_exp;         // INT8 only: per-column exponent of B (N)
  int exp_bias;               // INT8 only: C scale = 2^(row+col+bias)
  void *C;                    // column-major M x N (float or double)
  int64_t ldc;
  double alpha, beta;
  int tiles_m, tiles_n;
};

template <Elem E, typename OutT, int BM, int BN, int BK, int WM, int WN,
          int MI = 16>
__global__ __launch_bounds__(64 * WM * WN) void gemm_planes_kernel(
    PlaneParams p) {
  using Tr = ElemT<E>;
  using T = typename Tr::T;
  using AccV = typename AccOf<E, MI>::type;
  constexpr bool INT = E == Elem::I8;
  constexpr int VEC = Tr::VEC;
  constexpr int NG = 64 / MI;
  constexpr int CK = NG * VEC;
  constexpr int NVAL = MI * MI / 64;
  static_assert(Tr::STEPS == 1, "packed types only");
  static_assert(BK % CK == 0, "BK");
  constexpr int WTM = BM / WM, WTN = BN / WN;
  constexpr int FM = WTM / MI, FN = WTN / MI;
  constexpr int NTH = 64 * WM * WN;
  using LA = TileLoader<T, BM, BK, true, true, NTH>;
  using LB = TileLoader<T, BN, BK, true, true, NTH>;
  constexpr int ROWB = LA::ROWB;
  constexpr int A_BYTES = BM * ROWB, B_BYTES = BN * ROWB;
  __shared__ __attribute__((aligned(16))) char lds[2 * (A_BYTES + B_BYTES)];

  int tile_m, tile_n;
  tile_coords(blockIdx.x, p.tiles_m, p.tiles_n, tile_m, tile_n);
  const int m0 = tile_m * BM, n0 = tile_n * BN;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wm = wave / WN, wn = wave % WN;
  const int lrow = lane % MI, lgrp = lane / MI;
  const int64_t ld = (int64_t)p.S * p.Kp;
  const T *Ap = static_cast<const T *>(p.Ap);
  const T *Bp = static_cast<const T *>(p.Bp);

  AccV acc[FM][FN];
  double facc[INT ? FM : 1][INT ? FN : 1][INT ? NVAL : 1];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = AccV{};
  if constexpr (INT) {
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < NVAL; ++r) facc[i][j][r] = 0.0;
  }

  const int kt_per_plane = (int)(p.Kp / BK);
  // Flat tile sequence over all groups.
  int total = 0;
  for (int g = 0; g < p.ngroups; ++g) total += p.np[g] * kt_per_plane;

  LA la;
  LB lb;
  auto load_tile = [&](int g, int kt) {
    const int64_t ka = (int64_t)p.a0[g] * p.Kp + (int64_t)kt * BK;
    const int64_t kb = (int64_t)p.b0[g] * p.Kp + (int64_t)kt * BK;
    la.load_full(Ap + ka, ld, m0, 0, p.M);
    lb.load_full(Bp + kb, ld, n0, 0, p.N);
  };
  auto buf = [&](int t) { return lds + (t & 1) * (A_BYTES + B_BYTES); };
  auto store = [&](int t) {
    la.store(buf(t));
    lb.store(buf(t) + A_BYTES);
  };
  // Advances a (group, k-tile) cursor by one tile of the flat sequence.
  auto advance = [&](int &g, int &kt) {
    if (++kt == p.np[g] * kt_per_plane) {
      kt = 0;
      ++g;
    }
  };
  auto compute = [&](const char *cur) {
    const char *As = cur;
    const char *Bs = cur + A_BYTES;
#pragma unroll
    for (int c = 0; c < BK / CK; ++c) {
      vec4u af[FM], bf[FN];
      const int kbyte = (c * CK + lgrp * VEC) * (int)sizeof(T);
#pragma unroll
      for (int i = 0; i < FM; ++i)
        af[i] = *reinterpret_cast<const vec4u *>(
            As + (wm * WTM + i * MI + lrow) * ROWB + kbyte);
#pragma unroll
      for (int j = 0; j < FN; ++j)
        bf[j] = *reinterpret_cast<const vec4u *>(
            Bs + (wn * WTN + j * MI + lrow) * ROWB + kbyte);
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
          mfma_step<E, MI, 0>(af[i], bf[j], acc[i][j]);
    }
  };
  int cg = 0, ckt = 0;  // group / tile being computed
  // End of a group: INT32 group sums are exact; fold them into FP64.
  auto finish_tile = [&]() {
    if (++ckt == p.np[cg] * kt_per_plane) {
      if constexpr (INT) {
        const double w = p.weight[cg];
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) {
#pragma unroll
            for (int r = 0; r < NVAL; ++r)
              facc[i][j][r] = __builtin_fma((double)acc[i][j][r], w,
                                            facc[i][j][r]);
            acc[i][j] = AccV{};
          }
      }
      ckt = 0;
      ++cg;
    }
  };

  // Registers hold tile t+1 while tile t computes; they are written to the
  // other LDS buffer at the top of the step and refilled with tile t+2.
  int lg = 0, lkt = 0;  // cursor of the tile in the registers
  load_tile(0, 0);
  store(0);
  if (total > 1) {
    advance(lg, lkt);
    load_tile(lg, lkt);
  }
  __syncthreads();
  for (int t = 0; t < total; ++t) {
    if (t + 1 < total) store(t + 1);
    if (t + 2 < total) {
      advance(lg, lkt);
      load_tile(lg, lkt);
    }
    __builtin_amdgcn_sched_barrier(0);
    compute(buf(t));
    finish_tile();
    __syncthreads();
  }

  OutT *C = static_cast<OutT *>(p.C);
#pragma unroll
  for (int i = 0; i < FM; ++i) {
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int col = n0 + wn * WTN + j * MI + lrow;
      if (col >= p.N) continue;
      const int rowb = m0 + wm * WTM + i * MI;
      int ce = 0;
      if constexpr (INT) ce = p.col_exp[col];
#pragma unroll
      for (int r = 0; r < NVAL; ++r) {
        const int row = rowb + out_row<E, MI>(lgrp, r);
        if (row >= p.M) continue;
        OutT *dst = C + (int64_t)col * p.ldc + row;
        if constexpr (INT) {
          // An all-zero row / column has no exponent (INT_MIN) and a zero
          // accumulator; a row / column holding an Inf or NaN is marked
          // INT_MAX and its outputs are NaN.
          const int re = p.row_exp[row];
          const double v =
              (re == INT_MAX || ce == INT_MAX) ? __builtin_nan("")
              : (re == INT_MIN || ce == INT_MIN)
                  ? 0.0
                  : __builtin_ldexp(facc[i][j][r], re + ce + p.exp_bias);
          double o = p.alpha * v;
          if (p.beta != 0.0) o = __builtin_fma(p.beta, (double)*dst, o);
          *dst = (OutT)o;
        } else {
          const float a = (float)p.alpha, b = (float)p.beta;
          float o = a * (float)acc[i][j][r];
          if (b != 0.f) o = __builtin_fmaf(b, (float)*dst, o);
          *dst = (OutT)o;
        }
      }
    }
  }
}

}  // namespace hcb
