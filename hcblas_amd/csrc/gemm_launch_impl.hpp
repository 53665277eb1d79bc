// Definition of launch_gemm; included once per element type so the template
// instances compile in parallel translation units.
#pragma once

#include "gemm_launch.hpp"

namespace hcb {

template <Elem E>
struct TileCfg;
template <>
struct TileCfg<Elem::F32> {
  static constexpr int BK_BIG = 32, BK_SMALL = 32;
};
template <>
struct TileCfg<Elem::F64> {
  static constexpr int BK_BIG = 16, BK_SMALL = 16;
};
template <>
struct TileCfg<Elem::F16> {
  static constexpr int BK_BIG = 64, BK_SMALL = 64;
};
template <>
struct TileCfg<Elem::BF16> {
  static constexpr int BK_BIG = 64, BK_SMALL = 64;
};

template <Elem E, typename OutT, int BM, int BN, int BK, int WM, int WN,
          int MI = 16>
hcblasStatus_t launch_tiles(hipStream_t stream, GemmParams p, bool a_k,
                            bool b_k, bool vload, int64_t batch) {
  p.tiles_m = (int)cdiv(p.M, BM);
  p.tiles_n = (int)cdiv(p.N, BN);
  const unsigned tiles = (unsigned)(p.tiles_m * p.tiles_n);
  for (int64_t b0 = 0; b0 < batch; b0 += 65535) {
    GemmParams q = p;
    const int64_t nb = batch - b0 < 65535 ? batch - b0 : 65535;
    if (b0) {
      // Advance strided bases / pointer arrays to this chunk of the batch.
      if (q.Aptrs) q.Aptrs += b0;
      else q.A = (const char *)q.A + b0 * q.strideA * sizeof(typename ElemT<E>::T);
      if (q.Bptrs) q.Bptrs += b0;
      else q.B = (const char *)q.B + b0 * q.strideB * sizeof(typename ElemT<E>::T);
      if (q.Cptrs) q.Cptrs += b0;
      else q.C = (char *)q.C + b0 * q.strideC * sizeof(OutT);
    }
    dim3 grid(tiles, 1, (unsigned)nb);
    // Pointer-array batches with 16-byte loads check each problem's bases
    // on the device (CHK).
    const bool chk = vload && (q.Aptrs || q.Bptrs);
#define HCB_GEMM_CASE(AK, BKC, VL, CK)                                    \
  if (a_k == AK && b_k == BKC && vload == VL && chk == CK)                \
    gemm_mfma_kernel<E, OutT, AK, BKC, BM, BN, BK, WM, WN, VL, VL, MI, CK> \
        <<<grid, 64 * WM * WN, 0, stream>>>(q);
    HCB_GEMM_CASE(false, false, false, false)
    HCB_GEMM_CASE(false, true, false, false)
    HCB_GEMM_CASE(true, false, false, false)
    HCB_GEMM_CASE(true, true, false, false)
    HCB_GEMM_CASE(false, false, true, false)
    HCB_GEMM_CASE(false, true, true, false)
    HCB_GEMM_CASE(true, false, true, false)
    HCB_GEMM_CASE(true, true, true, false)
    HCB_GEMM_CASE(false, false, true, true)
    HCB_GEMM_CASE(false, true, true, true)
    HCB_GEMM_CASE(true, false, true, true)
    HCB_GEMM_CASE(true, true, true, true)
#undef HCB_GEMM_CASE
    hcblasStatus_t st = launch_status();
    if (st != HCBLAS_STATUS_SUCCESS) return st;
  }
  return HCBLAS_STATUS_SUCCESS;
}

// Macro-tile choice (measured at 8192^3 on MI355X, the largest tile whose
// grid still covers every CU is taken):
//   FP32  256x256x32, 8 waves of 128x64 on 32x32 MFMA blocks (146-148
//         TF/s, 93-94% of peak, with the step's LDS writes and loads issued
//         among the first chunk's MFMAs), then 256x128 8 waves, 128x128 4
//         waves, 64x64;
//   FP64  256x128x16, 8 waves of 64x64 (71.6 TF/s NN and TN, 91%, with the
//         same interleave; 66.9 / 64.5 before it; 128x256 2x4 was as fast
//         on NN, 128x128 and
//         smaller tiles stay at 75% or below; aligned operands only, the
//         element-wise loaders spill at this size), then 128x64x16 8 waves
//         (58.5), then 64x64 8 waves;
//   F16 / BF16 256x256x64, 8 waves on 32x32 blocks, then 128x128 8 waves,
//         then 64x64.
template <Elem E, typename OutT>
hcblasStatus_t launch_gemm(hipStream_t stream, int num_cus, GemmParams p,
                           bool a_k, bool b_k, bool vload, int64_t batch) {
  auto tiles = [&](int bm, int bn) {
    return cdiv(p.M, bm) * cdiv(p.N, bn) * batch;
  };
  // Tiles whose element-wise loaders do not fit the registers take aligned
  // strided operands only: a pointer-array batch may hold misaligned bases,
  // which its kernel sends to the element-wise loaders.
  const bool big_ok = vload && !p.Aptrs && !p.Bptrs;
  if constexpr (E == Elem::F32) {
    if (tiles(256, 256) >= num_cus)
      return launch_tiles<E, OutT, 256, 256, 32, 2, 4, 32>(stream, p, a_k, b_k,
                                                           vload, batch);
    if (tiles(256, 128) >= num_cus)
      return launch_tiles<E, OutT, 256, 128, 32, 2, 4>(stream, p, a_k, b_k,
                                                       vload, batch);
    if (tiles(128, 128) >= num_cus)
      return launch_tiles<E, OutT, 128, 128, 32, 2, 2>(stream, p, a_k, b_k,
                                                       vload, batch);
  } else if constexpr (E == Elem::F64) {
    if (big_ok && tiles(256, 128) >= num_cus)
      return launch_tiles<E, OutT, 256, 128, 16, 4, 2>(stream, p, a_k, b_k,
                                                       vload, batch);
    if (tiles(128, 64) >= num_cus)
      return launch_tiles<E, OutT, 128, 64, 16, 2, 4>(stream, p, a_k, b_k,
                                                      vload, batch);
  } else {
    // (the element-wise loaders of unaligned operands do not fit the 256x256
    // tile's registers: they take the 128x128 tile)
    if (big_ok && tiles(256, 256) >= num_cus)
      return launch_tiles<E, OutT, 256, 256, TileCfg<E>::BK_BIG, 2, 4, 32>(
          stream, p, a_k, b_k, vload, batch);
    // 8 waves of 64 x 32 (fragment read-ahead, two-vector transposed loads):
    // 2048^3 HGEMM NN 38.6 against 44.4 us with 4 waves of 64 x 64.
    if (tiles(128, 128) >= num_cus)
      return launch_tiles<E, OutT, 128, 128, TileCfg<E>::BK_BIG, 2, 4, 32>(
          stream, p, a_k, b_k, vload, batch);
  }
  // Small problems: 64x64 tiles, 8 waves of 32x16 for every type (more
  // waves per CU to cover the load latency when there are only about as
  // many tiles as CUs: 1024^3 SGEMM 0.0376 ms against 0.0382 with 4 waves,
  // the conv-net shape set 24.6 TF/s against 23.8; 1024^3 DGEMM 58.8 us
  // against 71.6). FP32 with K >= 512 takes 128-wide K steps:
  // a 32-wide step is 16 MFMAs per wave, too short to cover the global loads
  // issued one step ahead (1024^3 kernel 27.5 us against 31.6 with BK = 32,
  // 29.5 with 64); shorter K keeps 32 so the last step pads little.
  if constexpr (E == Elem::F32) {
    if (p.K >= 512)
      return launch_tiles<E, OutT, 64, 64, 128, 2, 4>(stream, p, a_k, b_k,
                                                      vload, batch);
  }
  return launch_tiles<E, OutT, 64, 64, TileCfg<E>::BK_SMALL, 2, 4>(
      stream, p, a_k, b_k, vload, batch);
}

template <Elem E, typename OutT>
hcblasStatus_t launch_gemm_splitk(hipStream_t stream, int num_cus,
                                  GemmParams p, bool a_k, bool b_k, int S,
                                  int ks, typename ElemT<E>::Acc *W) {
  using Acc = typename ElemT<E>::Acc;
  constexpr int BM = 64, BN = 64, BK = TileCfg<E>::BK_SMALL;
  // 8 waves of 32 x 16 for every type (FP64 256 x 128 x 16384 96 us
  // against 123 with 4 waves of 32 x 32).
  constexpr int WM = 2, WN = 4;
  p.tiles_m = (int)cdiv(p.M, BM);
  p.tiles_n = (int)cdiv(p.N, BN);
  dim3 grid((unsigned)(p.tiles_m * p.tiles_n), 1, (unsigned)S);
#define HCB_SPLITK_CASE(AK, BKC)                                           \
  if (a_k == AK && b_k == BKC)                                             \
    gemm_splitk_kernel<E, AK, BKC, BM, BN, BK, WM, WN, true, true, 16>     \
        <<<grid, 64 * WM * WN, 0, stream>>>(p, ks, W);
  HCB_SPLITK_CASE(false, false)
  HCB_SPLITK_CASE(false, true)
  HCB_SPLITK_CASE(true, false)
  HCB_SPLITK_CASE(true, true)
#undef HCB_SPLITK_CASE
  hcblasStatus_t st = launch_status();
  if (st != HCBLAS_STATUS_SUCCESS) return st;
  const int64_t MN = (int64_t)p.M * p.N;
  const int64_t nb = cdiv(MN, 256);
  const unsigned g = (unsigned)(nb < 4 * num_cus ? nb : 4 * num_cus);
  splitk_reduce_kernel<Acc, OutT><<<g, 256, 0, stream>>>(
      S, p.M, p.N, W, static_cast<OutT *>(p.C) + p.cOff, p.ldc,
      (Acc)p.alpha, (Acc)p.beta);
  return launch_status();
}

}  // namespace hcb
