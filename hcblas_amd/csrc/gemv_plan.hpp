// Launch plan of the GEMV kernels (host only, no HIP dependency): which
// kernel runs, its grid, the reduction split count and the bytes of split
// partials it needs. Shared by level2.hip and the CPU test that checks the
// partials of the allocation-free entry against its fixed scratch.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>

namespace hcb {

// Kernel geometry constants of level2.hip.
constexpr int kGemvThreads = 256;
constexpr int kBandNRows = 4;     // XR of the band-N kernel
constexpr int kBandNCols = 8;     // UC: columns in flight
constexpr int kCPW = 4;           // columns per wave of the column kernel
constexpr int kColsPerBlock = 4 * kCPW;
constexpr int kBandCols = 8;      // columns per workgroup of the band kernel

enum class GemvKernel { kBandN, kPanel, kBand16, kBand8, kColumn };

struct GemvPlan {
  GemvKernel kernel;
  int64_t oblocks;  // output blocks (grid x, or grid y for band-N)
  int splits;       // reduction splits (grid y, or grid x for band-N)
  size_t partial_bytes;  // split partials (0 when splits == 1)
};

inline int64_t plan_cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// panel: the reduction runs along the contiguous storage rows' columns
// (y has R entries); otherwise along the rows (y has Cc entries). vec: the
// 16-byte path is legal; vn: elements per 16 bytes; esize: element bytes.
inline GemvPlan plan_gemv(int num_cus, bool panel, int64_t R, int64_t Cc,
                          bool vec, int vn, size_t esize, int64_t batch) {
  const int64_t len_y = panel ? R : Cc;
  const int64_t target = (int64_t)num_cus * 4;  // workgroups to aim for
  GemvPlan p{GemvKernel::kColumn, 1, 1, 0};
  // Band kernel for the column case when a band of 16 (or 8) row steps
  // fits and the column blocks alone give the grid enough workgroups.
  int band_xr = 0;
  if (!panel && vec) {
    for (int xr : {16, 8}) {
      const int64_t bands = plan_cdiv(R, (int64_t)xr * kGemvThreads * vn);
      if (R >= (int64_t)xr * kGemvThreads * vn && bands <= 65535 &&
          plan_cdiv(Cc, kBandCols) * bands * batch >= target) {
        band_xr = xr;
        break;
      }
    }
  }
  // Band-N kernel for the row-contiguous case: whole 16 KiB bands and at
  // least 64 columns.
  const int64_t bandn_rows = (int64_t)kBandNRows * kGemvThreads * vn;
  const bool bandn = panel && vec && R >= bandn_rows && Cc >= 64 &&
                     plan_cdiv(R, bandn_rows) <= 65535;
  if (bandn) {
    p.kernel = GemvKernel::kBandN;
    p.oblocks = plan_cdiv(R, bandn_rows);
    const int64_t per_batch = std::max<int64_t>(
        1, plan_cdiv((int64_t)num_cus, p.oblocks * batch));
    p.splits = (int)std::max<int64_t>(
        1, std::min<int64_t>(plan_cdiv(Cc, 256), per_batch));
  } else if (panel) {
    p.kernel = GemvKernel::kPanel;
    p.oblocks = plan_cdiv(R, 64 * (int64_t)(vec ? vn : 1));
    // at least 32 columns per wave-interleaved split
    const int64_t max_split = std::max<int64_t>(1, Cc / 128);
    p.splits = (int)std::min<int64_t>(
        std::max<int64_t>(1, plan_cdiv(target, p.oblocks * batch)), max_split);
  } else if (band_xr) {
    p.kernel = band_xr == 16 ? GemvKernel::kBand16 : GemvKernel::kBand8;
    p.oblocks = plan_cdiv(Cc, kBandCols);
    p.splits = (int)plan_cdiv(R, (int64_t)band_xr * kGemvThreads * vn);
  } else {
    p.kernel = GemvKernel::kColumn;
    p.oblocks = plan_cdiv(Cc, kColsPerBlock);
    const int64_t rsteps = plan_cdiv(R, 64 * (vec ? vn : 1));
    const int64_t max_split = std::max<int64_t>(1, rsteps / 16);
    p.splits = (int)std::min<int64_t>(
        std::max<int64_t>(1, plan_cdiv(target, p.oblocks * batch)), max_split);
  }
  if (!band_xr && !bandn && p.splits > 256) p.splits = 256;
  if (p.splits > 1)
    p.partial_bytes = esize * (size_t)p.splits * (size_t)len_y *
                      (size_t)std::min<int64_t>(batch, 65535);
  return p;
}

}  // namespace hcb
