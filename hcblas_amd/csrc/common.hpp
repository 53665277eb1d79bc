// Internal definitions shared by every translation unit of libhcblas.
//
// The handle plays the role of the reference's Hcblaslibrary context
// (lib/include/hcblaslib.h:67-101): it owns the device, the stream, and the
// persistent device workspaces that the reference allocated and freed on every
// call (sdot_array.cpp:40-41, sgemv_array.cpp:41; here allocated once).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>

#include "../../include/hcblas.h"

namespace hcb {

constexpr uint32_t kHandleMagic = 0x48434d49u;  // "HCMI"

// Largest grid of the level-1 elementwise kernels (they grid-stride beyond).
constexpr int kMaxReduceBlocks = 1 << 21;
// Reduction scratch (dot / asum): one FP64 partial per workgroup of the
// bounded reduction grid.
constexpr int kReducePartials = 256;

struct Workspace {
  void *ptr = nullptr;
  size_t bytes = 0;
};

// Fixed per-GPU scratch of the allocation-free GEMV entry and the largest
// device index it serves.
constexpr size_t kNoAllocScratchBytes = HCBLAS_GEMV_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

// Device address of the no-allocation scratch on GPU `dev` (looked up once
// per device); nullptr if the lookup failed.
void *scratch_address(int dev);

}  // namespace hcb

struct hcblasContext {
  uint32_t magic = 0;
  int device = 0;
  int num_cus = 256;
  hipStream_t stream = nullptr;
  int math_mode = HCBLAS_DEFAULT_MATH;
  int ozaki_slices = 8;
  int ozaki_moduli = 0;  // 0: chosen per K (see gemm_split.hip)

  // dot / asum: per-workgroup partials.
  double *d_partials = nullptr;
  // Pinned, host-coherent block of the host-result reductions (dot / asum)
  // and its device address: h_result[0] holds the result and, as a 64-bit
  // word, h_result[1] the sequence number of the call that wrote it. The
  // reduction's last workgroup writes the result, then (system-scope
  // release) the sequence number; the host returns as soon as it reads its
  // call's number there, without waiting for the kernel's completion signal.
  double *h_result = nullptr;
  double *h_result_dev = nullptr;
  uint64_t result_seq = 0;

  // gemv split-reduction partials and per-row-block tickets (grow-only; the
  // ticket region is zeroed when it is (re)allocated and every launch leaves
  // it zero again).
  hcb::Workspace gemv_partials;
  hcb::Workspace gemv_tickets;

  // Real embedding of complex GEMM operands.
  hcb::Workspace gemm_ws;
  // BF16 planes / INT8 slices and exponents of the split GEMM paths (kept
  // apart from gemm_ws: a split complex GEMM reads the embedding while it
  // writes the planes).
  hcb::Workspace split_ws;
  // FP32 copies of BF16 operands for the adjoint GEMMs.
  hcb::Workspace adj_ws;
  // Split-K partial tiles of the plain GEMM.
  hcb::Workspace kpart_ws;
  // This device's address of the fixed scratch of hcblas_gemv_noalloc.
  void *noalloc_scratch = nullptr;
};

namespace hcb {

inline bool handle_ok(hcblasHandle_t h) {
  return h != nullptr && h->magic == kHandleMagic;
}

// Makes `ws` at least `bytes` large. Growing frees the old block, which
// synchronises the device first (hipFree semantics), so work still in flight
// on the old block is complete before it is released.
hipError_t ensure_workspace(Workspace &ws, size_t bytes);

// Makes the gemv ticket region at least `count` counters and zero.
hipError_t ensure_tickets(hcblasHandle_t h, size_t count);

inline hcblasStatus_t launch_status() {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? HCBLAS_STATUS_SUCCESS
                         : HCBLAS_STATUS_EXECUTION_FAILED;
}

// Operation code of an argument: HCBLAS_OP_N/T/C, with the reference's
// internal transpose codes NoTrans = 'n' and Trans = 't' (hcblasTranspose,
// hcblaslib.h:51) as aliases of OP_N and OP_T. -1 for any other value.
inline int canon_op(int t) {
  switch (t) {
    case HCBLAS_OP_N:
    case HCBLAS_NO_TRANS:
      return HCBLAS_OP_N;
    case HCBLAS_OP_T:
    case HCBLAS_TRANS:
      return HCBLAS_OP_T;
    case HCBLAS_OP_C:
      return HCBLAS_OP_C;
    default:
      return -1;
  }
}

inline bool aligned16(const void *p) {
  return (reinterpret_cast<uintptr_t>(p) & 15u) == 0;
}

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

__host__ __device__ inline int64_t cdiv_dev(int64_t a, int64_t b) {
  return (a + b - 1) / b;
}

}  // namespace hcb
