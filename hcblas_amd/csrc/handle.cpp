// Handle, stream, math mode and host<->device helpers.
// Reference: lib/src/hcblas.cpp:39-312 (hcblasCreate ... hcblasGetMatrix) and
// the hipBLAS stream semantics of lib/src/hcc_detail/hipblas.cpp:51-62.
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "common.hpp"

// Split partials of hcblas_gemv_noalloc: one fixed array on each GPU, used
// in place of the handle's grow-on-demand GEMV workspace so that the call
// allocates nothing. Host code never passes this name to a kernel or a copy:
// it uses the device address that hipGetSymbolAddress returns for the
// current GPU (scratch_address below).
static __device__ __align__(256) char scratch[hcb::kNoAllocScratchBytes];

namespace hcb {

void *scratch_address(int dev) {
  static std::mutex mu;
  static void *addr[kMaxDevices] = {};
  if (dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(mu);
  if (addr[dev] == nullptr) {
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) return nullptr;
    if (cur != dev && hipSetDevice(dev) != hipSuccess) return nullptr;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    addr[dev] = p;
  }
  return addr[dev];
}

hipError_t ensure_workspace(Workspace &ws, size_t bytes) {
  if (ws.bytes >= bytes) return hipSuccess;
  if (ws.ptr) {
    hipError_t e = hipFree(ws.ptr);
    ws.ptr = nullptr;
    ws.bytes = 0;
    if (e != hipSuccess) return e;
  }
  size_t want = bytes < 256 ? 256 : bytes;
  hipError_t e = hipMalloc(&ws.ptr, want);
  if (e != hipSuccess) {
    ws.ptr = nullptr;
    return e;
  }
  ws.bytes = want;
  return hipSuccess;
}

hipError_t ensure_tickets(hcblasHandle_t h, size_t count) {
  size_t bytes = count * sizeof(unsigned int);
  if (h->gemv_tickets.bytes >= bytes) return hipSuccess;
  hipError_t e = ensure_workspace(h->gemv_tickets, bytes);
  if (e != hipSuccess) return e;
  // Zeroed once at allocation; every kernel that takes a ticket leaves it 0.
  e = hipMemset(h->gemv_tickets.ptr, 0, h->gemv_tickets.bytes);
  if (e != hipSuccess) return e;
  return hipDeviceSynchronize();
}

}  // namespace hcb

using namespace hcb;

extern "C" {

hcblasStatus_t hcblasCreate(hcblasHandle_t *handle) {
  if (handle == nullptr) return HCBLAS_STATUS_INVALID_VALUE;
  *handle = nullptr;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return HCBLAS_STATUS_NOT_INITIALIZED;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess)
    return HCBLAS_STATUS_NOT_INITIALIZED;

  hcblasContext *h = new (std::nothrow) hcblasContext();
  if (h == nullptr) return HCBLAS_STATUS_ALLOC_FAILED;
  h->device = dev;
  h->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  // Device address of this GPU's no-allocation scratch (looked up once per
  // device, with the device current).
  h->noalloc_scratch = scratch_address(dev);
  if (h->noalloc_scratch == nullptr) {
    delete h;
    return HCBLAS_STATUS_INTERNAL_ERROR;
  }

  bool ok = hipMalloc(&h->d_partials, sizeof(double) * kReducePartials) ==
                hipSuccess &&
            hipHostMalloc(&h->h_result, 64,
                          hipHostMallocMapped | hipHostMallocCoherent) ==
                hipSuccess &&
            hipHostGetDevicePointer(reinterpret_cast<void **>(&h->h_result_dev),
                                    h->h_result, 0) == hipSuccess &&
            hipDeviceSynchronize() == hipSuccess;
  if (ok) std::memset(h->h_result, 0, 64);
  if (!ok) {
    if (h->d_partials) (void)hipFree(h->d_partials);
    if (h->h_result) (void)hipHostFree(h->h_result);
    delete h;
    return HCBLAS_STATUS_ALLOC_FAILED;
  }
  h->magic = kHandleMagic;
  *handle = h;
  return HCBLAS_STATUS_SUCCESS;
}

hcblasStatus_t hcblasDestroy(hcblasHandle_t *handle) {
  if (handle == nullptr || !handle_ok(*handle))
    return HCBLAS_STATUS_NOT_INITIALIZED;
  hcblasContext *h = *handle;
  // Work enqueued by this handle may still read the workspaces.
  (void)hipStreamSynchronize(h->stream);
  h->magic = 0;
  (void)hipFree(h->d_partials);
  (void)hipHostFree(h->h_result);
  if (h->gemv_partials.ptr) (void)hipFree(h->gemv_partials.ptr);
  if (h->gemv_tickets.ptr) (void)hipFree(h->gemv_tickets.ptr);
  if (h->gemm_ws.ptr) (void)hipFree(h->gemm_ws.ptr);
  if (h->split_ws.ptr) (void)hipFree(h->split_ws.ptr);
  if (h->adj_ws.ptr) (void)hipFree(h->adj_ws.ptr);
  if (h->kpart_ws.ptr) (void)hipFree(h->kpart_ws.ptr);
  delete h;
  *handle = nullptr;
  return HCBLAS_STATUS_SUCCESS;
}

hcblasStatus_t hcblasSetStream(hcblasHandle_t handle, hipStream_t stream) {
  if (!handle_ok(handle)) return HCBLAS_STATUS_NOT_INITIALIZED;
  handle->stream = stream;
  return HCBLAS_STATUS_SUCCESS;
}

hcblasStatus_t hcblasGetStream(hcblasHandle_t handle, hipStream_t *stream) {
  if (!handle_ok(handle)) return HCBLAS_STATUS_NOT_INITIALIZED;
  if (stream == nullptr) return HCBLAS_STATUS_INVALID_VALUE;
  *stream = handle->stream;
  return HCBLAS_STATUS_SUCCESS;
}

hcblasStatus_t hcblasSetMathMode(hcblasHandle_t handle, int mode) {
  if (!handle_ok(handle)) return HCBLAS_STATUS_NOT_INITIALIZED;
  if (mode & ~(HCBLAS_SPLIT_BF16_MATH | HCBLAS_OZAKI_INT8_MATH |
               HCBLAS_OZAKI_CRT_MATH | HCBLAS_OZAKI_CRT_F32_MATH))
    return HCBLAS_STATUS_INVALID_VALUE;
  handle->math_mode = mode;
  return HCBLAS_STATUS_SUCCESS;
}

hcblasStatus_t hcblasGetMathMode(hcblasHandle_t handle, int *mode) {
  if (!handle_ok(handle)) return HCBLAS_STATUS_NOT_INITIALIZED;
  if (mode == nullptr) return HCBLAS_STATUS_INVALID_VALUE;
  *mode = handle->math_mode;
  return HCBLAS_STATUS_SUCCESS;
}

hcblasStatus_t hcblasSetOzakiSlices(hcblasHandle_t handle, int slices) {
  if (!handle_ok(handle)) return HCBLAS_STATUS_NOT_INITIALIZED;
  if (slices < 3 || slices > 12) return HCBLAS_STATUS_INVALID_VALUE;
  handle->ozaki_slices = slices;
  return HCBLAS_STATUS_SUCCESS;
}

hcblasStatus_t hcblasSetOzakiModuli(hcblasHandle_t handle, int moduli) {
  if (!handle_ok(handle)) return HCBLAS_STATUS_NOT_INITIALIZED;
  if (moduli != 0 && (moduli < 6 || moduli > 20))
    return HCBLAS_STATUS_INVALID_VALUE;
  handle->ozaki_moduli = moduli;
  return HCBLAS_STATUS_SUCCESS;
}

// ---------------------------------------------------------------------------
// Set/Get Vector/Matrix. The reference copies n*elemSize bytes contiguously
// and reports MAPPING_ERROR when that overruns the device allocation
// (hcblas.cpp:160-168,252-260). Here strides and leading dimensions are
// honoured through one pitched copy (hipMemcpy2DAsync) on the handle's
// stream, and the overrun check uses the allocation range reported by
// hipMemGetAddressRange.
// ---------------------------------------------------------------------------
namespace {

// True when [p, p+bytes) lies inside one device allocation.
bool device_range_ok(const void *p, size_t bytes) {
  hipDeviceptr_t base = nullptr;
  size_t size = 0;
  if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  const char *b = static_cast<const char *>(base);
  const char *q = static_cast<const char *>(p);
  return q >= b && q + bytes <= b + size;
}

hcblasStatus_t copy_strided(hcblasHandle_t h, int rows, int cols, int elem,
                            const void *src, int64_t spitch_elems, void *dst,
                            int64_t dpitch_elems, bool to_device) {
  if (rows <= 0 || cols <= 0) return HCBLAS_STATUS_SUCCESS;
  const void *dev = to_device ? dst : src;
  const int64_t dev_pitch = to_device ? dpitch_elems : spitch_elems;
  const int64_t host_pitch = to_device ? spitch_elems : dpitch_elems;
  const size_t span = (size_t)((int64_t)(cols - 1) * dev_pitch + rows) * elem;
  if (!device_range_ok(dev, span)) return HCBLAS_STATUS_MAPPING_ERROR;
  const hipMemcpyKind kind =
      to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
  hipError_t e = hipSuccess;
  if (dev_pitch == rows && host_pitch == rows) {
    // Dense on both sides: one copy.
    e = hipMemcpyAsync(dst, src, (size_t)rows * cols * elem, kind, h->stream);
  } else {
    // Strided: one pitched copy of cols runs of rows elements, ordered on
    // the handle's stream after the work already enqueued there (so a
    // kernel still writing the gaps between the runs is not raced).
    const size_t spitch = (size_t)spitch_elems * elem;
    const size_t dpitch = (size_t)dpitch_elems * elem;
    e = hipMemcpy2DAsync(dst, dpitch, src, spitch, (size_t)rows * elem,
                         (size_t)cols, kind, h->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  return e == hipSuccess ? HCBLAS_STATUS_SUCCESS
                         : HCBLAS_STATUS_MAPPING_ERROR;
}

}  // namespace

hcblasStatus_t hcblasSetVector(hcblasHandle_t handle, int n, int elemSize,
                               const void *x, int incx, void *y, int incy) {
  if (!handle_ok(handle)) return HCBLAS_STATUS_NOT_INITIALIZED;
  if (incx <= 0 || incy <= 0 || elemSize <= 0)
    return HCBLAS_STATUS_INVALID_VALUE;
  if (x == nullptr || y == nullptr) return HCBLAS_STATUS_INVALID_VALUE;
  // A vector with stride inc is a 1-row matrix whose column pitch is inc.
  return copy_strided(handle, 1, n, elemSize, x, incx, y, incy, true);
}

hcblasStatus_t hcblasGetVector(hcblasHandle_t handle, int n, int elemSize,
                               const void *x, int incx, void *y, int incy) {
  if (!handle_ok(handle)) return HCBLAS_STATUS_NOT_INITIALIZED;
  if (incx <= 0 || incy <= 0 || elemSize <= 0)
    return HCBLAS_STATUS_INVALID_VALUE;
  if (x == nullptr || y == nullptr) return HCBLAS_STATUS_INVALID_VALUE;
  return copy_strided(handle, 1, n, elemSize, x, incx, y, incy, false);
}

hcblasStatus_t hcblasSetMatrix(hcblasHandle_t handle, int rows, int cols,
                               int elemSize, const void *A, int lda, void *B,
                               int ldb) {
  if (!handle_ok(handle)) return HCBLAS_STATUS_NOT_INITIALIZED;
  if (rows < 0 || cols < 0 || elemSize <= 0 || lda <= 0 || ldb <= 0)
    return HCBLAS_STATUS_INVALID_VALUE;
  if (A == nullptr || B == nullptr) return HCBLAS_STATUS_INVALID_VALUE;
  if (lda < rows || ldb < rows) return HCBLAS_STATUS_INVALID_VALUE;
  return copy_strided(handle, rows, cols, elemSize, A, lda, B, ldb, true);
}

hcblasStatus_t hcblasGetMatrix(hcblasHandle_t handle, int rows, int cols,
                               int elemSize, const void *A, int lda, void *B,
                               int ldb) {
  if (!handle_ok(handle)) return HCBLAS_STATUS_NOT_INITIALIZED;
  if (rows < 0 || cols < 0 || elemSize <= 0 || lda <= 0 || ldb <= 0)
    return HCBLAS_STATUS_INVALID_VALUE;
  if (A == nullptr || B == nullptr) return HCBLAS_STATUS_INVALID_VALUE;
  if (lda < rows || ldb < rows) return HCBLAS_STATUS_INVALID_VALUE;
  return copy_strided(handle, rows, cols, elemSize, A, lda, B, ldb, false);
}

hcblasStatus_t hcblasGetVersionString(char *buf, int len) {
  if (buf == nullptr || len <= 0) return HCBLAS_STATUS_INVALID_VALUE;
  std::snprintf(buf, (size_t)len, "hcblas-mi355x 0.1 gfx950");
  return HCBLAS_STATUS_SUCCESS;
}

}  // extern "C"
