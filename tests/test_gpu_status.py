"""Status-code contract with a live handle (reference: lib/src/hcblas.cpp
argument checks, test/unit-api/testHelperFunctions.cpp and the
NOT_INITIALIZED-after-destroy checks of test/unit-api/test-blas.cpp).

Documented differences (CBLAS semantics): empty problems return SUCCESS
(reference: EXECUTION_FAILED); negative sizes, zero increments and leading
dimensions below the stored rows return INVALID_VALUE.
"""
import ctypes

import numpy as np
import pytest

from gpu_util import to_dev, to_host

pytestmark = pytest.mark.gpu


def test_create_destroy_double_destroy(hc, torch_cuda):
    lib = hc.load()
    h = ctypes.c_void_p()
    assert lib.hcblasCreate(ctypes.byref(h)) == 0 and h.value
    assert lib.hcblasDestroy(ctypes.byref(h)) == 0
    assert h.value is None
    assert lib.hcblasDestroy(ctypes.byref(h)) == 1
    assert lib.hcblasSgemm(h, 0, 0, 1, 1, 1, None, None, 1, None, 1, None,
                           None, 1) == 1


def test_stream_roundtrip(hc, handle, torch_cuda):
    torch = torch_cuda
    lib = hc.load()
    s = torch.cuda.Stream()
    assert lib.hcblasSetStream(handle.value, ctypes.c_void_p(s.cuda_stream)) == 0
    out = ctypes.c_void_p()
    assert lib.hcblasGetStream(handle.value, ctypes.byref(out)) == 0
    assert out.value == s.cuda_stream
    # work runs on that stream: axpy then wait on the stream only
    x = torch.ones(1 << 20, device="cuda")
    y = torch.zeros(1 << 20, device="cuda")
    torch.cuda.synchronize()
    assert hc.api.hcblasSaxpy(handle, x.numel(), 2.0, x, 1, y, 1) == 0
    s.synchronize()
    assert torch.all(y == 2)
    assert lib.hcblasSetStream(handle.value, None) == 0
    assert lib.hcblasGetStream(handle.value, ctypes.byref(out)) == 0
    assert out.value is None


def test_argument_errors(hc, handle, torch_cuda):
    api = hc.api
    x = torch_cuda.zeros(64, device="cuda")
    # gemm: negative sizes (hcblas.cpp:1311), bad ld
    assert api.hcblasSgemm(handle, 0, 0, -1, 2, 2, 1.0, x, 2, x, 2, 0.0, x, 2) == 3
    assert api.hcblasSgemm(handle, 0, 0, 4, 4, 4, 1.0, x, 3, x, 4, 0.0, x, 4) == 3
    assert api.hcblasSgemm(handle, 7, 0, 4, 4, 4, 1.0, x, 4, x, 4, 0.0, x, 4) == 3
    # empty problems quick-return
    assert api.hcblasSgemm(handle, 0, 0, 0, 4, 4, 1.0, x, 1, x, 4, 0.0, x, 1) == 0
    # batched: negative batch (hcblas.cpp:1532)
    assert api.hcblasSgemmBatched(handle, 0, 0, 2, 2, 2, 1.0, x, 2, x, 2, 0.0,
                                  x, 2, -1) == 3
    # gemv / ger: inc == 0 or negative sizes (hcblas.cpp:974, 1136)
    assert api.hcblasSgemv(handle, 0, 4, 4, 1.0, x, 4, x, 0, 0.0, x, 1) == 3
    assert api.hcblasSgemv(handle, 0, -4, 4, 1.0, x, 4, x, 1, 0.0, x, 1) == 3
    assert api.hcblasSger(handle, 4, 4, 1.0, x, 1, x, 0, x, 4) == 3
    assert api.hcblasSger(handle, 4, 4, 1.0, x, 1, x, 1, x, 2) == 3
    # level 1: zero increment is invalid; n <= 0 is a no-op
    assert api.hcblasSaxpy(handle, 4, 1.0, x, 0, x, 1) == 3
    assert api.hcblasSaxpy(handle, 0, 1.0, x, 1, x, 1) == 0
    st, r = api.hcblasSdot(handle, 0, x, 1, x, 1, None)
    assert st == 0 and r == 0.0
    # math mode
    lib = hc.load()
    assert lib.hcblasSetMathMode(handle.value, 99) == 3
    assert lib.hcblasSetMathMode(handle.value, 1) == 0
    mode = ctypes.c_int()
    assert lib.hcblasGetMathMode(handle.value, ctypes.byref(mode)) == 0
    assert mode.value == 1
    assert lib.hcblasSetMathMode(handle.value, 0) == 0
    assert lib.hcblasSetOzakiSlices(handle.value, 2) == 3


def _hip():
    # By SONAME, so this resolves to the HIP runtime already loaded in the
    # process (torch's), the one libhcblas is bound to; the unversioned name
    # would load a second, separate runtime from /opt/rocm.
    lib = ctypes.CDLL("libamdhip64.so.7")
    lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.hipFree.argtypes = [ctypes.c_void_p]
    return lib


def test_set_get_vector_matrix(hc, handle, torch_cuda):
    """testHelperFunctions.cpp: SUCCESS, INVALID_VALUE (inc 0, elemSize 0),
    MAPPING_ERROR (a copy larger than the device allocation)."""
    lib = hc.load()
    hip = _hip()
    n = 10
    d = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(d), n * 4) == 0
    try:
        x = (ctypes.c_float * (2 * n))(*range(2 * n))
        y = (ctypes.c_float * (2 * n))()
        assert lib.hcblasSetVector(handle.value, n, 4, x, 1, d, 1) == 0
        assert lib.hcblasGetVector(handle.value, n, 4, d, 1, y, 1) == 0
        assert list(y)[:n] == list(range(n))
        # strided: every second host element into consecutive device slots
        assert lib.hcblasSetVector(handle.value, n, 4, x, 2, d, 1) == 0
        assert lib.hcblasGetVector(handle.value, n, 4, d, 1, y, 1) == 0
        assert list(y)[:n] == list(range(0, 2 * n, 2))
        assert lib.hcblasSetVector(handle.value, n, 4, x, 0, d, 1) == 3
        assert lib.hcblasSetVector(handle.value, n, 4, x, 1, d, 0) == 3
        assert lib.hcblasSetVector(handle.value, n, 0, x, 1, d, 1) == 3
        # 10 doubles into a 40-byte allocation
        assert lib.hcblasSetVector(handle.value, n, 8, x, 1, d, 1) == 4
        assert lib.hcblasGetVector(handle.value, n, 8, d, 1, y, 1) == 4
        # matrix 5 x 2 with ld 5 fits; ld 6 does not
        assert lib.hcblasSetMatrix(handle.value, 5, 2, 4, x, 5, d, 5) == 0
        assert lib.hcblasGetMatrix(handle.value, 5, 2, 4, d, 5, y, 5) == 0
        assert lib.hcblasSetMatrix(handle.value, 5, 2, 4, x, 5, d, 6) == 4
    finally:
        hip.hipFree(d)
    h2 = ctypes.c_void_p(None)
    assert lib.hcblasSetVector(h2, n, 4, x, 1, d, 1) == 1


def test_destroyed_handle_every_family(hc, torch_cuda):
    lib = hc.load()
    h = ctypes.c_void_p()
    assert lib.hcblasCreate(ctypes.byref(h)) == 0
    assert lib.hcblasDestroy(ctypes.byref(h)) == 0
    api = hc.api
    x = torch_cuda.zeros(16, device="cuda")
    assert api.hcblasSaxpy(h.value, 4, 1.0, x, 1, x, 1) == 1
    assert api.hcblasDgemv(h.value, 0, 2, 2, 1.0, x, 2, x, 1, 0.0, x, 1) == 1
    assert api.hcblasCgemm(h.value, 0, 0, 1, 1, 1, 1j, x, 1, x, 1, 0j, x, 1) == 1
    assert api.hcblasSdot(h.value, 4, x, 1, x, 1, None)[0] == 1
