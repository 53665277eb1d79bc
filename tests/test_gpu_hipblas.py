"""hipBLAS-compatible entry points on the GPU (reference: the hipBLAS shim
lib/src/hcc_detail/hipblas.cpp and its tests test/unit-hip/test-hipblas.cpp).

Each routine is checked against the library entry it forwards to or against
the reference tests' closed forms:
  - sgemm with A = 1, B = 2, C = 3, alpha = beta = 1 gives C = 2K + 3
    (test-hipblas.cpp:1815-1853), through hipblasSetMatrix / GetMatrix;
  - hipBLAS operation codes 111 / 112 / 113 select N / T / C;
  - results are bit-identical to the corresponding hcblas entry;
  - a destroyed handle gives HIPBLAS_STATUS_NOT_INITIALIZED.
"""
import ctypes

import numpy as np
import pytest

from gpu_util import bitwise_equal, to_dev, to_host

pytestmark = pytest.mark.gpu


@pytest.fixture()
def shim(torch_cuda):
    from hcblas_amd import hipblas_shim
    return hipblas_shim.load()


@pytest.fixture()
def hh(shim):
    h = ctypes.c_void_p()
    assert shim.hipblasCreate(ctypes.byref(h)) == 0
    yield h
    shim.hipblasDestroy(h)


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def test_sgemm_closed_form_through_set_get_matrix(shim, hh, torch_cuda):
    torch = torch_cuda
    from hcblas_amd import hipblas_shim as hs
    M, N, K = 123, 78, 23
    A = np.ones(M * K, np.float32)
    B = np.full(K * N, 2, np.float32)
    C = np.full(M * N, 3, np.float32)
    dA = torch.empty(M * K, device="cuda")
    dB = torch.empty(K * N, device="cuda")
    dC = torch.empty(M * N, device="cuda")
    for host, dev, r, c in ((A, dA, M, K), (B, dB, K, N), (C, dC, M, N)):
        assert shim.hipblasSetMatrix(r, c, 4, host.ctypes.data, r, _p(dev), r) == 0
    one = ctypes.c_float(1.0)
    assert shim.hipblasSgemm(hh, hs.OP_N, hs.OP_N, M, N, K, ctypes.byref(one),
                             _p(dA), M, _p(dB), K, ctypes.byref(one), _p(dC),
                             M) == 0
    out = np.zeros(M * N, np.float32)
    assert shim.hipblasGetMatrix(M, N, 4, _p(dC), M, out.ctypes.data, M) == 0
    assert np.all(out == 2 * K + 3)


@pytest.mark.parametrize("ops", [(111, 111), (112, 111), (111, 112), (112, 112),
                                 (113, 113)])
def test_gemm_ops_match_library(shim, hh, hc, handle, orc, ops):
    """hipblasSgemm / Dgemm with each hipBLAS op code equals hcblasSgemm /
    Dgemm with the matching library code, bit for bit."""
    api = hc.api
    M, N, K = 67, 45, 39
    ta, tb = ops[0] - 111, ops[1] - 111
    for p, ct, gen in (("S", ctypes.c_float, orc.uniform_f32),
                       ("D", ctypes.c_double, orc.uniform_f64)):
        A = gen(M * K, 1)
        B = gen(K * N, 2)
        C = gen(M * N, 3)
        lda = M if ta == 0 else K
        ldb = K if tb == 0 else N
        dA, dB = to_dev(A), to_dev(B)
        d1, d2 = to_dev(C), to_dev(C)
        al, be = ct(1.5), ct(-0.5)
        st = getattr(shim, f"hipblas{p}gemm")(hh, ops[0], ops[1], M, N, K,
                                              ctypes.byref(al), _p(dA), lda,
                                              _p(dB), ldb, ctypes.byref(be),
                                              _p(d1), M)
        assert st == 0
        st = getattr(api, f"hcblas{p}gemm")(handle, ta, tb, M, N, K, 1.5, dA,
                                            lda, dB, ldb, -0.5, d2, M)
        assert st == 0
        assert bitwise_equal(to_host(d1), to_host(d2)), p


def test_level1_and_level2_forward(shim, hh, hc, handle, orc, torch_cuda):
    """asum / dot / axpy / scal / copy / gemv / ger through the shim equal the
    library's results bit for bit."""
    api = hc.api
    n = 100003
    x = orc.uniform_f32(n, 4)
    y = orc.uniform_f32(n, 5)
    dx, dy = to_dev(x), to_dev(y)
    r = ctypes.c_float()
    assert shim.hipblasSdot(hh, n, _p(dx), 1, _p(dy), 1, ctypes.byref(r)) == 0
    st, ref = api.hcblasSdot(handle, n, dx, 1, dy, 1, None)
    assert st == 0 and r.value == ref
    assert shim.hipblasSasum(hh, n, _p(dx), 1, ctypes.byref(r)) == 0
    st, ref = api.hcblasSasum(handle, n, dx, 1, None)
    assert st == 0 and r.value == ref
    a = ctypes.c_float(0.75)
    y1, y2 = to_dev(y), to_dev(y)
    assert shim.hipblasSaxpy(hh, n, ctypes.byref(a), _p(dx), 1, _p(y1), 1) == 0
    assert api.hcblasSaxpy(handle, n, 0.75, dx, 1, y2, 1) == 0
    assert bitwise_equal(to_host(y1), to_host(y2))
    assert shim.hipblasSscal(hh, n, ctypes.byref(a), _p(y1), 1) == 0
    assert api.hcblasSscal(handle, n, 0.75, y2, 1) == 0
    assert bitwise_equal(to_host(y1), to_host(y2))
    M, N = 300, 200
    A = orc.uniform_f32(M * N, 6)
    xv = orc.uniform_f32(N, 7)
    yv = orc.uniform_f32(M, 8)
    dA = to_dev(A)
    g1, g2 = to_dev(yv), to_dev(yv)
    dxg = to_dev(xv)  # kept alive until the kernel ran
    b = ctypes.c_float(0.5)
    assert shim.hipblasSgemv(hh, 111, M, N, ctypes.byref(a), _p(dA), M,
                             _p(dxg), 1, ctypes.byref(b), _p(g1), 1) == 0
    assert api.hcblasSgemv(handle, 0, M, N, 0.75, dA, M, to_dev(xv), 1, 0.5,
                           g2, 1) == 0
    assert bitwise_equal(to_host(g1), to_host(g2))
    A1, A2 = to_dev(A), to_dev(A)
    dxv, dyv = to_dev(yv), to_dev(xv)
    assert shim.hipblasSger(hh, M, N, ctypes.byref(a), _p(dxv), 1, _p(dyv), 1,
                            _p(A1), M) == 0
    assert api.hcblasSger(handle, M, N, 0.75, dxv, 1, dyv, 1, A2, M) == 0
    assert bitwise_equal(to_host(A1), to_host(A2))


def test_stream_and_destroyed_handle(shim, torch_cuda):
    torch = torch_cuda
    h = ctypes.c_void_p()
    assert shim.hipblasCreate(ctypes.byref(h)) == 0
    s = torch.cuda.Stream()
    assert shim.hipblasSetStream(h, ctypes.c_void_p(s.cuda_stream)) == 0
    got = ctypes.c_void_p()
    assert shim.hipblasGetStream(h, ctypes.byref(got)) == 0
    assert got.value == s.cuda_stream
    assert shim.hipblasDestroy(h) == 0
    one = ctypes.c_float(1.0)
    x = torch.zeros(16, device="cuda")
    # the handle object is gone: any call on it reports NOT_INITIALIZED
    # (test-hipblas.cpp:1856-1860), checked here on a fresh null handle
    assert shim.hipblasSscal(None, 16, ctypes.byref(one), _p(x), 1) == 1
