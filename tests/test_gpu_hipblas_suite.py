"""The reference's hipBLAS wrapper suite (test/unit-hip/test-hipblas.cpp,
40 TEST cases) run against the hipBLAS-compatible library: same routines,
sizes, batch counts, increments, alpha / beta, integer-valued data and CBLAS
oracle (here oracle/refblas.c), same exact comparison (EXPECT_EQ), then
HIPBLAS_STATUS_NOT_INITIALIZED after the handle is gone.

Deviations, each for a reason:
  - data are integers drawn from a seeded numpy generator with the
    reference's moduli (rand_r % 10 / 15 / 20 / 25 there); the values differ,
    the property checked (exact integer results) does not;
  - the reference calls with the destroyed handle's dangling pointer; here
    the NOT_INITIALIZED check uses a null handle (a freed handle is not
    readable memory);
  - where the reference's alpha = 1 makes scal / axpy / ger the identity or
    a plain add, a second call with a non-trivial alpha is checked too;
  - hipblasSetMatrix is called with a destination leading dimension no
    smaller than the row count (the reference's sgemmBatched passes K < M,
    which its wrapper ignored);
  - plain hipblasDgemm, which the reference suite has no case for, is
    checked like hipblasSgemm.
"""
import ctypes

import numpy as np
import pytest

from gpu_util import to_dev, to_host

pytestmark = pytest.mark.gpu
N_OP, T_OP = 111, 112
COL = 1


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


def _ints(rng, n, mod, dt):
    return rng.integers(0, mod, n).astype(dt)


def _cints(rng, n, dt):
    # X[i].x = rand % 10, X[i].y = rand % 20
    return (rng.integers(0, 10, n) + 1j * rng.integers(0, 20, n)).astype(dt)


_R = {"S": (np.float32, ctypes.c_float), "D": (np.float64, ctypes.c_double)}


class _CF(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float), ("y", ctypes.c_float)]


class _CD(ctypes.Structure):
    _fields_ = [("x", ctypes.c_double), ("y", ctypes.c_double)]


def _set_get_vector(shim, host, dev):
    es = host.dtype.itemsize
    assert shim.hipblasSetVector(host.size, es, host.ctypes.data, 1, _p(dev), 1) == 0


def _get_vector(shim, dev, dt, n):
    out = np.zeros(n, dt)
    assert shim.hipblasGetVector(n, out.dtype.itemsize, _p(dev), 1,
                                 out.ctypes.data, 1) == 0
    return out


# -- asum (test-hipblas.cpp:49-205) -------------------------------------------
@pytest.mark.parametrize("p", ["S", "D"])
@pytest.mark.parametrize("batch", [0, 128])
def test_asum(shim, hh, orc, torch_cuda, p, batch):
    dt, ct = _R[p]
    rng = np.random.default_rng(1)
    n = 23
    X = _ints(rng, n * max(batch, 1), 10, dt)
    dX = torch_cuda.empty(X.size, dtype=getattr(torch_cuda, dt.__name__), device="cuda")
    _set_get_vector(shim, X, dX)
    r = ct()
    if batch:
        st = getattr(shim, f"hipblas{p}asumBatched")(hh, n, _p(dX), 1, ctypes.byref(r), batch)
        exp = sum(orc.asum(X[i * n:(i + 1) * n]) for i in range(batch))
    else:
        st = getattr(shim, f"hipblas{p}asum")(hh, n, _p(dX), 1, ctypes.byref(r))
        exp = orc.asum(X)
    assert st == 0 and r.value == exp
    fn = getattr(shim, f"hipblas{p}asumBatched" if batch else f"hipblas{p}asum")
    args = (None, n, _p(dX), 1, ctypes.byref(r)) + ((batch,) if batch else ())
    assert fn(*args) == 1


# -- scal (test-hipblas.cpp:207-825) ------------------------------------------
_SCAL = {  # routine: (data dtype, alpha ctype, reference alpha, extra alpha)
    "S": (np.float32, ctypes.c_float, 1.0, 2.5),
    "D": (np.float64, ctypes.c_double, 1.0, 2.5),
    "C": (np.complex64, _CF, 1 + 1j, 2 - 0.5j),
    "Z": (np.complex128, _CD, 1 + 1j, 2 - 0.5j),
    "Cs": (np.complex64, ctypes.c_float, 1.0, -1.5),
    "Zd": (np.complex128, ctypes.c_double, 1.0, -1.5),
}


@pytest.mark.parametrize("p", list(_SCAL))
@pytest.mark.parametrize("batch", [0, 128])
def test_scal(shim, hh, orc, p, batch):
    dt, act, a_ref, a_more = _SCAL[p]
    rng = np.random.default_rng(2)
    n = 123
    nb = max(batch, 1)
    X = (_cints(rng, n * nb, dt) if np.issubdtype(dt, np.complexfloating)
         else _ints(rng, n * nb, 10, dt))
    fn = getattr(shim, f"hipblas{p}scalBatched" if batch else f"hipblas{p}scal")
    for alpha in (a_ref, a_more):
        a = act(alpha.real, alpha.imag) if act in (_CF, _CD) else act(alpha)
        dX = to_dev(X)
        args = (hh, n, ctypes.byref(a), _p(dX), 1) + ((batch,) if batch else ())
        assert fn(*args) == 0
        got = to_host(dX)
        exp = np.concatenate([orc.scal(alpha, X[i * n:(i + 1) * n]) for i in range(nb)])
        assert np.array_equal(got, exp), alpha
    args = (None, n, ctypes.byref(a), _p(dX), 1) + ((batch,) if batch else ())
    assert fn(*args) == 1


# -- copy (test-hipblas.cpp:827-1039) -----------------------------------------
@pytest.mark.parametrize("p", ["S", "D"])
@pytest.mark.parametrize("batch", [0, 32])
def test_copy(shim, hh, orc, p, batch):
    dt, _ = _R[p]
    rng = np.random.default_rng(3)
    n = 123
    nb = max(batch, 1)
    X = _ints(rng, n * nb, 10, dt)
    Y = _ints(rng, n * nb, 15, dt)
    dX, dY = to_dev(X), to_dev(Y)
    fn = getattr(shim, f"hipblas{p}copyBatched" if batch else f"hipblas{p}copy")
    args = (hh, n, _p(dX), 1, _p(dY), 1) + ((batch,) if batch else ())
    assert fn(*args) == 0
    assert np.array_equal(to_host(dY), X)
    args = (None,) + args[1:]
    assert fn(*args) == 1


# -- dot (test-hipblas.cpp:1041-1237) -----------------------------------------
@pytest.mark.parametrize("p", ["S", "D"])
@pytest.mark.parametrize("batch", [0, 32])
def test_dot(shim, hh, orc, p, batch):
    dt, ct = _R[p]
    rng = np.random.default_rng(4)
    n = 123
    nb = max(batch, 1)
    X = _ints(rng, n * nb, 10, dt)
    Y = _ints(rng, n * nb, 15, dt)
    dX, dY = to_dev(X), to_dev(Y)
    r = ct()
    fn = getattr(shim, f"hipblas{p}dotBatched" if batch else f"hipblas{p}dot")
    args = (hh, n, _p(dX), 1, _p(dY), 1, ctypes.byref(r)) + ((batch,) if batch else ())
    assert fn(*args) == 0
    exp = sum(orc.dot(X[i * n:(i + 1) * n], Y[i * n:(i + 1) * n]) for i in range(nb))
    assert r.value == exp
    assert fn(*((None,) + args[1:])) == 1


# -- axpy (test-hipblas.cpp:1239-1346) ----------------------------------------
@pytest.mark.parametrize("batch", [0, 32])
def test_saxpy(shim, hh, orc, batch):
    rng = np.random.default_rng(5)
    n = 123
    nb = max(batch, 1)
    X = _ints(rng, n * nb, 10, np.float32)
    Y = _ints(rng, n * nb, 15, np.float32)
    fn = shim.hipblasSaxpyBatched if batch else shim.hipblasSaxpy
    for alpha in (1.0, 0.75):
        dX, dY = to_dev(X), to_dev(Y)
        a = ctypes.c_float(alpha)
        args = (hh, n, ctypes.byref(a), _p(dX), 1, _p(dY), 1) + ((batch,) if batch else ())
        assert fn(*args) == 0
        exp = np.concatenate([orc.axpy(alpha, X[i * n:(i + 1) * n], Y[i * n:(i + 1) * n])
                              for i in range(nb)])
        assert np.array_equal(to_host(dY), exp), alpha
    assert fn(*((None,) + args[1:])) == 1


# -- ger (test-hipblas.cpp:1348-1620) -----------------------------------------
@pytest.mark.parametrize("p", ["S", "D"])
@pytest.mark.parametrize("batch", [0, 32])
def test_ger(shim, hh, orc, p, batch):
    dt, ct = _R[p]
    rng = np.random.default_rng(6)
    m, n = (123, 67) if batch else (123, 78)
    nb = max(batch, 1)
    if batch:  # X = 1, Y = 2, A = 3 (test-hipblas.cpp:1501-1509)
        X = np.ones(m * nb, dt)
        Y = np.full(n * nb, 2, dt)
        A = np.full(m * n * nb, 3, dt)
    else:
        X = _ints(rng, m, 10, dt)
        Y = _ints(rng, n, 15, dt)
        A = _ints(rng, m * n, 25, dt)
    fn = getattr(shim, f"hipblas{p}gerBatched" if batch else f"hipblas{p}ger")
    dX, dY = to_dev(X), to_dev(Y)  # kept alive until the kernels ran
    for alpha in (1.0, 1.5):
        dA = to_dev(A)
        a = ct(alpha)
        args = (hh, m, n, ctypes.byref(a), _p(dX), 1, _p(dY), 1,
                _p(dA), m) + ((batch,) if batch else ())
        assert fn(*args) == 0
        exp = np.concatenate([orc.ger(COL, m, n, alpha, X[i * m:(i + 1) * m], 1,
                                      Y[i * n:(i + 1) * n], 1,
                                      A[i * m * n:(i + 1) * m * n], m)
                              for i in range(nb)])
        assert np.array_equal(to_host(dA), exp), alpha
    assert fn(*((None,) + args[1:])) == 1


# -- gemv (test-hipblas.cpp:1622-1789) ----------------------------------------
@pytest.mark.parametrize("batch", [0, 32])
@pytest.mark.parametrize("trans", [N_OP, T_OP])
def test_sgemv(shim, hh, orc, batch, trans):
    rng = np.random.default_rng(7)
    m, n = (123, 67) if batch else (123, 78)
    nb = max(batch, 1)
    lx, ly = (n, m) if trans == N_OP else (m, n)
    X = _ints(rng, lx * nb, 10, np.float32)
    Y = _ints(rng, ly * nb, 15, np.float32)
    A = _ints(rng, m * n * nb, 25, np.float32)
    dA, dX, dY = to_dev(A), to_dev(X), to_dev(Y)
    one = ctypes.c_float(1.0)
    fn = shim.hipblasSgemvBatched if batch else shim.hipblasSgemv
    args = (hh, trans, m, n, ctypes.byref(one), _p(dA), m, _p(dX), 1,
            ctypes.byref(one), _p(dY), 1) + ((batch,) if batch else ())
    assert fn(*args) == 0
    exp = np.concatenate([orc.gemv(COL, trans - N_OP, m, n, 1.0,
                                   A[i * m * n:(i + 1) * m * n], m,
                                   X[i * lx:(i + 1) * lx], 1, 1.0,
                                   Y[i * ly:(i + 1) * ly], 1) for i in range(nb)])
    assert np.array_equal(to_host(dY), exp)
    assert fn(*((None,) + args[1:])) == 1


# -- gemm (test-hipblas.cpp:1791-2700) ----------------------------------------
_GEMM = {"S": (np.float32, ctypes.c_float), "D": (np.float64, ctypes.c_double),
         "C": (np.complex64, _CF), "Z": (np.complex128, _CD)}


def _scalar(ct, v):
    return ct(v.real, v.imag) if ct in (_CF, _CD) else ct(v)


@pytest.mark.parametrize("p", ["S", "D", "C", "Z"])
@pytest.mark.parametrize("batch", [0, 32])
def test_gemm(shim, hh, orc, torch_cuda, p, batch):
    """A = 1, B = 2, C = 3 and alpha = beta = 1 give C = 2K + 3
    (test-hipblas.cpp:1815-1853), through SetMatrix / GetMatrix; complex
    cases use integer-valued complex data."""
    torch = torch_cuda
    dt, ct = _GEMM[p]
    rng = np.random.default_rng(8)
    M, N, K = 123, 78, 23
    nb = max(batch, 1)
    cplx = np.issubdtype(dt, np.complexfloating)
    if cplx:
        As = [_cints(rng, M * K, dt) for _ in range(nb)]
        Bs = [_cints(rng, K * N, dt) for _ in range(nb)]
        Cs = [_cints(rng, M * N, dt) for _ in range(nb)]
        alpha, beta = 1 + 1j, 1 - 1j
    else:
        As = [np.ones(M * K, dt) for _ in range(nb)]
        Bs = [np.full(K * N, 2, dt) for _ in range(nb)]
        Cs = [np.full(M * N, 3, dt) for _ in range(nb)]
        alpha, beta = 1.0, 1.0
    es = np.dtype(dt).itemsize
    dA = [torch.empty(M * K * es, dtype=torch.uint8, device="cuda") for _ in range(nb)]
    dB = [torch.empty(K * N * es, dtype=torch.uint8, device="cuda") for _ in range(nb)]
    dC = [torch.empty(M * N * es, dtype=torch.uint8, device="cuda") for _ in range(nb)]
    for i in range(nb):
        for host, dev, r, c in ((As[i], dA[i], M, K), (Bs[i], dB[i], K, N),
                                (Cs[i], dC[i], M, N)):
            assert shim.hipblasSetMatrix(r, c, es, host.ctypes.data, r, _p(dev), r) == 0
    a, b = _scalar(ct, alpha), _scalar(ct, beta)
    if batch:
        ptrs = [torch.tensor([t.data_ptr() for t in lst], dtype=torch.int64, device="cuda")
                for lst in (dA, dB, dC)]
        fn = getattr(shim, f"hipblas{p}gemmBatched")
        args = (hh, N_OP, N_OP, M, N, K, ctypes.byref(a), _p(ptrs[0]), M, _p(ptrs[1]), K,
                ctypes.byref(b), _p(ptrs[2]), M, batch)
    else:
        fn = getattr(shim, f"hipblas{p}gemm")
        args = (hh, N_OP, N_OP, M, N, K, ctypes.byref(a), _p(dA[0]), M, _p(dB[0]), K,
                ctypes.byref(b), _p(dC[0]), M)
    assert fn(*args) == 0
    for i in range(nb):
        out = np.zeros(M * N, dt)
        assert shim.hipblasGetMatrix(M, N, es, _p(dC[i]), M, out.ctypes.data, M) == 0
        exp = orc.gemm(COL, 0, 0, M, N, K, alpha, As[i], M, Bs[i], K, beta, Cs[i], M)
        if not cplx:
            assert np.all(exp == 2 * K + 3)
        assert np.array_equal(out, exp), i
    assert fn(*((None,) + args[1:])) == 1


def test_hgemm(shim, hh, orc, torch_cuda):
    """test-hipblas.cpp:2396-2480 (compiled out in the reference): integer
    binary16 data, FP32 accumulation, exact while every sum fits binary16."""
    torch = torch_cuda
    rng = np.random.default_rng(9)
    M, N, K = 123, 78, 23
    A = rng.integers(0, 4, M * K).astype(np.float16)
    B = rng.integers(0, 4, K * N).astype(np.float16)
    C = rng.integers(0, 8, M * N).astype(np.float16)
    dA, dB, dC = (to_dev(x.view(np.uint16)) for x in (A, B, C))
    one = ctypes.c_uint16(np.float16(1.0).view(np.uint16))
    assert shim.hipblasHgemm(hh, N_OP, N_OP, M, N, K, ctypes.byref(one), _p(dA), M,
                             _p(dB), K, ctypes.byref(one), _p(dC), M) == 0
    got = to_host(dC).view(np.float16).astype(np.float64)
    exp = (A.astype(np.float64).reshape(K, M).T @ B.astype(np.float64).reshape(N, K).T
           + C.astype(np.float64).reshape(N, M).T)
    assert np.array_equal(got.reshape(N, M).T, exp)
    assert shim.hipblasHgemm(None, N_OP, N_OP, M, N, K, ctypes.byref(one), _p(dA), M,
                             _p(dB), K, ctypes.byref(one), _p(dC), M) == 1
