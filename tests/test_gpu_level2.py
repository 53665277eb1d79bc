"""GEMV / GER parity on the GPU through the C ABI.

Reference tests mirrored: test/unit/hcblas_sgemv.cpp, hcblas_sger.cpp
(integer recipes, both storage orders, compared exactly with CBLAS) and the
batched forms; plus seeded real-valued data with leading-dimension padding,
offsets, increments and the alpha == 0 / beta == 0 paths. GER is bit-exact
(fma(x_i, alpha*y_j, a)); GEMV is within a c*n*eps bound of the
double-double value.
"""
import numpy as np
import pytest

from gpu_util import bitwise_equal, to_dev, to_host

pytestmark = pytest.mark.gpu
ROW, COL = 0, 1
N_, T_ = 0, 1


@pytest.mark.parametrize("p", ["s", "d"])
def test_reference_recipes(hc, handle, golden, p):
    api = hc.api
    g = golden("level2_ref_int.npz")
    M, N = (int(v) for v in g["dims"])
    gemv = getattr(api, f"hcblas_{p}gemv")
    for trans in (N_, T_):
        for order in (COL, ROW):
            key = f"{p}gemv_t{trans}_o{order}"
            lda = M if order == COL else N
            A, x, y = (to_dev(g[key + s]) for s in ("_A", "_x", "_y"))
            st = gemv(handle, order, trans, M, N, 1.0, A, 0, lda, x, 0, 1, 1.0,
                      y, 0, 1)
            assert st == 0
            assert bitwise_equal(to_host(y), g[key + "_expect"]), key
    ger = getattr(api, f"hcblas_{p}ger")
    for order in (COL, ROW):
        lda = M if order == COL else N
        for alpha in (1, 3):
            key = f"{p}ger_o{order}_a{alpha}"
            x, y, A = (to_dev(g[key + s]) for s in ("_x", "_y", "_A"))
            st = ger(handle, order, M, N, float(alpha), x, 0, 1, y, 0, 1, A, 0,
                     lda)
            assert st == 0
            assert bitwise_equal(to_host(A), g[key + "_expect"]), key


SHAPES = [(1, 1), (7, 3), (64, 64), (255, 257), (1000, 33), (33, 1000),
          (4096, 96), (96, 4096)]


@pytest.mark.parametrize("p", ["s", "d"])
@pytest.mark.parametrize("order", [COL, ROW])
@pytest.mark.parametrize("trans", [N_, T_])
@pytest.mark.parametrize("M,N", SHAPES)
def test_gemv_seeded(hc, handle, orc, p, order, trans, M, N):
    api = hc.api
    dt = np.float32 if p == "s" else np.float64
    gen = orc.uniform_f32 if p == "s" else orc.uniform_f64
    rows = M if order == COL else N
    cols = N if order == COL else M
    for lda, off, incx, incy in ((rows, 0, 1, 1), (rows + 3, 1, 2, -1)):
        A = gen(lda * cols + off, M * 7 + N)
        lenx, leny = (N, M) if trans == N_ else (M, N)
        x = gen(lenx * abs(incx), 3)
        y = gen(leny * abs(incy), 4)
        for alpha, beta in ((1.5, 0.5), (1.0, 0.0), (0.0, 2.0)):
            yy = y.copy()
            if beta == 0.0:
                yy[:] = np.nan  # must not propagate
            dA, dx, dy = to_dev(A), to_dev(x), to_dev(yy)
            st = getattr(api, f"hcblas_{p}gemv")(
                handle, order, trans, M, N, alpha, dA, off, lda, dx, 0, incx,
                beta, dy, 0, incy)
            assert st == 0
            got = to_host(dy)
            ref = orc.gemv_exact(order, trans, M, N, alpha, A[off:], lda, x,
                                 incx, beta, yy, incy)
            # ref is in BLAS order of y elements; map increments back
            idx = (np.arange(leny) * incy if incy > 0 else
                   (leny - 1 - np.arange(leny)) * (-incy))
            eps = np.finfo(dt).eps
            tol = 2 * (lenx + 2) * eps * (abs(alpha) * lenx + abs(beta) + 1)
            assert np.all(np.isfinite(got[idx]))
            assert np.max(np.abs(got[idx] - ref)) <= tol, (alpha, beta)


@pytest.mark.parametrize("p", ["s", "d"])
@pytest.mark.parametrize("order", [COL, ROW])
@pytest.mark.parametrize("M,N", SHAPES)
def test_ger_bit_exact(hc, handle, orc, p, order, M, N):
    api = hc.api
    gen = orc.uniform_f32 if p == "s" else orc.uniform_f64
    rows = M if order == COL else N
    cols = N if order == COL else M
    for lda, off, incx, incy in ((rows, 0, 1, 1), (rows + 5, 3, -2, 3)):
        A = gen(lda * cols + off, 99)
        x = gen(M * abs(incx), 1)
        y = gen(N * abs(incy), 2)
        dA = to_dev(A)
        st = getattr(api, f"hcblas_{p}ger")(handle, order, M, N, 1.37,
                                            to_dev(x), 0, incx, to_dev(y), 0,
                                            incy, dA, off, lda)
        assert st == 0
        exp = A.copy()
        exp[off:] = orc.ger(order, M, N, 1.37, x, incx, y, incy, A[off:], lda)
        assert bitwise_equal(to_host(dA), exp)


@pytest.mark.parametrize("p,M,N,batch", [("d", 6144, 6144, 1),
                                         ("s", 8448, 8448, 1),
                                         ("d", 2048, 2048, 9)])
def test_ger_streaming_bit_exact(hc, handle, orc, p, M, N, batch):
    """Matrices above 256 MiB (single and batched) take the non-temporal
    load / store path; bit-exact against the oracle like the cached one."""
    api = hc.api
    gen = orc.uniform_f32 if p == "s" else orc.uniform_f64
    A = gen(M * N * batch, 7)
    assert A.nbytes > 256 << 20
    x, y = gen(M * batch, 8), gen(N * batch, 9)
    dA = to_dev(A)
    if batch == 1:
        st = getattr(api, f"hcblas_{p}ger")(handle, COL, M, N, -0.75,
                                            to_dev(x), 0, 1, to_dev(y), 0, 1,
                                            dA, 0, M)
    else:
        st = getattr(api, f"hcblas{p.upper()}gerBatched")(
            handle, M, N, -0.75, to_dev(x), 1, to_dev(y), 1, dA, M, batch)
    assert st == 0
    got = to_host(dA)
    for i in range(batch):
        sl = slice(i * M * N, (i + 1) * M * N)
        exp = orc.ger(COL, M, N, -0.75, x[i * M:(i + 1) * M], 1,
                      y[i * N:(i + 1) * N], 1, A[sl], M)
        assert bitwise_equal(got[sl], exp)


# Shapes that reach the band kernels: (p, order, trans, M, N, batch).
# Column-contiguous case (col-major T, row-major N): bands of 16 or 8 row
# steps of 256 x 16 bytes with enough column blocks of 8 to fill the GPU.
# Row-contiguous case (col-major N, row-major T): bands of 4 row steps, at
# least 64 columns, the columns split over blocks of >= 256. The single
# cases have a partial last band.
BAND_CASES = [("s", COL, T_, 18384, 4096, 1), ("s", ROW, N_, 4096, 9192, 1),
              ("d", COL, T_, 12000, 4096, 1), ("d", ROW, N_, 8192, 5000, 1),
              ("s", COL, T_, 16384, 64, 128),
              ("s", COL, N_, 6000, 3000, 1), ("d", ROW, T_, 1500, 5000, 1),
              ("d", COL, N_, 3000, 700, 1), ("s", COL, N_, 4096, 256, 16)]


@pytest.mark.parametrize("case", BAND_CASES, ids=lambda c: "-".join(map(str, c)))
def test_gemv_band_shapes(hc, handle, orc, case):
    p, order, trans, M, N, b = case
    api = hc.api
    dt = np.float32 if p == "s" else np.float64
    gen = orc.uniform_f32 if p == "s" else orc.uniform_f64
    rows = M if order == COL else N
    cols = N if order == COL else M
    lenx, leny = (N, M) if trans == N_ else (M, N)
    eps = np.finfo(dt).eps
    if b > 1:
        A = gen(M * N * b, 11)
        x, y = gen(lenx * b, 12), gen(leny * b, 13)
        dA, dx, dy = to_dev(A), to_dev(x), to_dev(y)
        st = getattr(api, f"hcblas{p.upper()}gemvBatched")(
            handle, trans, M, N, 0.5, dA, M, dx, 1, 2.0, dy, 1, b)
        assert st == 0
        got = to_host(dy)
        for i in range(b):
            ref = orc.gemv_exact(COL, trans, M, N, 0.5, A[i * M * N:], M,
                                 x[i * lenx:], 1, 2.0, y[i * leny:], 1)
            tol = 2 * (lenx + 2) * eps * (0.5 * lenx + 3)
            assert np.max(np.abs(got[i * leny:(i + 1) * leny] - ref)) <= tol
        return
    lda = rows + 4
    A = gen(lda * cols, M + N)
    for incx, incy, alpha, beta in ((1, 1, 1.0, 0.0), (2, -1, 1.5, 0.5)):
        x = gen(lenx * incx, 3)
        y = gen(leny * abs(incy), 4)
        yy = y.copy()
        if beta == 0.0:
            yy[:] = np.nan  # must not propagate
        dA, dx, dy = to_dev(A), to_dev(x), to_dev(yy)
        st = getattr(api, f"hcblas_{p}gemv")(
            handle, order, trans, M, N, alpha, dA, 0, lda, dx, 0, incx, beta,
            dy, 0, incy)
        assert st == 0
        got = to_host(dy)
        ref = orc.gemv_exact(order, trans, M, N, alpha, A, lda, x, incx, beta,
                             yy, incy)
        idx = (np.arange(leny) * incy if incy > 0 else
               (leny - 1 - np.arange(leny)) * (-incy))
        tol = 2 * (lenx + 2) * eps * (abs(alpha) * lenx + abs(beta) + 1)
        assert np.all(np.isfinite(got[idx]))
        assert np.max(np.abs(got[idx] - ref)) <= tol, (incx, incy)


@pytest.mark.parametrize("p", ["s", "d"])
def test_batched(hc, handle, orc, p):
    api = hc.api
    gen = orc.uniform_f32 if p == "s" else orc.uniform_f64
    P = p.upper()
    M, N, b = 45, 61, 9
    A = gen(M * N * b, 1)
    for trans in (N_, T_):
        lenx, leny = (N, M) if trans == N_ else (M, N)
        x, y = gen(lenx * b, 2), gen(leny * b, 3)
        dy = to_dev(y)
        st = getattr(api, f"hcblas{P}gemvBatched")(handle, trans, M, N, 0.5,
                                                   to_dev(A), M, to_dev(x), 1,
                                                   2.0, dy, 1, b)
        assert st == 0
        got = to_host(dy)
        for i in range(b):
            ref = orc.gemv_exact(COL, trans, M, N, 0.5, A[i * M * N:], M,
                                 x[i * lenx:], 1, 2.0, y[i * leny:], 1)
            tol = 4 * lenx * np.finfo(A.dtype).eps * (lenx + 4)
            assert np.max(np.abs(got[i * leny:(i + 1) * leny] - ref)) <= tol
    x, y = gen(M * b, 4), gen(N * b, 5)
    dA = to_dev(A)
    st = getattr(api, f"hcblas{P}gerBatched")(handle, M, N, 1.5, to_dev(x), 1,
                                              to_dev(y), 1, dA, M, b)
    assert st == 0
    got = to_host(dA)
    for i in range(b):
        exp = orc.ger(COL, M, N, 1.5, x[i * M:(i + 1) * M], 1,
                      y[i * N:(i + 1) * N], 1, A[i * M * N:(i + 1) * M * N], M)
        assert bitwise_equal(got[i * M * N:(i + 1) * M * N], exp)


@pytest.mark.parametrize("p", ["s", "d"])
def test_gemv_full_size_unit_vector(hc, handle, torch_cuda, p):
    """M = N = 16384: with x a unit vector, y = alpha*column (N) or
    alpha*row (T) of A exactly -- an exact, size-independent property."""
    torch = torch_cuda
    api = hc.api
    dt = torch.float32 if p == "s" else torch.float64
    n = 16384
    g = torch.Generator(device="cuda").manual_seed(1)
    A = torch.rand(n * n, generator=g, device="cuda", dtype=dt) * 2 - 1
    j = 12345
    x = torch.zeros(n, device="cuda", dtype=dt)
    x[j] = 1.0
    y = torch.full((n,), float("nan"), device="cuda", dtype=dt)
    fn = getattr(api, f"hcblas{p.upper()}gemv")
    assert fn(handle, 0, n, n, 1.5, A, n, x, 1, 0.0, y, 1) == 0
    torch.cuda.synchronize()
    Am = A.view(n, n)  # Am[c, r] = A(r, c) column-major
    assert torch.equal(y, 1.5 * Am[j, :])
    assert fn(handle, 1, n, n, 1.5, A, n, x, 1, 0.0, y, 1) == 0
    torch.cuda.synchronize()
    assert torch.equal(y, 1.5 * Am[:, j])
