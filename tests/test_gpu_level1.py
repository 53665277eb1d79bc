"""Level-1 parity on the GPU, through the C ABI, against the oracle.

Reference tests mirrored: test/unit/hcblas_{s,d}{axpy,scal,copy,dot,asum}.cpp
and test/unit-api/test-blas.cpp (integer recipes, exact comparison), plus
seeded real-valued data, ragged sizes, misaligned offsets, strides and
batches. copy/scal/axpy are bit-exact; dot/asum are within n*eps*sum|x*y|
of the double-double value (and exact on integer data).
"""
import numpy as np
import pytest

from gpu_util import bitwise_equal, empty, to_dev, to_host

pytestmark = pytest.mark.gpu

SIZES = [1, 3, 17, 255, 1023, 4097, (1 << 20) + 3]


def _api(hc):
    return hc.api


@pytest.mark.parametrize("p", ["s", "d"])
def test_reference_recipes(hc, handle, golden, p):
    api = _api(hc)
    g = golden("level1_ref_int.npz")
    dt = np.float32 if p == "s" else np.float64
    axpy = getattr(api, f"hcblas{p.upper()}axpy")
    for alpha in (1, 3):
        x, y = to_dev(g[f"{p}axpy_x"]), to_dev(g[f"{p}axpy_y"])
        assert axpy(handle, x.numel(), float(alpha), x, 1, y, 1) == 0
        assert bitwise_equal(to_host(y), g[f"{p}axpy_a{alpha}"])
    scal = getattr(api, f"hcblas{p.upper()}scal")
    for alpha in (1, 3, 0):
        x = to_dev(g[f"{p}scal_x"])
        assert scal(handle, x.numel(), float(alpha), x, 1) == 0
        assert bitwise_equal(to_host(x), g[f"{p}scal_a{alpha}"])
    copy = getattr(api, f"hcblas{p.upper()}copy")
    x, y = to_dev(g[f"{p}copy_x"]), to_dev(g[f"{p}copy_y"])
    assert copy(handle, x.numel(), x, 1, y, 1) == 0
    assert bitwise_equal(to_host(y), g[f"{p}copy_x"])
    dot = getattr(api, f"hcblas{p.upper()}dot")
    for n in (23, 1000):
        x, y = to_dev(g[f"{p}dot{n}_x"]), to_dev(g[f"{p}dot{n}_y"])
        st, r = dot(handle, n, x, 1, y, 1, None)
        assert st == 0 and r == float(g[f"{p}dot{n}"])
    asum = getattr(api, f"hcblas{p.upper()}asum")
    x = to_dev(g[f"{p}asum_x"])
    st, r = asum(handle, x.numel(), x, 1, None)
    assert st == 0 and r == float(g[f"{p}asum"])
    # batched asum returns one scalar over the whole batch (test-blas.cpp:116)
    asumb = getattr(api, f"hcblas{p.upper()}asumBatched")
    x = to_dev(g[f"{p}asumb_x"])
    st, r = asumb(handle, 23, x, 1, None, 128)
    assert st == 0 and r == float(g[f"{p}asumb"])
    assert dt


@pytest.mark.parametrize("p", ["s", "d"])
@pytest.mark.parametrize("n", SIZES)
def test_elementwise_bit_exact(hc, handle, orc, p, n):
    api = _api(hc)
    dt = np.float32 if p == "s" else np.float64
    gen = orc.uniform_f32 if p == "s" else orc.uniform_f64
    P = p.upper()
    for off in (0, 1):  # off = 1 breaks 16-byte alignment -> scalar path
        x = gen(n + off, 11 + n)
        y = gen(n + off, 12 + n)
        dx, dy = to_dev(x), to_dev(y)
        xo, yo = dx[off:], dy[off:]
        alpha = 1.37
        assert getattr(api, f"hcblas{P}axpy")(handle, n, alpha, xo, 1, yo, 1) == 0
        exp = orc.axpy(dt(alpha), x[off:], y[off:])
        assert bitwise_equal(to_host(dy)[off:], exp)
        assert getattr(api, f"hcblas{P}scal")(handle, n, -0.625, xo, 1) == 0
        exp = orc.scal(dt(-0.625), x[off:])
        assert bitwise_equal(to_host(dx)[off:], exp)
        dz = to_dev(np.zeros(n + off, dt))
        assert getattr(api, f"hcblas{P}copy")(handle, n, xo, 1, dz[off:], 1) == 0
        assert bitwise_equal(to_host(dz)[off:], exp)


@pytest.mark.parametrize("p", ["s", "d"])
@pytest.mark.parametrize("incx,incy", [(2, 1), (1, 3), (-1, 2), (-2, -1)])
def test_strided(hc, handle, orc, p, incx, incy):
    api = _api(hc)
    dt = np.float32 if p == "s" else np.float64
    gen = orc.uniform_f32 if p == "s" else orc.uniform_f64
    P = p.upper()
    n = 1001
    x = gen(n * abs(incx), 5)
    y = gen(n * abs(incy), 6)
    dx, dy = to_dev(x), to_dev(y)
    assert getattr(api, f"hcblas{P}axpy")(handle, n, 0.5, dx, incx, dy, incy) == 0
    assert bitwise_equal(to_host(dy), orc.axpy(dt(0.5), x, y, incx, incy, n))
    dy = to_dev(y)
    assert getattr(api, f"hcblas{P}copy")(handle, n, dx, incx, dy, incy) == 0
    assert bitwise_equal(to_host(dy), orc.copy(x, y, incx, incy, n))
    st, r = getattr(api, f"hcblas{P}dot")(handle, n, dx, incx, to_dev(y), incy,
                                          None)
    ex = orc.dot_exact(x, y, incx, incy, n)
    bound = n * np.finfo(dt).eps * float(np.sum(np.abs(x) * 1.0)) * 1.0
    assert st == 0 and abs(r - ex) <= bound
    if incx > 0:
        dx = to_dev(x)
        assert getattr(api, f"hcblas{P}scal")(handle, n, 3.0, dx, incx) == 0
        assert bitwise_equal(to_host(dx), orc.scal(dt(3.0), x, incx, n))
        st, r = getattr(api, f"hcblas{P}asum")(handle, n, to_dev(x), incx, None)
        assert st == 0
        assert abs(r - orc.asum_exact(x, incx, n)) <= n * np.finfo(dt).eps * r
    else:
        # BLAS: non-positive increment -> scal no-op, asum = 0
        dx = to_dev(x)
        assert getattr(api, f"hcblas{P}scal")(handle, n, 3.0, dx, incx) == 0
        assert bitwise_equal(to_host(dx), x)
        st, r = getattr(api, f"hcblas{P}asum")(handle, n, dx, incx, None)
        assert st == 0 and r == 0.0


@pytest.mark.parametrize("p", ["s", "d"])
@pytest.mark.parametrize("n", SIZES + [(1 << 22) + 5])
def test_dot_asum_accuracy(hc, handle, orc, p, n):
    api = _api(hc)
    dt = np.float32 if p == "s" else np.float64
    gen = orc.uniform_f32 if p == "s" else orc.uniform_f64
    P = p.upper()
    x, y = gen(n, 101), gen(n, 102)
    dx, dy = to_dev(x), to_dev(y)
    eps = np.finfo(dt).eps
    st, r = getattr(api, f"hcblas{P}dot")(handle, n, dx, 1, dy, 1, None)
    ex = orc.dot_exact(x, y)
    scale = float(np.sum(np.abs(x.astype(np.float64) * y)))
    assert st == 0 and abs(r - ex) <= n * eps * scale + eps * abs(ex)
    # deterministic: identical bits on a repeat
    st2, r2 = getattr(api, f"hcblas{P}dot")(handle, n, dx, 1, dy, 1, None)
    assert st2 == 0 and np.array(r, dt).tobytes() == np.array(r2, dt).tobytes()
    st, r = getattr(api, f"hcblas{P}asum")(handle, n, dx, 1, None)
    ex = orc.asum_exact(x)
    assert st == 0 and abs(r - ex) <= n * eps * ex


@pytest.mark.parametrize("p", ["s", "d"])
def test_batched(hc, handle, orc, p):
    api = _api(hc)
    dt = np.float32 if p == "s" else np.float64
    gen = orc.uniform_f32 if p == "s" else orc.uniform_f64
    P = p.upper()
    n, b = 37, 19
    x, y = gen(n * b, 1), gen(n * b, 2)
    dx, dy = to_dev(x), to_dev(y)
    if p == "s":
        assert api.hcblasSaxpyBatched(handle, n, 2.5, dx, 1, dy, 1, b) == 0
    else:
        assert api.hcblasDaxpyBatched(handle, n, 2.5, dx, 1, dy, 1, b) == 0
    assert bitwise_equal(to_host(dy), orc.axpy(dt(2.5), x, y))
    dz = to_dev(np.zeros(n * b, dt))
    assert getattr(api, f"hcblas{P}copyBatched")(handle, n, dx, 1, dz, 1, b) == 0
    assert bitwise_equal(to_host(dz), x)
    assert getattr(api, f"hcblas{P}scalBatched")(handle, n, 0.75, dz, 1, b) == 0
    assert bitwise_equal(to_host(dz), orc.scal(dt(0.75), x))
    st, r = getattr(api, f"hcblas{P}dotBatched")(handle, n, dx, 1, to_dev(y), 1,
                                                 None, b)
    ex = sum(orc.dot_exact(x[i * n:(i + 1) * n], y[i * n:(i + 1) * n])
             for i in range(b))
    assert st == 0 and abs(r - ex) <= n * b * np.finfo(dt).eps * n * b


@pytest.mark.parametrize("p", ["c", "z"])
def test_complex_scal_bit_exact(hc, handle, orc, p):
    api = _api(hc)
    dt = np.complex64 if p == "c" else np.complex128
    gen = orc.uniform_f32 if p == "c" else orc.uniform_f64
    n = 3001
    v = gen(2 * n, 7)
    x = v[0::2] + 1j * v[1::2]
    x = x.astype(dt)
    a = complex(0.5, -1.25)
    dx = to_dev(x)
    assert getattr(api, f"hcblas{p.upper()}scal")(handle, n, a, dx, 1) == 0
    assert bitwise_equal(to_host(dx), orc.scal(a, x))
    dx = to_dev(x)
    real = getattr(api, "hcblasCsscal" if p == "c" else "hcblasZdscal")
    assert real(handle, n, 1.5, dx, 1) == 0
    assert bitwise_equal(to_host(dx), orc.scal(1.5, x))
    # batched: batch stride n (hcblas.cpp:376), two batches
    xx = np.concatenate([x, x[::-1]])
    dx = to_dev(xx)
    bat = getattr(api, f"hcblas{p.upper()}scalBatched")
    assert bat(handle, n, a, dx, 1, 2) == 0
    assert bitwise_equal(to_host(dx), orc.scal(a, xx))
    # strided inside one vector
    dx = to_dev(x)
    fn = getattr(api, f"hcblas{p.upper()}scal")
    assert fn(handle, n // 2, a, dx, 2) == 0
    assert bitwise_equal(to_host(dx), orc.scal(a, x, 2, n // 2))


def test_full_size_properties(hc, handle, torch_cuda):
    """n = 2^28 fp32 (the bandwidth configuration): exact checks through
    properties that do not need a CPU pass over 1 GiB."""
    torch = torch_cuda
    api = hc.api
    n = 1 << 28
    g = torch.Generator(device="cuda").manual_seed(0)
    x = torch.randint(-8, 8, (n,), generator=g, device="cuda").float()
    y = torch.randint(-8, 8, (n,), generator=g, device="cuda").float()
    z = torch.empty_like(x)
    assert api.hcblasScopy(handle, n, x, 1, z, 1) == 0
    torch.cuda.synchronize()
    assert torch.equal(z, x)
    # integer data: dot and asum are exact (|sum| < 2^24 per partial is not
    # needed: FP64 cross-workgroup sums; per-thread fp32 partials stay exact
    # because every partial is an integer below 2^24)
    st, r = api.hcblasSdot(handle, n, x, 1, y, 1, None)
    exact = float((x.double() * y.double()).sum())
    # the result is an fp32 scalar: exact value rounded once
    assert st == 0 and r == float(np.float32(exact))
    st, r = api.hcblasSasum(handle, n, x, 1, None)
    assert st == 0 and r == float(np.float32(x.double().abs().sum().item()))
    assert api.hcblasSaxpy(handle, n, 2.0, x, 1, z, 1) == 0
    torch.cuda.synchronize()
    assert torch.equal(z, 3 * x)
    assert api.hcblasSscal(handle, n, 0.5, z, 1) == 0
    torch.cuda.synchronize()
    assert torch.equal(z, 1.5 * x)
