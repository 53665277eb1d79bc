"""hcblas_gemv_noalloc (the allocation-free GEMV) against hcblas_sgemv /
hcblas_dgemv through the C ABI: bit-identical results on two calls in a row,
the second on a smaller problem (the fixed scratch keeps the first call's
partials, which must not leak into the second), for both types, both
orders and both transposes, including the 16384^2 shapes whose split
partials fill a quarter of the scratch. Which problems fit the scratch
(every single GEMV up to 16 GiB of A) is checked on the CPU from the launch
plan itself (tests/test_gemv_plan.py).
"""
import numpy as np
import pytest

from gpu_util import bitwise_equal, to_dev, to_host

pytestmark = pytest.mark.gpu
ROW, COL = 0, 1
N_, T_ = 0, 1
R_32F, R_64F = 0, 1


def _pair(hc, handle, orc, dt, order, trans, M, N, seed):
    api = hc.api
    gen = orc.uniform_f32 if dt == np.float32 else orc.uniform_f64
    lda = M if order == COL else N
    A = gen(M * N, seed)
    lenx, leny = (N, M) if trans == N_ else (M, N)
    x = gen(lenx, seed + 1)
    y = gen(leny, seed + 2)
    dA, dx = to_dev(A), to_dev(x)
    ref_y, got_y = to_dev(y), to_dev(y)
    plain = api.hcblas_sgemv if dt == np.float32 else api.hcblas_dgemv
    assert plain(handle, order, trans, M, N, 1.5, dA, 0, lda, dx, 0, 1, 0.5,
                 ref_y, 0, 1) == 0
    code = R_32F if dt == np.float32 else R_64F
    st = api.hcblas_gemv_noalloc(handle, code, order, trans, M, N, 1.5, dA, 0,
                                 lda, dx, 0, 1, 0.5, got_y, 0, 1)
    assert st == 0
    return to_host(ref_y), to_host(got_y)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("order", [COL, ROW])
@pytest.mark.parametrize("trans", [N_, T_])
def test_noalloc_matches_gemv_two_calls(hc, handle, orc, dt, order, trans):
    # first a large problem (many split partials), then a smaller one
    for (M, N, seed) in ((4096, 8192, 1), (1000, 3000, 7)):
        ref, got = _pair(hc, handle, orc, dt, order, trans, M, N, seed)
        assert bitwise_equal(ref, got), (M, N)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_noalloc_bandwidth_shapes(hc, handle, orc, dt):
    """The 16384^2 GEMV of the bandwidth set, then a small problem."""
    for trans in (N_, T_):
        ref, got = _pair(hc, handle, orc, dt, COL, trans, 16384, 16384, 11)
        assert bitwise_equal(ref, got)
    ref, got = _pair(hc, handle, orc, dt, COL, N_, 300, 257, 5)
    assert bitwise_equal(ref, got)


def test_noalloc_argument_checks(hc, handle):
    api = hc.api
    st = api.hcblas_gemv_noalloc(handle, 2, COL, N_, 4, 4, 1.0, 0, 0, 4, 0, 0,
                                 1, 0.0, 0, 0, 1)
    assert st == 3  # complex dtype is not a GEMV type
    st = api.hcblas_gemv_noalloc(None, R_32F, COL, N_, 4, 4, 1.0, 0, 0, 4, 0,
                                 0, 1, 0.0, 0, 0, 1)
    assert st == 1
