"""Pins the CPU oracle (oracle/refblas.c) before it is trusted as the checker.

Each oracle function is checked against at least one source that is not the
oracle itself:
  - golden fixtures: bundled CBLAS (OpenBLAS) outputs on the reference's own
    integer-valued unit-test recipes (test/unit/hcblas_*.cpp, glibc rand_r
    seed 100) -- exact, so compared bit for bit;
  - real-valued CBLAS outputs (tolerance) and, for axpy/ger, bitwise (the
    bundled OpenBLAS uses the same contraction);
  - exact rational arithmetic on tiny problems;
  - closed forms from the reference tests (C = 2K+2, C = 2K+3);
  - a mutation check: a transposed operand or a dropped term changes the
    result against the same golden, so the goldens discriminate.
"""
from fractions import Fraction

import numpy as np
import pytest

ROW, COL = 0, 1
N_, T_, C_ = 0, 1, 2


def _gemm_case(g, name, i, M, N, K, orc, alpha, beta, ta=None, tb=None):
    o, ta0, tb0, lda, ldb, ldc = (int(v) for v in g[f"{name}_case{i}"])
    ta = ta0 if ta is None else ta
    tb = tb0 if tb is None else tb
    return orc.gemm(o, ta, tb, M, N, K, alpha, g[f"{name}_A"], lda,
                    g[f"{name}_B"], ldb, beta, g[f"{name}_C"], ldc)


@pytest.mark.parametrize("name", ["s", "d"])
def test_gemm_reference_recipe_bit_exact(golden, orc, name):
    g = golden("gemm_ref_int.npz")
    M, N, K = (int(v) for v in g["dims"])
    for i in range(8):
        got = _gemm_case(g, name, i, M, N, K, orc, 1.0, 1.0)
        assert np.array_equal(got, g[f"{name}_expect{i}"]), f"case {i}"


@pytest.mark.parametrize("name", ["c", "z"])
def test_complex_gemm_reference_recipe_bit_exact(golden, orc, name):
    g = golden("gemm_ref_int.npz")
    M, N, K = (int(v) for v in g["dims"])
    for i in range(11):
        got = _gemm_case(g, name, i, M, N, K, orc, 1 + 1j, 1 + 1j)
        assert np.array_equal(got, g[f"{name}_expect{i}"]), f"case {i}"


def test_gemm_golden_discriminates_mistakes(golden, orc):
    """A transposed operand or a swapped order must fail the golden."""
    g = golden("gemm_ref_int.npz")
    M, N, K = (int(v) for v in g["dims"])
    good = g["s_expect0"]
    # Col-major NN case read as if B were transposed (ldb=K still valid for
    # the K x N buffer only as N x K when N <= K: 9 <= 19).
    bad = _gemm_case(g, "s", 0, M, N, K, orc, 1.0, 1.0, tb=T_)
    assert not np.array_equal(bad, good)
    # dropped beta*C term
    bad2 = _gemm_case(g, "s", 0, M, N, K, orc, 1.0, 0.0)
    assert not np.array_equal(bad2, good)
    # complex: conjugation must matter
    gc = g["c_expect8"]
    badc = _gemm_case(g, "c", 8, M, N, K, orc, 1 + 1j, 1 + 1j, ta=T_)
    assert not np.array_equal(badc, gc)


@pytest.mark.parametrize("name", ["s", "d"])
def test_gemv_ger_reference_recipe_bit_exact(golden, orc, name):
    g = golden("level2_ref_int.npz")
    M, N = (int(v) for v in g["dims"])
    for trans in (N_, T_):
        for order in (COL, ROW):
            key = f"{name}gemv_t{trans}_o{order}"
            lda = M if order == COL else N
            got = orc.gemv(order, trans, M, N, 1.0, g[key + "_A"], lda,
                           g[key + "_x"], 1, 1.0, g[key + "_y"], 1)
            assert np.array_equal(got, g[key + "_expect"]), key
    for order in (COL, ROW):
        lda = M if order == COL else N
        for alpha in (1, 3):
            key = f"{name}ger_o{order}_a{alpha}"
            got = orc.ger(order, M, N, float(alpha), g[key + "_x"], 1,
                          g[key + "_y"], 1, g[key + "_A"], lda)
            assert np.array_equal(got, g[key + "_expect"]), key


@pytest.mark.parametrize("name", ["s", "d"])
def test_level1_reference_recipe_bit_exact(golden, orc, name):
    g = golden("level1_ref_int.npz")
    for alpha in (1, 3):
        got = orc.axpy(float(alpha), g[f"{name}axpy_x"], g[f"{name}axpy_y"])
        assert np.array_equal(got, g[f"{name}axpy_a{alpha}"])
    for alpha in (1, 3, 0):
        got = orc.scal(float(alpha), g[f"{name}scal_x"])
        assert np.array_equal(got, g[f"{name}scal_a{alpha}"])
    got = orc.copy(g[f"{name}copy_x"], g[f"{name}copy_y"])
    assert np.array_equal(got, g[f"{name}copy_x"])
    for n in (23, 1000):
        assert orc.dot(g[f"{name}dot{n}_x"], g[f"{name}dot{n}_y"]) == float(
            g[f"{name}dot{n}"])
    assert orc.asum(g[f"{name}asum_x"]) == float(g[f"{name}asum"])
    xb = g[f"{name}asumb_x"]
    tot = sum(orc.asum(xb[i * 23:(i + 1) * 23]) for i in range(128))
    assert tot == float(g[f"{name}asumb"])


@pytest.mark.parametrize("name", ["s", "d", "c", "z"])
def test_gemm_real_valued_matches_cblas(golden, orc, name):
    g = golden("real_valued.npz")
    M, N, K = (int(v) for v in g["dims"])
    a, b = g[f"{name}_ab"]
    if name in "sd":
        a, b = a.real, b.real
    eps = np.finfo(g[f"{name}_C"].dtype).eps
    for i in range(8):
        got = _gemm_case(g, name, i, M, N, K, orc, a, b)
        exp = g[f"{name}_expect{i}"]
        scale = np.abs(exp).max()
        assert np.max(np.abs(got - exp)) <= 4 * K * eps * scale, f"case {i}"


@pytest.mark.parametrize("name", ["s", "d"])
def test_axpy_ger_bitwise_vs_openblas(golden, orc, name):
    g = golden("real_valued.npz")
    got = orc.axpy(1.37, g[f"{name}axpy_x"], g[f"{name}axpy_y"])
    assert np.array_equal(got, g[f"{name}axpy"])
    m, n = (int(v) for v in g[f"{name}ger_dims"])
    for order in (COL, ROW):
        lda = m if order == COL else n
        got = orc.ger(order, m, n, 1.37, g[f"{name}ger_x"], 1,
                      g[f"{name}ger_y"], 1, g[f"{name}ger_A"], lda)
        assert np.array_equal(got, g[f"{name}ger_o{order}"]), order


@pytest.mark.parametrize("name", ["s", "d"])
def test_gemv_real_valued_vs_cblas_and_exact(golden, orc, name):
    g = golden("real_valued.npz")
    m, n = (int(v) for v in g[f"{name}gemv_dims"])
    A, x, y = g[f"{name}gemv_A"], g[f"{name}gemv_x"], g[f"{name}gemv_y"]
    eps = np.finfo(A.dtype).eps
    for trans in (N_, T_):
        lenx, leny = (n, m) if trans == N_ else (m, n)
        for order in (COL, ROW):
            lda = m if order == COL else n
            got = orc.gemv(order, trans, m, n, 1.5, A, lda, x[:lenx], 1, 0.5,
                           y[:leny], 1)
            exp = g[f"{name}gemv_t{trans}_o{order}"]
            ref = orc.gemv_exact(order, trans, m, n, 1.5, A, lda, x[:lenx], 1,
                                 0.5, y[:leny], 1)
            tol = 4 * lenx * eps * (1.5 * lenx + 1)
            assert np.max(np.abs(got - exp)) <= tol
            assert np.max(np.abs(got - ref)) <= tol
            assert np.max(np.abs(exp - ref)) <= tol


def test_exact_rational_references(golden, orc):
    g = golden("exact_rational.npz")
    M, N, K = (int(v) for v in g["gemm_dims"])
    a, b = (float(v) for v in g["gemm_ab"])
    ref = orc.sgemm_ref64(COL, N_, N_, M, N, K, a, g["gemm_A"], M, g["gemm_B"],
                          K, b, g["gemm_C"], M)
    assert np.array_equal(ref, g["gemm_exact"])
    rows = np.repeat(np.arange(M), N)
    cols = np.tile(np.arange(N), M)
    dd = orc.gemm_sample_exact(COL, N_, N_, M, N, K, a, g["gemm_A"], M,
                               g["gemm_B"], K, b, g["gemm_C"], M, rows, cols)
    assert np.array_equal(dd, g["gemm_exact"][rows + cols * M])
    al = float(g["axpy_alpha"][0])
    got = orc.axpy(al, g["axpy_x"], g["axpy_y"])
    assert np.array_equal(got, g["axpy_fma"])  # exactly one rounding
    m, n = (int(v) for v in g["ger_dims"])
    got = orc.ger(COL, m, n, al, g["ger_x"], 1, g["ger_y"], 1, g["ger_A"], m)
    assert np.array_equal(got, g["ger_exact"])
    assert orc.dot_exact(g["dot_x"], g["dot_y"]) == float(g["dot_exact"])


def test_double_double_oracle_against_rationals(orc):
    rng = np.random.default_rng(3)
    M, N, K = 6, 5, 33
    A = rng.standard_normal(M * K) * 2.0 ** rng.integers(-30, 30, M * K)
    B = rng.standard_normal(K * N) * 2.0 ** rng.integers(-30, 30, K * N)
    C = rng.standard_normal(M * N)
    rows = np.repeat(np.arange(M), N)
    cols = np.tile(np.arange(N), M)
    dd = orc.gemm_sample_exact(COL, N_, N_, M, N, K, 0.5, A, M, B, K, 2.0, C,
                               M, rows, cols)
    for s in range(rows.size):
        i, j = rows[s], cols[s]
        q = Fraction(0.5) * sum(Fraction(A[i + l * M]) * Fraction(B[l + j * K])
                                for l in range(K)) + 2 * Fraction(C[i + j * M])
        exact = float(q)
        assert abs(dd[s] - exact) <= 2 * np.spacing(abs(exact)), s
    # complex double-double
    Az = (rng.standard_normal(M * K) + 1j * rng.standard_normal(M * K))
    Bz = (rng.standard_normal(K * N) + 1j * rng.standard_normal(K * N))
    Cz = np.zeros(M * N, np.complex128)
    ddz = orc.gemm_sample_exact(COL, C_, N_, K, N, M, 1 + 0j, Az, M, Bz[:M * N],
                                M, 0j, Cz, K, np.arange(3), np.arange(3))
    for s in range(3):
        q_re = sum(Fraction(Az[l + s * M].real) * Fraction(Bz[l + s * M].real) +
                   Fraction(Az[l + s * M].imag) * Fraction(Bz[l + s * M].imag)
                   for l in range(M))
        q_im = sum(Fraction(Az[l + s * M].real) * Fraction(Bz[l + s * M].imag) -
                   Fraction(Az[l + s * M].imag) * Fraction(Bz[l + s * M].real)
                   for l in range(M))
        assert abs(ddz[s].real - float(q_re)) <= 2 * np.spacing(abs(float(q_re)))
        assert abs(ddz[s].imag - float(q_im)) <= 2 * np.spacing(abs(float(q_im)))


@pytest.mark.parametrize("order", [ROW, COL])
def test_gemm_closed_forms(orc, order):
    # test/unit-api/test-blas.cpp:1861-1920: A=B=C=1, alpha=beta=2, 123x78x23
    M, N, K = 123, 78, 23
    for dt in (np.float32, np.float64, np.complex64, np.complex128):
        A = np.ones(M * K, dt)
        B = np.ones(K * N, dt)
        C = np.ones(M * N, dt)
        lda, ldb, ldc = (M, K, M) if order == COL else (K, N, N)
        got = orc.gemm(order, N_, N_, M, N, K, 2, A, lda, B, ldb, 2, C, ldc)
        assert np.all(got == 2 * K + 2)
        # test-hipblas.cpp:1815-1853: A=1, B=2, C=3, alpha=beta=1 -> 2K+3
        got = orc.gemm(order, T_, T_, M, N, K, 1, A, K if order == COL else M,
                       2 * B, N if order == COL else K, 1, 3 * C, ldc)
        assert np.all(got == 2 * K + 3)
    H = orc.hgemm(order, N_, N_, M, N, K, 2.0, np.ones(M * K, np.float16), lda,
                  np.ones(K * N, np.float16), ldb, 2.0,
                  np.ones(M * N, np.float16), ldc)
    assert np.all(H == 2 * K + 2)


def test_gemm_special_cases(orc):
    M, N, K = 7, 5, 3
    A = orc.uniform_f32(M * K, 1)
    B = orc.uniform_f32(K * N, 2)
    C = np.full(M * N, np.nan, np.float32)
    # beta == 0: C is overwritten, NaN does not propagate
    got = orc.gemm(COL, N_, N_, M, N, K, 1.0, A, M, B, K, 0.0, C, M)
    assert np.all(np.isfinite(got))
    # alpha == 0: C = beta*C, A and B not read
    C2 = orc.uniform_f32(M * N, 3)
    got = orc.gemm(COL, N_, N_, M, N, K, 0.0, A, M, B, K, 2.0, C2, M)
    assert np.array_equal(got, 2 * C2)
    # K == 0: C = beta*C
    got = orc.gemm(COL, N_, N_, M, N, 0, 1.0, A, M, B, 1, 0.5, C2, M)
    assert np.array_equal(got, 0.5 * C2)


def test_generator_properties(orc):
    x = orc.uniform_f32(100000, 0)
    assert x.min() >= -1 and x.max() < 1
    assert np.array_equal(x, orc.uniform_f32(100000, 0))
    assert np.array_equal(x[500:600], orc.uniform_f32(100, 0, start=500))
    # on the 2^-23 grid (exact in fp32)
    assert np.all((x.astype(np.float64) * 2 ** 23) % 1 == 0)
    d = orc.uniform_f64(10000, 5)
    assert np.all((d * 2 ** 52) % 1 == 0)
    w = orc.wide_f64(100000, 9)
    e = np.log2(np.abs(w))
    assert e.min() >= -20 and e.max() <= 20 and np.std(e) > 10
    # continuous exponents (integer exponents would make splits exact)
    assert np.mean(e % 1 != 0) > 0.99


def test_scal_alpha_zero_keeps_nan(orc):
    x = np.array([np.nan, np.inf, 1.0, -2.0], np.float32)
    got = orc.scal(0.0, x)
    assert np.isnan(got[0]) and np.isnan(got[1])
    assert got[2] == 0 and np.signbit(got[3])


def test_complex_scal_definition(orc):
    rng = np.random.default_rng(0)
    x = (rng.uniform(-1, 1, 50) + 1j * rng.uniform(-1, 1, 50)).astype(np.complex64)
    a = complex(0.5, -1.25)
    got = orc.scal(a, x)
    ref = (np.complex128(a) * x.astype(np.complex128))
    assert np.max(np.abs(got - ref)) < 1e-6
    got2 = orc.scal(1.5, x)
    assert np.array_equal(got2, (np.float32(1.5) * x).astype(np.complex64))
