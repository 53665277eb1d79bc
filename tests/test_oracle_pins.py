"""Pins the oracle's accuracy references (fp64 dot, asum, sampled complex
GEMM, HGEMM) with exact arithmetic, and shows each pin discriminates.

Each reference is compared with a value the mathematics fixes:
  - ``dot_exact`` (fp64, double-double) and ``asum_exact``: exact rational
    sums (``fractions.Fraction``) of wide-exponent vectors, positive and
    negative strides, and an ill-conditioned dot where plain fp64
    accumulation is off by many ulps;
  - ``gemm_sample_exact`` on complex64 / complex128: the exact complex
    rational product for every order x transa x transb (N, T, C), padded
    leading dimensions, complex alpha and beta;
  - ``hgemm``: ``float16(sgemm_ref64(...))`` (the C restatement, exact on
    binary16 grid data) for every order x transa x transb with padded
    lda/ldb/ldc, and a case where accumulating in half would give another
    answer (the reference's HGEMM accumulates in half,
    lib/src/blas/hgemm/hgemm_NN_kernels.cpp:1372; its host reference is
    compiled out, test/unit/hcblas_hgemm.cpp:34-141).

Mutation checks: each test also evaluates a plausible mistake (a dropped
fabs, an ignored stride, a missing conjugation, a flipped sign, a transposed
operand, an ignored leading dimension, half accumulation) and asserts that
the same comparison rejects it, so a pin that passes is not vacuous.
Reference-side anchors: test/unit/hcblas_ddot.cpp:228-260 and
test/unit/hcblas_dasum.cpp (the CBLAS comparisons these references stand in
for at fp64 accuracy).
"""
from fractions import Fraction

import numpy as np
import pytest

ROW, COL = 0, 1
N_, T_, C_ = 0, 1, 2


def _fr(v):
    return Fraction(float(v))


def _strided(vec, n, inc):
    """Elements of a BLAS strided vector in logical order."""
    if inc > 0:
        return [vec[i * inc] for i in range(n)]
    return [vec[(n - 1 - i) * (-inc)] for i in range(n)]


def _close(got, exact, scale_abs):
    """|got - exact| within one rounding of the exact value plus 2^-100 of the
    absolute sum (double-double residue)."""
    err = abs(Fraction(got) - exact)
    return err <= Fraction(np.spacing(abs(float(exact)))) + scale_abs * Fraction(
        1, 2 ** 100)


def _wide(rng, n, emin=-60, emax=60):
    return (rng.uniform(-1, 1, n) * 2.0 ** rng.integers(emin, emax, n))


# ---------------------------------------------------------------- dot (fp64)
@pytest.mark.parametrize("incx,incy", [(1, 1), (3, 1), (1, -2), (-3, 2)])
def test_ddot_dd_exact_rational(orc, incx, incy):
    rng = np.random.default_rng(100 + 7 * incx + incy)
    n = 301
    x = _wide(rng, n * abs(incx))
    y = _wide(rng, n * abs(incy))
    xs, ys = _strided(x, n, incx), _strided(y, n, incy)
    exact = sum(_fr(a) * _fr(b) for a, b in zip(xs, ys))
    absum = sum(abs(_fr(a) * _fr(b)) for a, b in zip(xs, ys))
    got = orc.dot_exact(x, y, incx, incy, n)
    assert _close(got, exact, absum)
    # mutation: the stride of x ignored (contiguous read) is rejected
    if incx != 1:
        bad = orc.dot_exact(np.ascontiguousarray(x[:n]), y, 1, incy, n)
        assert not _close(bad, exact, absum)
    # mutation: a reversed (negative-stride) vector read forwards is rejected
    if incy < 0:
        bad = orc.dot_exact(x, y, incx, -incy, n)
        assert not _close(bad, exact, absum)


def test_ddot_dd_ill_conditioned(orc):
    """Terms that cancel to a tiny remainder: plain fp64 accumulation (the
    netlib ddot order, orc.dot) is wrong by far more than an ulp; the
    double-double reference is correctly rounded."""
    rng = np.random.default_rng(5)
    n = 64
    big = rng.uniform(1, 2, n) * 2.0 ** 50
    x = np.concatenate([big, -big, [3.0, 2.0 ** -30]])
    rng.shuffle(x)
    y = np.ones_like(x)
    y[::3] = 1.0 + 2.0 ** -40  # products carry bits below the big terms' ulp
    exact = sum(_fr(a) * _fr(b) for a, b in zip(x, y))
    absum = sum(abs(_fr(a) * _fr(b)) for a, b in zip(x, y))
    got = orc.dot_exact(x, y)
    assert _close(got, exact, absum)
    assert abs(Fraction(got) - exact) <= Fraction(np.spacing(float(exact)))
    naive = orc.dot(x, y)
    assert not _close(naive, exact, Fraction(0))


def test_sdot_f64_exact_rational(orc):
    rng = np.random.default_rng(2)
    n = 777
    x = _wide(rng, 2 * n, -20, 20).astype(np.float32)
    y = _wide(rng, n, -20, 20).astype(np.float32)
    exact = sum(_fr(a) * _fr(b) for a, b in zip(_strided(x, n, -2), y))
    absum = sum(abs(_fr(a) * _fr(b)) for a, b in zip(_strided(x, n, -2), y))
    assert _close(orc.dot_exact(x, y, -2, 1, n), exact, absum)


# -------------------------------------------------------------------- asum
@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("inc", [1, 2, 5])
def test_asum_exact_rational(orc, dt, inc):
    rng = np.random.default_rng(int(inc) + (0 if dt == np.float32 else 100))
    n = 513
    x = _wide(rng, n * inc, -40, 40).astype(dt)
    xs = _strided(x, n, inc)
    exact = sum(abs(_fr(v)) for v in xs)
    got = orc.asum_exact(x, inc, n)
    assert _close(got, exact, exact)
    # mutation: a dropped fabs (signed sum) is rejected
    signed = orc.dot_exact(x.astype(np.float64), np.ones(x.size), inc, inc, n)
    assert not _close(signed, exact, exact)
    # mutation: the stride ignored is rejected
    if inc != 1:
        assert not _close(orc.asum_exact(x, 1, n), exact, exact)


def test_asum_exact_catches_lost_low_bits(orc):
    """One large value and many values below its ulp: a plain fp64 running
    sum drops them all, the double-double sum keeps them."""
    x = np.concatenate([[2.0 ** 60], np.full(1000, 3.0)])
    exact = Fraction(2 ** 60) + 3000
    assert Fraction(orc.asum_exact(x)) == Fraction(float(exact))
    assert Fraction(orc.asum(x)) != Fraction(float(exact))


def test_asum_exact_nonpositive_inc_is_zero(orc):
    x = np.arange(1, 10, dtype=np.float64)
    assert orc.asum_exact(x, 0, 9) == 0.0
    assert orc.asum_exact(x, -1, 9) == 0.0


# ----------------------------------------------------- complex sampled GEMM
def _cfr(v):
    return (_fr(v.real), _fr(v.imag))


def _cmul(a, b):
    return (a[0] * b[0] - a[1] * b[1], a[0] * b[1] + a[1] * b[0])


def _cadd(a, b):
    return (a[0] + b[0], a[1] + b[1])


def _elem(X, order, ld, r, c):
    return X[r + c * ld] if order == COL else X[r * ld + c]


def _op_elem(X, order, ld, trans, r, c, flip_conj=False):
    """op(X)(r, c) as a rational pair; `flip_conj` plants a mistake (the
    conjugation of OP_C dropped)."""
    if trans == N_:
        v = _elem(X, order, ld, r, c)
    else:
        v = _elem(X, order, ld, c, r)
    re, im = _cfr(v)
    if trans == C_ and not flip_conj:
        im = -im
    return (re, im)


def _complex_exact(order, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C,
                   ldc, i, j, mutate=None):
    acc = (Fraction(0), Fraction(0))
    absacc = Fraction(0)
    for l in range(K):
        a = _op_elem(A, order, lda, ta, i, l, mutate == "noconj")
        b = _op_elem(B, order, ldb, tb, l, j, mutate == "noconj")
        p = _cmul(a, b)
        if mutate == "sign":  # wrong sign on the imaginary x imaginary term
            p = (a[0] * b[0] + a[1] * b[1], p[1])
        acc = _cadd(acc, p)
        absacc += abs(a[0] * b[0]) + abs(a[1] * b[1]) + abs(a[0] * b[1]) + abs(
            a[1] * b[0])
    al = (_fr(alpha.real), _fr(alpha.imag))
    r = _cmul(al, acc)
    if mutate != "nobeta" and (beta.real != 0 or beta.imag != 0):
        r = _cadd(r, _cmul((_fr(beta.real), _fr(beta.imag)),
                           _cfr(_elem(C, order, ldc, i, j))))
    scale = (abs(al[0]) + abs(al[1])) * absacc + 4
    return r, scale


def _cclose(got, exact, scale):
    return (_close(got.real, exact[0], scale) and
            _close(got.imag, exact[1], scale))


@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
@pytest.mark.parametrize("order", [COL, ROW])
def test_complex_dd_sample_exact_rational(orc, dt, order):
    rng = np.random.default_rng(31 if dt == np.complex64 else 32)
    M, N, K = 5, 4, 9
    rdt = np.float32 if dt == np.complex64 else np.float64
    alpha, beta = complex(0.75, -1.5), complex(-0.5, 0.25)
    for ta in (N_, T_, C_):
        for tb in (N_, T_, C_):
            ar, ac = (M, K) if ta == N_ else (K, M)
            br, bc = (K, N) if tb == N_ else (N, K)
            lda = (ar if order == COL else ac) + 3
            ldb = (br if order == COL else bc) + 2
            ldc = (M if order == COL else N) + 1
            na = lda * (ac if order == COL else ar)
            nb = ldb * (bc if order == COL else br)
            nc = ldc * (N if order == COL else M)

            def mk(n):
                return (_wide(rng, n, -12, 12).astype(rdt) +
                        1j * _wide(rng, n, -12, 12).astype(rdt)).astype(dt)
            A, B, C = mk(na), mk(nb), mk(nc)
            rows = np.array([0, M - 1, 2, 3, 1], np.int64)
            cols = np.array([0, N - 1, 1, 3, 2], np.int64)
            got = orc.gemm_sample_exact(order, ta, tb, M, N, K, alpha, A, lda,
                                        B, ldb, beta, C, ldc, rows, cols)
            rejected = {"noconj": False, "sign": False, "nobeta": False}
            for s in range(rows.size):
                i, j = int(rows[s]), int(cols[s])
                ex, sc = _complex_exact(order, ta, tb, M, N, K, alpha, A, lda,
                                        B, ldb, beta, C, ldc, i, j)
                assert _cclose(got[s], ex, sc), (ta, tb, s)
                for mut in rejected:
                    if mut == "noconj" and C_ not in (ta, tb):
                        continue
                    bad, _ = _complex_exact(order, ta, tb, M, N, K, alpha, A,
                                            lda, B, ldb, beta, C, ldc, i, j,
                                            mutate=mut)
                    if not _cclose(got[s], bad, sc):
                        rejected[mut] = True
            for mut, hit in rejected.items():
                if mut == "noconj" and C_ not in (ta, tb):
                    continue
                assert hit, (mut, ta, tb)


# ------------------------------------------------------------------- HGEMM
def _grid_half(rng, n):
    """binary16 values k/256, |k| <= 512: every product is a multiple of
    2^-16 below 4, so any sum of fewer than 2^30 of them is exact in fp64."""
    return (rng.integers(-512, 513, n) / 256.0).astype(np.float16)


def _hgemm_shapes(order, ta, tb, M, N, K, pad):
    ar, ac = (M, K) if ta == N_ else (K, M)
    br, bc = (K, N) if tb == N_ else (N, K)
    lda = (ar if order == COL else ac) + pad
    ldb = (br if order == COL else bc) + pad
    ldc = (M if order == COL else N) + pad
    na = lda * (ac if order == COL else ar)
    nb = ldb * (bc if order == COL else br)
    nc = ldc * (N if order == COL else M)
    return lda, ldb, ldc, na, nb, nc


def _c_view(Cflat, order, M, N, ldc):
    if order == COL:
        idx = np.arange(M)[:, None] + ldc * np.arange(N)[None, :]
    else:
        idx = ldc * np.arange(M)[:, None] + np.arange(N)[None, :]
    return Cflat[idx]


@pytest.mark.parametrize("order", [COL, ROW])
@pytest.mark.parametrize("ta", [N_, T_])
@pytest.mark.parametrize("tb", [N_, T_])
def test_hgemm_matches_c_restatement(orc, order, ta, tb):
    rng = np.random.default_rng(100 + 8 * order + 2 * ta + tb)
    M, N, K = 23, 17, 41
    lda, ldb, ldc, na, nb, nc = _hgemm_shapes(order, ta, tb, M, N, K, 3)
    A, B, C = _grid_half(rng, na), _grid_half(rng, nb), _grid_half(rng, nc)
    alpha, beta = 1.5, -0.75
    got = orc.hgemm(order, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C,
                    ldc)
    ref64 = orc.sgemm_ref64(order, ta, tb, M, N, K, alpha, A.astype(np.float32),
                            lda, B.astype(np.float32), ldb, beta,
                            C.astype(np.float32), ldc)
    ref = ref64.reshape((N, M)).T if order == COL else ref64.reshape((M, N))
    expect = ref.astype(np.float16)
    gv = _c_view(got, order, M, N, ldc)
    assert np.array_equal(gv.view(np.uint16), expect.view(np.uint16))
    # entries outside the M x N view (ld padding) are untouched
    mask = np.ones(nc, bool)
    mask[_c_view(np.arange(nc), order, M, N, ldc).ravel()] = False
    assert np.array_equal(got[mask].view(np.uint16), C[mask].view(np.uint16))
    # mutation: transposed A (ta flipped, same lda) is rejected; the buffer
    # is extended (repeating A) so the flipped read stays in bounds
    bad = orc.hgemm(order, 1 - ta, tb, M, N, K, alpha,
                    np.resize(A, lda * (M + K)), lda, B, ldb, beta, C, ldc)
    assert not np.array_equal(_c_view(bad, order, M, N, ldc), expect)
    # mutation: the padded leading dimension of B ignored is rejected
    ldb_bad = ldb - 3
    bad = orc.hgemm(order, ta, tb, M, N, K, alpha, A, lda, B, ldb_bad, beta,
                    C, ldc)
    assert not np.array_equal(_c_view(bad, order, M, N, ldc), expect)


def test_hgemm_exact_rational_tiny(orc):
    """Arbitrary binary16 values (not on a grid), K = 7: the oracle is the
    exact rational result rounded once to binary16."""
    rng = np.random.default_rng(77)
    M, N, K = 4, 3, 7
    A = (rng.uniform(-1, 1, M * K) * 2.0 ** rng.integers(-6, 4, M * K)).astype(
        np.float16)
    B = (rng.uniform(-1, 1, K * N) * 2.0 ** rng.integers(-6, 4, K * N)).astype(
        np.float16)
    C = rng.uniform(-1, 1, M * N).astype(np.float16)
    got = orc.hgemm(COL, N_, N_, M, N, K, 1.0, A, M, B, K, 0.5, C, M)
    for i in range(M):
        for j in range(N):
            q = sum(_fr(A[i + l * M]) * _fr(B[l + j * K]) for l in range(K))
            q += Fraction(1, 2) * _fr(C[i + j * M])
            # q is a multiple of 2^-36 below 2^12: exact in fp64, so the
            # fp64 -> binary16 conversion is the single correct rounding.
            assert Fraction(float(q)) == q
            assert got[i + j * M] == np.float16(float(q)), (i, j)


def test_hgemm_half_accumulation_differs(orc):
    """A row [2048, 1 x 16] times ones: the exact sum 2064 is a binary16
    value, while a binary16 running sum stays at 2048 (2049 rounds to even).
    The oracle returns 2064; a half-accumulating mutant returns 2048."""
    K = 17
    A = np.concatenate([[2048.0], np.ones(K - 1)]).astype(np.float16)
    B = np.ones(K, np.float16)
    C = np.zeros(1, np.float16)
    got = orc.hgemm(COL, N_, N_, 1, 1, K, 1.0, A, 1, B, K, 0.0, C, 1)
    assert float(got[0]) == 2064.0
    acc = np.float16(0)
    for a, b in zip(A, B):
        acc = np.float16(acc + np.float16(a * b))
    assert float(acc) == 2048.0 and float(acc) != float(got[0])
