"""Split-precision GEMM on the GPU (no reference counterpart):
BF16 three-way split SGEMM / CGEMM ("s"), Ozaki scheme II on INT8 for SGEMM /
CGEMM with norm-scaled operands ("sc"), Ozaki INT8 slice DGEMM / ZGEMM ("d")
and Ozaki scheme II (CRT over INT8 moduli) DGEMM / ZGEMM ("dc").

Checks:
  - bit-exact on the reference's integer recipes (test/unit/hcblas_sgemm.cpp
    189x9x19, all order x trans cases): small integers split / slice exactly;
  - every order x trans case within the plain kernel's error bound of the
    double-double oracle;
  - measured error against the double-double oracle no worse than the plain
    FP32 / FP64 MFMA kernel on the same inputs, for uniform[-1,1] and for the
    wide-exponent distribution sign*2^u, u ~ U[-20, 20];
  - the handle's math mode routes hcblasSgemm / hcblasDgemm / C / Z GEMM to
    the split kernels with results identical to the explicit entries.
"""
import ctypes

import numpy as np
import pytest

from gemm_bounds import check_sampled
from gpu_util import bitwise_equal, to_dev, to_host

pytestmark = pytest.mark.gpu
ROW, COL = 0, 1
N_, T_, C_ = 0, 1, 2


def _split(hc, handle, p, o, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C,
           ldc):
    api = hc.api
    dA, dB, dC = to_dev(A), to_dev(B), to_dev(C)
    fn = {"s": api.hcblas_sgemm_split_bf16, "sc": api.hcblas_sgemm_ozaki_crt,
          "d": api.hcblas_dgemm_ozaki_int8,
          "dc": api.hcblas_dgemm_ozaki_crt}[p]
    st = fn(handle, o, ta, tb, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc,
            0, 0, 0)
    assert st == 0
    return to_host(dC)


def _plain(hc, handle, p, o, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C,
           ldc):
    api = hc.api
    dA, dB, dC = to_dev(A), to_dev(B), to_dev(C)
    st = getattr(api, f"hcblas_{p[0]}gemm")(handle, o, ta, tb, M, N, K, alpha, dA,
                                         lda, dB, ldb, beta, dC, ldc, 0, 0, 0)
    assert st == 0
    return to_host(dC)


@pytest.mark.parametrize("p", ["s", "sc", "d", "dc"])
def test_reference_recipe_bit_exact(hc, handle, golden, p):
    g = golden("gemm_ref_int.npz")
    M, N, K = (int(v) for v in g["dims"])
    q = p[0]
    for i in range(8):
        o, ta, tb, lda, ldb, ldc = (int(v) for v in g[f"{q}_case{i}"])
        got = _split(hc, handle, p, o, ta, tb, M, N, K, 1.0, g[f"{q}_A"], lda,
                     g[f"{q}_B"], ldb, 1.0, g[f"{q}_C"], ldc)
        assert bitwise_equal(got, g[f"{q}_expect{i}"]), f"case {i}"


def _gen(orc, p, n, seed, wide):
    if p[0] == "s":
        return orc.wide_f32(n, seed) if wide else orc.uniform_f32(n, seed)
    return orc.wide_f64(n, seed) if wide else orc.uniform_f64(n, seed)


def _rel_errors(orc, o, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                got, rows, cols):
    ref = orc.gemm_sample_exact(o, ta, tb, M, N, K, alpha, A, lda, B, ldb,
                                beta, C, ldc, rows, cols)
    idx = rows + cols * ldc if o == COL else rows * ldc + cols
    return np.abs(got[idx] - ref) / np.maximum(np.abs(ref), 1e-300)


@pytest.mark.parametrize("p", ["s", "sc", "d", "dc"])
@pytest.mark.parametrize("M,N,K", [(33, 17, 45), (128, 256, 512), (300, 129, 1000),
                                   (272, 300, 130)])
def test_all_cases_within_plain_bound(hc, handle, orc, p, M, N, K):
    eps = np.finfo(np.float32 if p[0] == "s" else np.float64).eps
    for o in (COL, ROW):
        for ta in (N_, T_):
            for tb in (N_, T_):
                ar, ac = (M, K) if ta == N_ else (K, M)
                br, bc = (K, N) if tb == N_ else (N, K)
                lda = (ar if o == COL else ac) + 1
                ldb = (br if o == COL else bc) + 2
                ldc = (M if o == COL else N)
                A = _gen(orc, p, lda * (ac if o == COL else ar), 1, False)
                B = _gen(orc, p, ldb * (bc if o == COL else br), 2, False)
                C = _gen(orc, p, ldc * (N if o == COL else M), 3, False)
                got = _split(hc, handle, p, o, ta, tb, M, N, K, 0.5, A, lda, B,
                             ldb, -1.5, C, ldc)
                rng = np.random.default_rng(0)
                rows = rng.integers(0, M, 256)
                cols = rng.integers(0, N, 256)
                check_sampled(orc, o, ta, tb, M, N, K, 0.5, A, lda, B, ldb,
                              -1.5, C, ldc, got, rows, cols, eps, c=4.0)


@pytest.mark.parametrize("p", ["s", "sc", "d", "dc"])
@pytest.mark.parametrize("wide", [False, True])
@pytest.mark.parametrize("ta", [N_, T_])
def test_error_no_worse_than_plain(hc, handle, orc, p, wide, ta):
    """Measured max / mean relative error vs the double-double oracle on
    sampled entries: split <= plain (K = 4096)."""
    M = N = 384
    K = 4096
    lda = M if ta == N_ else K
    A = _gen(orc, p, M * K, 11, wide)
    B = _gen(orc, p, K * N, 12, wide)
    C = np.zeros(M * N, A.dtype)
    rng = np.random.default_rng(1)
    rows = rng.integers(0, M, 1024)
    cols = rng.integers(0, N, 1024)
    plain = _plain(hc, handle, p, COL, ta, N_, M, N, K, 1.0, A, lda, B, K, 0.0,
                   C, M)
    split = _split(hc, handle, p, COL, ta, N_, M, N, K, 1.0, A, lda, B, K, 0.0,
                   C, M)
    ep = _rel_errors(orc, COL, ta, N_, M, N, K, 1.0, A, lda, B, K, 0.0, C, M,
                     plain, rows, cols)
    es = _rel_errors(orc, COL, ta, N_, M, N, K, 1.0, A, lda, B, K, 0.0, C, M,
                     split, rows, cols)
    print(f"{p} wide={wide} ta={ta}: plain max {ep.max():.3e} mean {ep.mean():.3e}"
          f" | split max {es.max():.3e} mean {es.mean():.3e}")
    assert es.mean() <= ep.mean()
    assert np.median(es) <= np.median(ep)


def test_math_mode_routes_plain_entries(hc, handle, orc):
    lib = hc.load()
    api = hc.api
    M, N, K = 200, 150, 300
    for p, mode in (("s", 1), ("sc", 8), ("sc", 9), ("d", 2), ("dc", 4),
                    ("dc", 6)):
        A = _gen(orc, p, M * K, 5, False)
        B = _gen(orc, p, K * N, 6, False)
        C = _gen(orc, p, M * N, 7, False)
        exp = _split(hc, handle, p, COL, N_, T_, M, N, K, 1.25, A, M,
                     B[:N * K], N, 0.5, C, M)
        assert lib.hcblasSetMathMode(handle.value, mode) == 0
        dC = to_dev(C)
        st = getattr(api, f"hcblas{p[0].upper()}gemm")(handle, N_, T_, M, N, K,
                                                    1.25, to_dev(A), M,
                                                    to_dev(B[:N * K]), N, 0.5,
                                                    dC, M)
        assert st == 0
        assert bitwise_equal(to_host(dC), exp)
        assert lib.hcblasSetMathMode(handle.value, 0) == 0


@pytest.mark.parametrize("p,mode", [("c", 1), ("c", 8), ("z", 2), ("z", 4)])
def test_complex_split_modes(hc, handle, orc, p, mode):
    lib = hc.load()
    api = hc.api
    M, N, K = 96, 80, 200
    gen = orc.uniform_f32 if p == "c" else orc.uniform_f64
    dt = np.complex64 if p == "c" else np.complex128

    def cx(n, seed):
        v = gen(2 * n, seed)
        return (v[0::2] + 1j * v[1::2]).astype(dt)

    A, B, C = cx(M * K, 1), cx(K * N, 2), cx(M * N, 3)
    alpha, beta = 0.5 - 1j, 0.25 + 0.5j
    assert lib.hcblasSetMathMode(handle.value, mode) == 0
    for ta in (N_, C_):
        dC = to_dev(C)
        lda = M if ta == N_ else K
        st = getattr(api, f"hcblas{p.upper()}gemm")(handle, ta, N_, M, N, K,
                                                    alpha, to_dev(A), lda,
                                                    to_dev(B), K, beta, dC, M)
        assert st == 0
        got = to_host(dC)
        rng = np.random.default_rng(2)
        rows = rng.integers(0, M, 256)
        cols = rng.integers(0, N, 256)
        ref = orc.gemm_sample_exact(COL, ta, N_, M, N, K, alpha, A, lda, B, K,
                                    beta, C, M, rows, cols)
        eps = np.finfo(dt).eps
        tol = 2 * (K + 2) * eps * (abs(alpha) * K * 2 + abs(beta) * 2 + 1)
        assert np.max(np.abs(got[rows + cols * M] - ref)) <= tol
    assert lib.hcblasSetMathMode(handle.value, 0) == 0


def test_ozaki_slices_control_accuracy(hc, handle, orc):
    lib = hc.load()
    M = N = 128
    K = 2048
    A = orc.uniform_f64(M * K, 1)
    B = orc.uniform_f64(K * N, 2)
    C = np.zeros(M * N)
    rng = np.random.default_rng(3)
    rows = rng.integers(0, M, 512)
    cols = rng.integers(0, N, 512)
    errs = []
    for s in (4, 6, 8, 10):
        assert lib.hcblasSetOzakiSlices(handle.value, s) == 0
        got = _split(hc, handle, "d", COL, N_, N_, M, N, K, 1.0, A, M, B, K,
                     0.0, C, M)
        errs.append(_rel_errors(orc, COL, N_, N_, M, N, K, 1.0, A, M, B, K,
                                0.0, C, M, got, rows, cols).max())
    assert lib.hcblasSetOzakiSlices(handle.value, 8) == 0
    print("ozaki max rel err by slices 4,6,8,10:", errs)
    assert errs[0] > errs[1] > errs[2]
    assert errs[3] <= errs[2]


def test_crt_moduli_control_accuracy(hc, handle, orc):
    """More moduli -> more operand bits (25, 41, 56 and 62 at K = 2048 for
    8, 12, 16, 20 moduli) -> smaller error; 16 moduli are at least as
    accurate as the plain FP64 kernel (the default count is checked against
    it in test_error_no_worse_than_plain)."""
    lib = hc.load()
    M = N = 128
    K = 2048
    A = orc.uniform_f64(M * K, 1)
    B = orc.uniform_f64(K * N, 2)
    C = np.zeros(M * N)
    rng = np.random.default_rng(3)
    rows = rng.integers(0, M, 512)
    cols = rng.integers(0, N, 512)
    errs = []
    for n in (8, 12, 16, 20):
        assert lib.hcblasSetOzakiModuli(handle.value, n) == 0
        got = _split(hc, handle, "dc", COL, N_, N_, M, N, K, 1.0, A, M, B, K,
                     0.0, C, M)
        errs.append(_rel_errors(orc, COL, N_, N_, M, N, K, 1.0, A, M, B, K,
                                0.0, C, M, got, rows, cols).max())
    assert lib.hcblasSetOzakiModuli(handle.value, 0) == 0  # back to auto
    plain = _plain(hc, handle, "d", COL, N_, N_, M, N, K, 1.0, A, M, B, K, 0.0,
                   C, M)
    ep = _rel_errors(orc, COL, N_, N_, M, N, K, 1.0, A, M, B, K, 0.0, C, M,
                     plain, rows, cols).max()
    print("crt max rel err by moduli 8,12,16,20:", errs, "plain", ep)
    assert errs[0] > errs[1] > errs[2]
    assert errs[3] <= errs[2]
    assert errs[2] <= ep
    assert lib.hcblasSetOzakiModuli(handle.value, 5) == 3
    assert lib.hcblasSetOzakiModuli(handle.value, 21) == 3


@pytest.mark.parametrize("wide", [False, True])
def test_crt_f32_moduli_control_accuracy(hc, handle, orc, wide):
    """FP32 CRT: 6, 7, 8, 10 moduli scale the operand norms to 2^23, 2^27,
    2^30, 2^30; the error falls with the first three and stays put at 10 (the
    scale is capped at 2^30); the default (8) is below the FP32 kernel's
    error in max, mean and median."""
    lib = hc.load()
    M = N = 192
    K = 2048
    gen = orc.wide_f32 if wide else orc.uniform_f32
    A = gen(M * K, 41)
    B = gen(K * N, 42)
    C = np.zeros(M * N, np.float32)
    rng = np.random.default_rng(4)
    rows = rng.integers(0, M, 1024)
    cols = rng.integers(0, N, 1024)

    def err(got):
        return _rel_errors(orc, COL, N_, N_, M, N, K, 1.0, A, M, B, K, 0.0, C,
                           M, got, rows, cols)

    errs = {}
    for n in (6, 7, 8, 10, 0):
        assert lib.hcblasSetOzakiModuli(handle.value, n) == 0
        errs[n] = err(_split(hc, handle, "sc", COL, N_, N_, M, N, K, 1.0, A, M,
                             B, K, 0.0, C, M))
    assert lib.hcblasSetOzakiModuli(handle.value, 0) == 0
    ep = err(_plain(hc, handle, "s", COL, N_, N_, M, N, K, 1.0, A, M, B, K,
                    0.0, C, M))
    print("crt f32 mean rel err by moduli:",
          {n: float(e.mean()) for n, e in errs.items()}, "plain", ep.mean())
    assert errs[6].mean() > errs[7].mean() > errs[8].mean()
    assert errs[10].mean() <= 1.01 * errs[8].mean()
    assert np.array_equal(errs[0], errs[8])
    for f in (np.max, np.mean, np.median):
        assert f(errs[0]) <= f(ep)


@pytest.mark.parametrize("p", ["dc", "sc"])
def test_crt_zero_rows_and_k_chunks(hc, handle, orc, p):
    """All-zero rows / columns (no exponent) give exact zeros; K above the
    65536 chunk runs in two chunks with beta applied once."""
    M, N, K = 40, 24, 70000
    gen = orc.uniform_f64 if p == "dc" else orc.uniform_f32
    A = gen(M * K, 21)
    B = gen(K * N, 22)
    A = A.reshape(K, M)
    A[:, 5] = 0.0  # row 5 of A (column-major M x K)
    A = A.reshape(-1)
    B = B.reshape(N, K)
    B[7, :] = 0.0  # column 7 of B
    B = B.reshape(-1)
    C = gen(M * N, 23)
    got = _split(hc, handle, p, COL, N_, N_, M, N, K, 1.0, A, M, B, K, 0.5,
                 C, M)
    rows = np.repeat(np.arange(M), N)
    cols = np.tile(np.arange(N), M)
    ref = orc.gemm_sample_exact(COL, N_, N_, M, N, K, 1.0, A, M, B, K, 0.5, C,
                                M, rows, cols)
    idx = rows + cols * M
    ulp = 2.0 ** -52 if p == "dc" else 2.0 ** -23
    assert np.max(np.abs(got[idx] - ref)) <= 4 * K * ulp
    Cm = got.reshape(N, M)
    half = (C * np.asarray(0.5, C.dtype)).reshape(N, M)
    assert np.array_equal(Cm[:, 5], half[:, 5])
    assert np.array_equal(Cm[7, :], half[7, :])


@pytest.mark.parametrize("p", ["s", "sc", "d", "dc"])
def test_non_finite_inputs_propagate(hc, handle, orc, p):
    """A NaN or Inf in a row of A or a column of B makes the outputs of that
    row / column non-finite (NaN in the Ozaki modes, where the operand scaling
    has no finite value to work with); every other output keeps its
    accuracy."""
    M, N, K = 64, 48, 96
    gen = orc.uniform_f32 if p[0] == "s" else orc.uniform_f64
    A = gen(M * K, 31)
    B = gen(K * N, 32)
    A[3 + 10 * M] = np.nan  # A(3, 10), column-major M x K
    A[7 + 20 * M] = np.inf  # A(7, 20)
    B[5 + 2 * K] = -np.inf  # B(5, 2), column-major K x N
    C = np.zeros(M * N, A.dtype)
    got = _split(hc, handle, p, COL, N_, N_, M, N, K, 1.0, A, M, B, K, 0.0,
                 C, M).reshape(N, M)
    bad = np.zeros((N, M), bool)
    bad[:, 3] = bad[:, 7] = True
    bad[2, :] = True
    assert not np.any(np.isfinite(got[bad]))
    if p in ("sc", "d", "dc"):
        assert np.all(np.isnan(got[bad]))
    Af = A.copy()
    Bf = B.copy()
    Af[3 + 10 * M] = Af[7 + 20 * M] = 0.0
    Bf[5 + 2 * K] = 0.0
    cols, rows = np.nonzero(~bad)
    ref = orc.gemm_sample_exact(COL, N_, N_, M, N, K, 1.0, Af, M, Bf, K, 0.0,
                                C, M, rows, cols)
    eps = np.finfo(A.dtype).eps
    assert np.max(np.abs(got[cols, rows] - ref)) <= 4 * K * eps
