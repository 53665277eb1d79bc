"""Generates the golden fixtures under tests/golden (run once, in a CPU
container; the .npz files are committed and the GPU tests only read them).

Sources of truth, none of them this repository's own code:
  1. The reference's own unit-test recipes (rocmarchive/hcBLAS test/unit/*.cpp):
     integer-valued inputs from glibc rand_r(&seed) with seed 100, the same
     sizes and alpha/beta; inputs are regenerated here by calling the C
     library's rand_r through ctypes, in the reference's call order.
  2. Expected outputs from CBLAS -- the reference tests' oracle -- as bundled
     in this image (scipy.linalg.blas, OpenBLAS). Integer-valued data make the
     results exact and independent of summation order.
  3. Exact rational arithmetic (fractions.Fraction) for tiny real-valued
     problems, giving the exactly-rounded answer.

Usage:  python tests/golden/gen_golden.py
"""
from __future__ import annotations

import ctypes
import os
from fractions import Fraction

import numpy as np
from scipy.linalg import blas

OUT = os.path.dirname(os.path.abspath(__file__))
ROW, COL = 0, 1
N_, T_, C_ = 0, 1, 2

_libc = ctypes.CDLL("libc.so.6")
_libc.rand_r.argtypes = [ctypes.POINTER(ctypes.c_uint)]
_libc.rand_r.restype = ctypes.c_int


class RandR:
    """glibc rand_r stream (test/unit/*.cpp: unsigned int global_seed = 100)."""

    def __init__(self, seed=100):
        self.s = ctypes.c_uint(seed)

    def __call__(self, mod):
        return _libc.rand_r(ctypes.byref(self.s)) % mod


# ---------------------------------------------------------------------------
# CBLAS through scipy: dense storage (ld == rows of the stored array).
# ---------------------------------------------------------------------------
def _as2d(flat, rows, cols, ld, order):
    """Stored matrix (rows x cols as seen by the caller in `order`)."""
    if order == COL:
        return flat[: ld * cols].reshape(cols, ld).T[:rows, :]
    return flat[: ld * rows].reshape(rows, ld)[:, :cols]


def cblas_gemm(order, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc):
    """CBLAS semantics: row-major is the column-major call on the transposed
    storage with swapped operands, C^T = op(B)^T op(A)^T."""
    if order == ROW:
        return cblas_gemm(COL, tb, ta, N, M, K, alpha, B, ldb, A, lda, beta, C,
                          ldc)
    dt = C.dtype
    fn = {np.dtype(np.float32): blas.sgemm, np.dtype(np.float64): blas.dgemm,
          np.dtype(np.complex64): blas.cgemm,
          np.dtype(np.complex128): blas.zgemm}[dt]
    a_rows, a_cols = (M, K) if ta == N_ else (K, M)
    b_rows, b_cols = (K, N) if tb == N_ else (N, K)
    a2 = _as2d(A, a_rows, a_cols, lda, COL)
    b2 = _as2d(B, b_rows, b_cols, ldb, COL)
    c2 = _as2d(C, M, N, ldc, COL).copy()
    r = fn(alpha, np.asfortranarray(a2), np.asfortranarray(b2), beta,
           np.asfortranarray(c2), trans_a=ta, trans_b=tb)
    out = C.copy()
    out[: ldc * N].reshape(N, ldc).T[:M, :] = r
    return out


def cblas_gemv(order, trans, M, N, alpha, A, lda, x, beta, y):
    """CBLAS: row-major gemv is column-major gemv of A^T (N x M) with the
    transpose flag flipped."""
    fn = blas.sgemv if A.dtype == np.float32 else blas.dgemv
    if order == ROW:
        a2 = _as2d(A, N, M, lda, COL)
        trans = 0 if trans != N_ else 1
    else:
        a2 = _as2d(A, M, N, lda, COL)
    return fn(alpha, np.asfortranarray(a2), x, beta=beta, y=y.copy(),
              trans=trans)


def cblas_ger(order, M, N, alpha, x, y, A, lda):
    """CBLAS: row-major ger is column-major ger of A^T with x and y swapped."""
    fn = blas.sger if A.dtype == np.float32 else blas.dger
    if order == ROW:
        a2 = np.asfortranarray(_as2d(A, N, M, lda, COL))
        r = fn(alpha, y, x, a=a2)
        out = A.copy()
        out[: lda * M].reshape(M, lda).T[:N, :] = r
        return out
    a2 = np.asfortranarray(_as2d(A, M, N, lda, COL))
    r = fn(alpha, x, y, a=a2)
    out = A.copy()
    out[: lda * N].reshape(N, lda).T[:M, :] = r
    return out


# The eight (order, transA, transB, lda, ldb, ldc) cases of
# test/unit/hcblas_sgemm.cpp:73-148 for M=189, N=9, K=19.
def _gemm_cases(M, N, K):
    return [
        (COL, N_, N_, M, K, M), (ROW, N_, N_, K, N, N),
        (COL, N_, T_, M, N, M), (ROW, N_, T_, K, K, N),
        (COL, T_, N_, K, K, M), (ROW, T_, N_, M, N, N),
        (COL, T_, T_, K, N, M), (ROW, T_, T_, M, K, N),
    ]


def gen_gemm_ref_int():
    M, N, K = 189, 9, 19
    out = {}
    for dt, name in ((np.float32, "s"), (np.float64, "d")):
        r = RandR(100)  # hcblas_sgemm.cpp:32 / hcblas_dgemm.cpp
        A = np.array([r(100) for _ in range(M * K)], dt)
        B = np.array([r(15) for _ in range(K * N)], dt)
        C = np.array([r(25) for _ in range(M * N)], dt)
        out[f"{name}_A"], out[f"{name}_B"], out[f"{name}_C"] = A, B, C
        for i, (o, ta, tb, lda, ldb, ldc) in enumerate(_gemm_cases(M, N, K)):
            out[f"{name}_case{i}"] = np.array([o, ta, tb, lda, ldb, ldc])
            out[f"{name}_expect{i}"] = cblas_gemm(o, ta, tb, M, N, K, 1.0, A,
                                                  lda, B, ldb, 1.0, C, ldc)
    for dt, name in ((np.complex64, "c"), (np.complex128, "z")):
        r = RandR(100)  # hcblas_cgemm.cpp:69-79
        A = np.empty(M * K, dt)
        for i in range(M * K):
            A[i] = complex(r(10), r(20))
        B = np.empty(K * N, dt)
        for i in range(K * N):
            B[i] = complex(r(15), r(25))
        C = np.empty(M * N, dt)
        for i in range(M * N):
            r(18)
            r(28)
            v = r(25)  # C[i] = rand_r % 25 sets both components
            C[i] = complex(v, v)
        out[f"{name}_A"], out[f"{name}_B"], out[f"{name}_C"] = A, B, C
        alpha, beta = 1 + 1j, 1 + 1j  # cAlpha = cBeta = (1, 1)
        cases = _gemm_cases(M, N, K)
        # plus the conjugate-transpose forms (CBLAS semantics)
        cases += [(COL, C_, N_, K, K, M), (ROW, N_, C_, K, K, N),
                  (COL, C_, C_, K, N, M)]
        for i, (o, ta, tb, lda, ldb, ldc) in enumerate(cases):
            out[f"{name}_case{i}"] = np.array([o, ta, tb, lda, ldb, ldc])
            out[f"{name}_expect{i}"] = cblas_gemm(o, ta, tb, M, N, K, alpha, A,
                                                  lda, B, ldb, beta, C, ldc)
    out["dims"] = np.array([M, N, K])
    np.savez_compressed(os.path.join(OUT, "gemm_ref_int.npz"), **out)


def gen_level2_ref_int():
    M, N = 179, 19
    out = {"dims": np.array([M, N])}
    for dt, name in ((np.float32, "s"), (np.float64, "d")):
        # gemv: hcblas_sgemv.cpp:70-80 (x %10, A %25, y %15), alpha = beta = 1
        for trans in (N_, T_):
            lenx, leny = (N, M) if trans == N_ else (M, N)
            r = RandR(100)
            x = np.array([r(10) for _ in range(lenx)], dt)
            A = np.array([r(25) for _ in range(M * N)], dt)
            y = np.array([r(15) for _ in range(leny)], dt)
            for order in (COL, ROW):
                lda = M if order == COL else N
                key = f"{name}gemv_t{trans}_o{order}"
                out[key + "_x"], out[key + "_A"], out[key + "_y"] = x, A, y
                out[key + "_expect"] = cblas_gemv(order, trans, M, N, 1.0, A,
                                                  lda, x, 1.0, y)
        # ger: hcblas_sger.cpp:60-70 (x %10, y %15, A %25), alpha = 1; plus
        # alpha = 3 to exercise the alpha*y rounding path
        r = RandR(100)
        x = np.array([r(10) for _ in range(M)], dt)
        y = np.array([r(15) for _ in range(N)], dt)
        A = np.array([r(25) for _ in range(M * N)], dt)
        for order in (COL, ROW):
            lda = M if order == COL else N
            for alpha in (1.0, 3.0):
                key = f"{name}ger_o{order}_a{int(alpha)}"
                out[key + "_x"], out[key + "_y"], out[key + "_A"] = x, y, A
                out[key + "_expect"] = cblas_ger(order, M, N, alpha, x, y, A,
                                                 lda)
    np.savez_compressed(os.path.join(OUT, "level2_ref_int.npz"), **out)


def gen_level1_ref_int():
    out = {}
    for dt, name in ((np.float32, "s"), (np.float64, "d")):
        b = blas
        axpy = b.saxpy if name == "s" else b.daxpy
        scal = b.sscal if name == "s" else b.dscal
        dot = b.sdot if name == "s" else b.ddot
        asum = b.sasum if name == "s" else b.dasum
        # axpy N=279: X %10, Y %15 (hcblas_saxpy.cpp:36-57)
        r = RandR(100)
        n = 279
        X = np.array([r(10) for _ in range(n)], dt)
        Y = np.array([r(15) for _ in range(n)], dt)
        out[f"{name}axpy_x"], out[f"{name}axpy_y"] = X, Y
        for alpha in (1.0, 3.0):
            out[f"{name}axpy_a{int(alpha)}"] = axpy(X, Y.copy(), a=alpha)
        # scal N=23: X %10 (hcblas_sscal.cpp:36-48)
        r = RandR(100)
        X = np.array([r(10) for _ in range(23)], dt)
        out[f"{name}scal_x"] = X
        for alpha in (1.0, 3.0, 0.0):
            out[f"{name}scal_a{int(alpha)}"] = scal(alpha, X.copy())
        # copy N=23: X %10, Y %15
        r = RandR(100)
        X = np.array([r(10) for _ in range(23)], dt)
        Y = np.array([r(15) for _ in range(23)], dt)
        out[f"{name}copy_x"], out[f"{name}copy_y"] = X, Y
        # dot: X %10, Y %15; sizes from the reference test (N=23) and a
        # larger one
        for n in (23, 1000):
            r = RandR(100)
            X = np.array([r(10) for _ in range(n)], dt)
            Y = np.array([r(15) for _ in range(n)], dt)
            out[f"{name}dot{n}_x"], out[f"{name}dot{n}_y"] = X, Y
            out[f"{name}dot{n}"] = np.array(dot(X, Y), np.float64)
        # asum N=119: X %10 (hcblas_sasum.cpp:36-48); batched API test n=23
        # x 128 (test-blas.cpp:85-121) returns one sum over the batch
        r = RandR(100)
        X = np.array([r(10) for _ in range(119)], dt)
        out[f"{name}asum_x"] = X
        out[f"{name}asum"] = np.array(asum(X), np.float64)
        r = RandR(100)
        X = np.array([r(10) for _ in range(23 * 128)], dt)
        out[f"{name}asumb_x"] = X
        out[f"{name}asumb"] = np.array(
            sum(float(asum(X[i * 23:(i + 1) * 23])) for i in range(128)),
            np.float64)
    np.savez_compressed(os.path.join(OUT, "level1_ref_int.npz"), **out)


def gen_real_valued():
    """Real-valued problems (uniform[-1,1]) with CBLAS outputs, compared with
    a tolerance (summation order differs between implementations)."""
    rng = np.random.default_rng(20240611)
    out = {}
    M, N, K = 37, 29, 41
    out["dims"] = np.array([M, N, K])
    for dt, name in ((np.float32, "s"), (np.float64, "d"),
                     (np.complex64, "c"), (np.complex128, "z")):
        cplx = np.iscomplexobj(np.zeros(1, dt))

        def rnd(n):
            v = rng.uniform(-1, 1, n)
            if cplx:
                v = v + 1j * rng.uniform(-1, 1, n)
            return v.astype(dt)

        A, B, C = rnd(M * K), rnd(K * N), rnd(M * N)
        alpha = (0.75 - 0.5j) if cplx else 0.75
        beta = (-1.25 + 0.25j) if cplx else -1.25
        out[f"{name}_A"], out[f"{name}_B"], out[f"{name}_C"] = A, B, C
        out[f"{name}_ab"] = np.array([alpha, beta], np.complex128)
        for i, (o, ta, tb, lda, ldb, ldc) in enumerate(_gemm_cases(M, N, K)):
            out[f"{name}_case{i}"] = np.array([o, ta, tb, lda, ldb, ldc])
            out[f"{name}_expect{i}"] = cblas_gemm(o, ta, tb, M, N, K, alpha, A,
                                                  lda, B, ldb, beta, C, ldc)
    # axpy / ger real-valued: OpenBLAS contraction = the defined contraction,
    # so these are bit-exact references.
    for dt, name in ((np.float32, "s"), (np.float64, "d")):
        n = 4099
        x = rng.uniform(-1, 1, n).astype(dt)
        y = rng.uniform(-1, 1, n).astype(dt)
        out[f"{name}axpy_x"], out[f"{name}axpy_y"] = x, y
        axpy = blas.saxpy if name == "s" else blas.daxpy
        out[f"{name}axpy"] = axpy(x, y.copy(), a=1.37)
        m, nn = 67, 45
        gx = rng.uniform(-1, 1, m).astype(dt)
        gy = rng.uniform(-1, 1, nn).astype(dt)
        GA = rng.uniform(-1, 1, m * nn).astype(dt)
        out[f"{name}ger_x"], out[f"{name}ger_y"], out[f"{name}ger_A"] = gx, gy, GA
        for order in (COL, ROW):
            lda = m if order == COL else nn
            out[f"{name}ger_o{order}"] = cblas_ger(order, m, nn, 1.37, gx, gy,
                                                   GA, lda)
        out[f"{name}ger_dims"] = np.array([m, nn])
        # gemv real-valued
        gm, gn = 53, 71
        Av = rng.uniform(-1, 1, gm * gn).astype(dt)
        xv = rng.uniform(-1, 1, max(gm, gn)).astype(dt)
        yv = rng.uniform(-1, 1, max(gm, gn)).astype(dt)
        out[f"{name}gemv_A"], out[f"{name}gemv_x"], out[f"{name}gemv_y"] = Av, xv, yv
        out[f"{name}gemv_dims"] = np.array([gm, gn])
        for trans in (N_, T_):
            lenx, leny = (gn, gm) if trans == N_ else (gm, gn)
            for order in (COL, ROW):
                lda = gm if order == COL else gn
                out[f"{name}gemv_t{trans}_o{order}"] = cblas_gemv(
                    order, trans, gm, gn, 1.5, Av, lda, xv[:lenx], 0.5,
                    yv[:leny])
    np.savez_compressed(os.path.join(OUT, "real_valued.npz"), **out)


def _exact_round_f32(q: Fraction) -> np.float32:
    """Correctly rounded (to nearest even) binary32 value of a rational."""
    f = np.float32(float(q))  # may be off by one ulp from double rounding
    best = f
    for cand in (np.nextafter(f, np.float32(-np.inf)), f,
                 np.nextafter(f, np.float32(np.inf))):
        if abs(Fraction(float(cand)) - q) < abs(Fraction(float(best)) - q):
            best = cand
    # ties: keep the even mantissa
    lo = np.nextafter(best, np.float32(-np.inf))
    hi = np.nextafter(best, np.float32(np.inf))
    for other in (lo, hi):
        if abs(Fraction(float(other)) - q) == abs(Fraction(float(best)) - q):
            if int(np.float32(other).view(np.uint32)) % 2 == 0:
                best = other
    return np.float32(best)


def gen_exact_rational():
    """Tiny GEMM/GEMV/dot/axpy/ger on dyadic fp32 inputs, exact results."""
    rng = np.random.default_rng(7)
    out = {}
    M, N, K = 5, 4, 3
    A = (rng.integers(-2**20, 2**20, M * K) / 2.0**18).astype(np.float32)
    B = (rng.integers(-2**20, 2**20, K * N) / 2.0**19).astype(np.float32)
    C = (rng.integers(-2**20, 2**20, M * N) / 2.0**17).astype(np.float32)
    alpha, beta = np.float32(1.5), np.float32(-0.75)
    exact = np.empty(M * N, np.float64)
    for j in range(N):
        for i in range(M):
            s = sum(Fraction(float(A[i + l * M])) * Fraction(float(B[l + j * K]))
                    for l in range(K))
            v = Fraction(float(alpha)) * s + Fraction(float(beta)) * Fraction(
                float(C[i + j * M]))
            exact[i + j * M] = float(v)  # exact: few bits, fits a double
    out.update(gemm_A=A, gemm_B=B, gemm_C=C, gemm_exact=exact,
               gemm_ab=np.array([alpha, beta], np.float32),
               gemm_dims=np.array([M, N, K]))
    # axpy / ger single-rounding references: round(alpha*x + y) exactly
    n = 257
    x = rng.uniform(-1, 1, n).astype(np.float32)
    y = rng.uniform(-1, 1, n).astype(np.float32)
    a = np.float32(1.37)
    fma = np.array([_exact_round_f32(Fraction(float(a)) * Fraction(float(xi)) +
                                     Fraction(float(yi)))
                    for xi, yi in zip(x, y)], np.float32)
    out.update(axpy_x=x, axpy_y=y, axpy_alpha=np.array([a]), axpy_fma=fma)
    m, nn = 9, 7
    gx = rng.uniform(-1, 1, m).astype(np.float32)
    gy = rng.uniform(-1, 1, nn).astype(np.float32)
    GA = rng.uniform(-1, 1, m * nn).astype(np.float32)
    ger = GA.copy()
    for j in range(nn):
        t = np.float32(a * gy[j])  # netlib temp = alpha*y(j), one rounding
        for i in range(m):
            ger[i + j * m] = _exact_round_f32(
                Fraction(float(gx[i])) * Fraction(float(t)) +
                Fraction(float(GA[i + j * m])))
    out.update(ger_x=gx, ger_y=gy, ger_A=GA, ger_exact=ger,
               ger_dims=np.array([m, nn]))
    # dot exact value
    dx = rng.uniform(-1, 1, 100).astype(np.float32)
    dy = rng.uniform(-1, 1, 100).astype(np.float32)
    out.update(dot_x=dx, dot_y=dy, dot_exact=np.array(float(sum(
        Fraction(float(p)) * Fraction(float(q)) for p, q in zip(dx, dy)))))
    np.savez_compressed(os.path.join(OUT, "exact_rational.npz"), **out)


if __name__ == "__main__":
    gen_gemm_ref_int()
    gen_level2_ref_int()
    gen_level1_ref_int()
    gen_real_valued()
    gen_exact_rational()
    for f in sorted(os.listdir(OUT)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(OUT, f)))
