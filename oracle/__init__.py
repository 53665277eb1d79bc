"""CPU oracle for the hcBLAS hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this package, and only as the checker: the product path
(``hcblas_amd``) never calls it.

It wraps ``oracle/lib/liboracle.so`` (``refblas.c``: netlib-order restatement
of CBLAS, FP64 and double-double references, the seeded generator) with
numpy-array helpers, and adds the numpy reference for HGEMM. Each function
names the reference behaviour it restates in ``refblas.c``.

Parity pinning (tests/test_oracle.py, tests/test_oracle_pins.py): every
function here is checked against the golden fixtures under ``tests/golden``
(bundled-CBLAS outputs on the reference tests' own integer recipes,
exact-rational brute force), against closed forms, and the accuracy
references (double-double dot / asum / sampled GEMM, hgemm) against exact
rational arithmetic, with mutation checks that a plausible mistake fails.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "liboracle.so")

ROW_MAJOR, COL_MAJOR = 0, 1
OP_N, OP_T, OP_C = 0, 1, 2

_lib = None

_P = ctypes.c_void_p
_I = ctypes.c_int
_L = ctypes.c_int64
_F = ctypes.c_float
_D = ctypes.c_double
_U = ctypes.c_uint64

_SIGS = {
    "orc_fill_uniform_f32": [_P, _L, _U, _U],
    "orc_fill_uniform_f64": [_P, _L, _U, _U],
    "orc_fill_wide_f32": [_P, _L, _U, _U, _D],
    "orc_fill_wide_f64": [_P, _L, _U, _U, _D],
    "orc_saxpy": [_L, _F, _P, _L, _P, _L],
    "orc_daxpy": [_L, _D, _P, _L, _P, _L],
    "orc_sscal": [_L, _F, _P, _L],
    "orc_dscal": [_L, _D, _P, _L],
    "orc_cscal": [_L, _F, _F, _P, _L],
    "orc_zscal": [_L, _D, _D, _P, _L],
    "orc_csscal": [_L, _F, _P, _L],
    "orc_zdscal": [_L, _D, _P, _L],
    "orc_scopy": [_L, _P, _L, _P, _L],
    "orc_dcopy": [_L, _P, _L, _P, _L],
    "orc_sdot": [_L, _P, _L, _P, _L],
    "orc_ddot": [_L, _P, _L, _P, _L],
    "orc_sasum": [_L, _P, _L],
    "orc_dasum": [_L, _P, _L],
    "orc_sdot_f64": [_L, _P, _L, _P, _L],
    "orc_ddot_dd": [_L, _P, _L, _P, _L, _P, _P],
    "orc_dasum_dd": [_L, _P, _L, _P, _P],
    "orc_sgemv": [_I, _I, _L, _L, _F, _P, _L, _P, _L, _F, _P, _L],
    "orc_dgemv": [_I, _I, _L, _L, _D, _P, _L, _P, _L, _D, _P, _L],
    "orc_sgemv_ref": [_I, _I, _L, _L, _F, _P, _L, _P, _L, _F, _P, _L, _P],
    "orc_dgemv_ref": [_I, _I, _L, _L, _D, _P, _L, _P, _L, _D, _P, _L, _P],
    "orc_sger": [_I, _L, _L, _F, _P, _L, _P, _L, _P, _L],
    "orc_dger": [_I, _L, _L, _D, _P, _L, _P, _L, _P, _L],
    "orc_sgemm": [_I, _I, _I, _L, _L, _L, _F, _P, _L, _P, _L, _F, _P, _L],
    "orc_dgemm": [_I, _I, _I, _L, _L, _L, _D, _P, _L, _P, _L, _D, _P, _L],
    "orc_sgemm_ref64": [_I, _I, _I, _L, _L, _L, _F, _P, _L, _P, _L, _F, _P,
                        _L, _P],
    "orc_dgemm_dd_sample": [_I, _I, _I, _L, _L, _L, _D, _P, _L, _P, _L, _D, _P,
                            _L, _L, _P, _P, _P],
    "orc_sgemm_dd_sample": [_I, _I, _I, _L, _L, _L, _D, _P, _L, _P, _L, _D, _P,
                            _L, _L, _P, _P, _P],
    "orc_cgemm": [_I, _I, _I, _L, _L, _L, _F, _F, _P, _L, _P, _L, _F, _F, _P,
                  _L],
    "orc_zgemm": [_I, _I, _I, _L, _L, _L, _D, _D, _P, _L, _P, _L, _D, _D, _P,
                  _L],
    "orc_zgemm_dd_sample": [_I, _I, _I, _L, _L, _L, _D, _D, _P, _L, _P, _L, _D,
                            _D, _P, _L, _L, _P, _P, _P],
    "orc_cgemm_dd_sample": [_I, _I, _I, _L, _L, _L, _D, _D, _P, _L, _P, _L, _D,
                            _D, _P, _L, _L, _P, _P, _P],
}
_RET = {"orc_sdot": _F, "orc_ddot": _D, "orc_sasum": _F, "orc_dasum": _D,
        "orc_sdot_f64": _D}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built (make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RET.get(name, None)
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert isinstance(a, np.ndarray) and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


# -- generator ---------------------------------------------------------------
def uniform_f32(n, seed, start=0):
    out = np.empty(int(n), np.float32)
    lib().orc_fill_uniform_f32(_p(out), out.size, seed, start)
    return out


def uniform_f64(n, seed, start=0):
    out = np.empty(int(n), np.float64)
    lib().orc_fill_uniform_f64(_p(out), out.size, seed, start)
    return out


def wide_f32(n, seed, start=0, emax=20.0):
    out = np.empty(int(n), np.float32)
    lib().orc_fill_wide_f32(_p(out), out.size, seed, start, emax)
    return out


def wide_f64(n, seed, start=0, emax=20.0):
    out = np.empty(int(n), np.float64)
    lib().orc_fill_wide_f64(_p(out), out.size, seed, start, emax)
    return out


# -- level 1 -----------------------------------------------------------------
def axpy(alpha, x, y, incx=1, incy=1, n=None):
    """y <- fma(alpha, x, y) (netlib saxpy.f order; OpenBLAS contraction)."""
    y = y.copy()
    n = x.size if n is None else n
    f = lib().orc_saxpy if y.dtype == np.float32 else lib().orc_daxpy
    f(n, alpha, _p(x), incx, _p(y), incy)
    return y


def scal(alpha, x, incx=1, n=None):
    x = x.copy()
    n = x.size if n is None else n
    if x.dtype == np.float32:
        lib().orc_sscal(n, alpha, _p(x), incx)
    elif x.dtype == np.float64:
        lib().orc_dscal(n, alpha, _p(x), incx)
    elif x.dtype == np.complex64:
        v = x.view(np.float32)
        if isinstance(alpha, complex):
            lib().orc_cscal(n, alpha.real, alpha.imag, _p(v), incx)
        else:
            lib().orc_csscal(n, alpha, _p(v), incx)
    else:
        v = x.view(np.float64)
        if isinstance(alpha, complex):
            lib().orc_zscal(n, alpha.real, alpha.imag, _p(v), incx)
        else:
            lib().orc_zdscal(n, alpha, _p(v), incx)
    return x


def copy(x, y, incx=1, incy=1, n=None):
    y = y.copy()
    n = x.size if n is None else n
    f = lib().orc_scopy if y.dtype == np.float32 else lib().orc_dcopy
    f(n, _p(x), incx, _p(y), incy)
    return y


def dot(x, y, incx=1, incy=1, n=None):
    """netlib-order dot (sequential sum in the input precision)."""
    n = x.size if n is None else n
    f = lib().orc_sdot if x.dtype == np.float32 else lib().orc_ddot
    return f(n, _p(x), incx, _p(y), incy)


def asum(x, incx=1, n=None):
    n = x.size if n is None else n
    f = lib().orc_sasum if x.dtype == np.float32 else lib().orc_dasum
    return f(n, _p(x), incx)


def dot_exact(x, y, incx=1, incy=1, n=None):
    """Accuracy reference: double-double accumulation, rounded once."""
    n = x.size if n is None else n
    if x.dtype == np.float32:
        return lib().orc_sdot_f64(n, _p(x), incx, _p(y), incy)
    hi, lo = ctypes.c_double(), ctypes.c_double()
    lib().orc_ddot_dd(n, _p(x), incx, _p(y), incy, ctypes.byref(hi),
                      ctypes.byref(lo))
    return hi.value + lo.value


def asum_exact(x, incx=1, n=None):
    n = x.size if n is None else n
    hi, lo = ctypes.c_double(), ctypes.c_double()
    xd = np.ascontiguousarray(x, dtype=np.float64)
    lib().orc_dasum_dd(n, _p(xd), incx, ctypes.byref(hi), ctypes.byref(lo))
    return hi.value + lo.value


# -- level 2 -----------------------------------------------------------------
def gemv(order, trans, M, N, alpha, A, lda, x, incx, beta, y, incy):
    y = y.copy()
    f = lib().orc_sgemv if y.dtype == np.float32 else lib().orc_dgemv
    f(order, trans, M, N, alpha, _p(A), lda, _p(x), incx, beta, _p(y), incy)
    return y


def gemv_exact(order, trans, M, N, alpha, A, lda, x, incx, beta, y, incy):
    leny = (N if trans != OP_N else M)
    out = np.empty(max(leny, 0), np.float64)
    f = lib().orc_sgemv_ref if A.dtype == np.float32 else lib().orc_dgemv_ref
    f(order, trans, M, N, alpha, _p(A), lda, _p(x), incx, beta, _p(y), incy,
      _p(out))
    return out


def ger(order, M, N, alpha, x, incx, y, incy, A, lda):
    A = A.copy()
    f = lib().orc_sger if A.dtype == np.float32 else lib().orc_dger
    f(order, M, N, alpha, _p(x), incx, _p(y), incy, _p(A), lda)
    return A


# -- level 3 -----------------------------------------------------------------
def gemm(order, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc):
    """netlib-order GEMM; complex dtypes use cgemm.f semantics (OP_C
    conjugates)."""
    C = C.copy()
    if C.dtype == np.float32:
        lib().orc_sgemm(order, ta, tb, M, N, K, alpha, _p(A), lda, _p(B), ldb,
                        beta, _p(C), ldc)
    elif C.dtype == np.float64:
        lib().orc_dgemm(order, ta, tb, M, N, K, alpha, _p(A), lda, _p(B), ldb,
                        beta, _p(C), ldc)
    elif C.dtype == np.complex64:
        a, b = complex(alpha), complex(beta)
        lib().orc_cgemm(order, ta, tb, M, N, K, a.real, a.imag, _p(A), lda,
                        _p(B), ldb, b.real, b.imag, _p(C), ldc)
    elif C.dtype == np.complex128:
        a, b = complex(alpha), complex(beta)
        lib().orc_zgemm(order, ta, tb, M, N, K, a.real, a.imag, _p(A), lda,
                        _p(B), ldb, b.real, b.imag, _p(C), ldc)
    else:
        raise TypeError(C.dtype)
    return C


def sgemm_ref64(order, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc):
    """FP64 value of an fp32 GEMM, shaped like C's storage (M x N in the
    caller's order, densely packed: ld = M col-major / N row-major)."""
    out = np.empty(M * N, np.float64)
    lib().orc_sgemm_ref64(order, ta, tb, M, N, K, alpha, _p(A), lda, _p(B),
                          ldb, beta, _p(C), ldc, _p(out))
    return out


def gemm_sample_exact(order, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C,
                      ldc, rows, cols):
    """Double-double value of sampled entries (caller's (row, col) in the
    caller's order). Complex dtypes return complex128."""
    rows = np.ascontiguousarray(rows, np.int64)
    cols = np.ascontiguousarray(cols, np.int64)
    ns = rows.size
    if A.dtype in (np.float32, np.float64):
        out = np.empty(ns, np.float64)
        f = (lib().orc_sgemm_dd_sample if A.dtype == np.float32 else
             lib().orc_dgemm_dd_sample)
        f(order, ta, tb, M, N, K, float(alpha), _p(A), lda, _p(B), ldb,
          float(beta), _p(C), ldc, ns, _p(rows), _p(cols), _p(out))
        return out
    out = np.empty(2 * ns, np.float64)
    a, b = complex(alpha), complex(beta)
    f = (lib().orc_cgemm_dd_sample if A.dtype == np.complex64 else
         lib().orc_zgemm_dd_sample)
    f(order, ta, tb, M, N, K, a.real, a.imag, _p(A), lda, _p(B), ldb, b.real,
      b.imag, _p(C), ldc, ns, _p(rows), _p(cols), _p(out))
    return out[0::2] + 1j * out[1::2]


def _op_view(X, trans, rows, cols, ld, order):
    """numpy view of op(X) (rows x cols) for a flat storage array X."""
    if order == COL_MAJOR:
        if trans == OP_N:
            return X[: ld * (cols - 1) + rows].reshape(-1)[
                np.arange(rows)[:, None] + ld * np.arange(cols)[None, :]]
        return X[: ld * (rows - 1) + cols].reshape(-1)[
            np.arange(cols)[None, :] + ld * np.arange(rows)[:, None]]
    if trans == OP_N:
        return X[: ld * (rows - 1) + cols].reshape(-1)[
            ld * np.arange(rows)[:, None] + np.arange(cols)[None, :]]
    return X[: ld * (cols - 1) + rows].reshape(-1)[
        ld * np.arange(cols)[None, :] + np.arange(rows)[:, None]]


def hgemm(order, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc):
    """HGEMM with fp32-or-better accumulation: binary16 inputs, the exact
    product sum in float64, one rounding to binary16. (Reference HGEMM
    accumulates in half, hgemm_NN_kernels.cpp:1372; its tests are compiled
    out, hcblas_hgemm.cpp:34.) Pinned in tests/test_oracle_pins.py against
    float16(sgemm_ref64(...)) for every order and transpose with padded
    leading dimensions, exact rationals, and a case where half accumulation
    differs."""
    a = _op_view(A, ta, M, K, lda, order).astype(np.float64)
    b = _op_view(B, tb, K, N, ldb, order).astype(np.float64)
    r = float(alpha) * (a @ b)
    C = C.copy()
    if order == COL_MAJOR:
        idx = np.arange(M)[:, None] + ldc * np.arange(N)[None, :]
    else:
        idx = ldc * np.arange(M)[:, None] + np.arange(N)[None, :]
    if beta != 0:
        r = r + float(beta) * C[idx].astype(np.float64)
    C[idx] = r.astype(np.float16)
    return C
