"""Componentwise error bound for sampled GEMM results, shared by the GEMM
tests.

For sampled entries (i, j) of C = alpha * op(A) op(B) + beta * C0 the bound
is

    |C_ij - C_ref_ij| <= c * sqrt(K) * u * (|alpha| (|op(A)| |op(B)|)_ij
                                            + |beta| |C0_ij|) + 2 u |C_ref_ij|

with u the unit roundoff of the result type and C_ref the double-double
value. Rounding errors of a K-term dot product in any summation order are
bounded by K u (|a| |b|); for inputs whose rounding errors do not all align,
they grow as sqrt(K) u, which is what this bound allows (c = 4 by default).
A kernel that rounded its inputs to BF16 (u = 2^-8) misses it by orders of
magnitude, which tests/test_tolerance.py checks on the CPU.
"""
import numpy as np

N_, T_, C_ = 0, 1, 2


def abs_product(orc, o, ta, tb, M, N, K, A, lda, B, ldb, rows, cols):
    """(|op(A)| |op(B)|)_ij on the sampled entries (complex: moduli)."""
    Aa = np.ascontiguousarray(np.abs(A), dtype=np.float64)
    Ba = np.ascontiguousarray(np.abs(B), dtype=np.float64)
    t = (lambda x: T_ if x == C_ else x)
    return orc.gemm_sample_exact(o, t(ta), t(tb), M, N, K, 1.0, Aa, lda, Ba,
                                 ldb, 0.0, np.zeros(1), 1, rows, cols)


def sample_bound(orc, o, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                 rows, cols, ref, u, c=4.0):
    P = abs_product(orc, o, ta, tb, M, N, K, A, lda, B, ldb, rows, cols)
    idx = rows + cols * ldc if o == 1 else rows * ldc + cols
    c0 = np.abs(C[idx]).astype(np.float64) if beta != 0 else 0.0
    return (c * np.sqrt(max(K, 1)) * u * (abs(alpha) * P + abs(beta) * c0) +
            2 * u * np.abs(ref))


def check_sampled(orc, o, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C,
                  ldc, got, rows, cols, u, c=4.0):
    """Asserts the bound on the sampled entries of `got` (flat storage of C
    with leading dimension ldc in order o). Returns max(err / bound)."""
    ref = orc.gemm_sample_exact(o, ta, tb, M, N, K, alpha, A, lda, B, ldb,
                                beta, C, ldc, rows, cols)
    idx = rows + cols * ldc if o == 1 else rows * ldc + cols
    g = got[idx]
    assert np.all(np.isfinite(g))
    err = np.abs(g - ref)
    bound = sample_bound(orc, o, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta,
                         C, ldc, rows, cols, ref, u, c)
    ratio = float(np.max(err / np.maximum(bound, np.finfo(np.float64).tiny)))
    assert ratio <= 1.0, (ratio, float(err.max()))
    return ratio
