"""The componentwise GEMM bound of tests/gemm_bounds.py discriminates: an
FP32 GEMM accumulated in FP32 (the oracle's netlib order) passes it, the same
GEMM on BF16-rounded inputs fails it, for the shapes the GPU tests use."""
import numpy as np
import pytest

from gemm_bounds import check_sampled

COL, ROW = 1, 0
N_, T_ = 0, 1


def _bf16(x):
    b = np.ascontiguousarray(x, np.float32).view(np.uint32)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32)


@pytest.mark.parametrize("M,N,K", [(65, 63, 67), (300, 17, 515),
                                   (64, 64, 1024)])
def test_bound_accepts_fp32_rejects_bf16_inputs(orc, M, N, K):
    A = orc.uniform_f32(M * K, 1)
    B = orc.uniform_f32(K * N, 2)
    C = orc.uniform_f32(M * N, 3)
    rng = np.random.default_rng(0)
    rows = rng.integers(0, M, 256)
    cols = rng.integers(0, N, 256)
    u = np.finfo(np.float32).eps
    good = orc.gemm(COL, N_, N_, M, N, K, 0.75, A, M, B, K, -1.25, C, M)
    check_sampled(orc, COL, N_, N_, M, N, K, 0.75, A, M, B, K, -1.25, C, M,
                  good, rows, cols, u)
    bad = orc.gemm(COL, N_, N_, M, N, K, 0.75, _bf16(A), M, _bf16(B), K,
                   -1.25, C, M)
    with pytest.raises(AssertionError):
        check_sampled(orc, COL, N_, N_, M, N, K, 0.75, A, M, B, K, -1.25, C,
                      M, bad, rows, cols, u)


def test_bound_rejects_bf16_complex(orc):
    M, N, K = 40, 30, 256
    rng = np.random.default_rng(4)
    A = (rng.uniform(-1, 1, M * K) + 1j * rng.uniform(-1, 1, M * K)).astype(
        np.complex64)
    B = (rng.uniform(-1, 1, K * N) + 1j * rng.uniform(-1, 1, K * N)).astype(
        np.complex64)
    C = np.zeros(M * N, np.complex64)
    rows = rng.integers(0, M, 128)
    cols = rng.integers(0, N, 128)
    u = np.finfo(np.float32).eps
    good = orc.gemm(COL, N_, N_, M, N, K, 1 + 0j, A, M, B, K, 0j, C, M)
    check_sampled(orc, COL, N_, N_, M, N, K, 1 + 0j, A, M, B, K, 0j, C, M,
                  good, rows, cols, u)

    def bfc(z):
        return (_bf16(z.real) + 1j * _bf16(z.imag)).astype(np.complex64)
    bad = orc.gemm(COL, N_, N_, M, N, K, 1 + 0j, bfc(A), M, bfc(B), K, 0j, C,
                   M)
    with pytest.raises(AssertionError):
        check_sampled(orc, COL, N_, N_, M, N, K, 1 + 0j, A, M, B, K, 0j, C, M,
                      bad, rows, cols, u)
