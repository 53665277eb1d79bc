"""Column-major SGEMM over the shape corpus of the reference's
convolution-network benchmark (benchmark/BLAS_benchmark_Convolution_Networks/
*.txt, driven through test/src/sgemm_cn_test.cpp:42-123: column-major,
transA / transB, lda / ldb / ldc, alpha, beta and aOffset / bOffset / cOffset
added to the operand pointers), from tests/golden/convnet_shapes.json.

Every distinct shape of the four network lists (AlexNet, GoogLeNet, OverFeat,
Oxford: 155 shapes) and every 12th shape of the rectangular / square sweeps
with M*N*K <= 2^30 run with the corpus's largest offsets for that shape.
Inputs are integers in [0, 100) like the reference's (sgemm_cn_test.cpp:104,
rand_r % 100), so while K * 99^2 + 99 < 2^24 every result is an exact FP32
integer and must equal the double-double oracle exactly; past that the
bound is the plain kernel's c*K*eps. 256 sampled entries per shape.
"""
import json
import os

import numpy as np
import pytest

from gpu_util import to_dev, to_host

COL = 1
HERE = os.path.dirname(os.path.abspath(__file__))
NETS = ("AlexnetBlasDimensions", "GoogleBlasDimensions",
        "OverfeatBlasDimensions", "OxfordBlasDimensions")


def _shapes():
    with open(os.path.join(HERE, "golden", "convnet_shapes.json")) as f:
        return json.load(f)["shapes"]


def _select(group):
    rows = _shapes()
    if group == "nets":
        return [r for r in rows if r[13] in NETS]
    sweep = [r for r in rows if r[13] not in NETS and r[0] * r[1] * r[2] <= 2 ** 30]
    return sweep[::12]


def _extent(r):
    M, N, K, ta, tb, lda, ldb, ldc = r[:8]
    a = lda * (M if ta else K) + r[10]
    b = ldb * (K if tb else N) + r[11]
    c = ldc * N + r[12]
    return a, b, c


def test_corpus_fixture_matches_reference_lists():
    rows = _shapes()
    assert len(rows) == 2757
    assert sum(r[13] in NETS for r in rows) == 155
    # every row is a valid column-major call (ld >= rows of the stored operand)
    for r in rows:
        M, N, K, ta, tb, lda, ldb, ldc = r[:8]
        assert lda >= (K if ta else M) and ldb >= (N if tb else K) and ldc >= M


@pytest.mark.gpu
@pytest.mark.parametrize("group", ["nets", "sweep"])
def test_convnet_shapes(hc, handle, orc, group):
    api = hc.api
    rng = np.random.default_rng(11)
    sel = _select(group)
    done = 0
    for r in sel:
        M, N, K, ta, tb, lda, ldb, ldc = r[:8]
        alpha, beta = r[8], r[9]
        ao, bo, co = r[10:13]
        na, nb, nc = _extent(r)
        if na + nb + nc > 96 * 2 ** 20:
            continue
        A = np.zeros(na, np.float32)
        B = np.zeros(nb, np.float32)
        C = np.zeros(nc, np.float32)
        A[ao:] = rng.integers(0, 100, na - ao)
        B[bo:] = rng.integers(0, 100, nb - bo)
        C[co:] = rng.integers(0, 100, nc - co)
        dC = to_dev(C)
        st = api.hcblas_sgemm(handle, COL, ta, tb, M, N, K, alpha, to_dev(A),
                              lda, to_dev(B), ldb, beta, dC, ldc, ao, bo, co)
        assert st == 0, r
        got = to_host(dC)
        rows = rng.integers(0, M, 256)
        cols = rng.integers(0, N, 256)
        ref = orc.gemm_sample_exact(COL, ta, tb, M, N, K, alpha, A[ao:], lda,
                                    B[bo:], ldb, beta, C[co:], ldc, rows, cols)
        val = got[co + rows + cols * ldc].astype(np.float64)
        if K * 99 * 99 + 99 < 2 ** 24:
            assert np.array_equal(val, ref), r
        else:
            tol = (K + 2) * 2.0 ** -23 * (abs(alpha) * K * 99 * 99 + abs(beta) * 99)
            assert np.max(np.abs(val - ref)) <= tol, r
        # entries before the C offset are untouched
        assert not np.any(got[:co]), r
        done += 1
    print(f"{group}: {done} of {len(sel)} shapes checked")
    assert done >= 0.9 * len(sel)
