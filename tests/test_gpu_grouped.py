"""Grouped GEMM on the GPU (no reference counterpart; the reference's batched
GEMM takes one shape for the whole batch).

Checks on the grouped configuration's problem generator (log-uniform sizes,
mixed fp32 / fp16, random storage order):
  - the persistent, graph-captured launch is bit-identical to per-problem
    hcblas_sgemm / hcblas_hgemm launches (same tile code and K step), for
    every autotuned schedule;
  - graph replays, interleaved with eager work, reproduce the same bits;
  - sampled entries agree with the oracle (double-double for fp32, exactly
    rounded binary16 of the FP64 product for fp16);
  - invalid tables are rejected with INVALID_VALUE.
"""
import numpy as np
import pytest

import bench

pytestmark = pytest.mark.gpu


def test_grouped_matches_plain_launches_bitwise(hc, handle, torch_cuda):
    torch = torch_cuda
    problems, meta, arenas = bench.grouped_problem_table(torch, hc, 192, seed=7)
    plan = hc.GroupedPlan(handle, problems)
    try:
        assert plan.run() == 0
        torch.cuda.synchronize()
        got = [t.clone() for t in arenas]
        ref = (arenas[0].clone(), arenas[1].clone())
        bench.run_plain_group(hc, handle, meta, arenas, ref)
        torch.cuda.synchronize()
        # compare only the C regions (A / B regions are the inputs)
        for (d, o, m, n, k, ao, bo, co, lda, ldb, ldc) in meta:
            g = got[d][co:co + m * n]
            r = ref[d][co:co + m * n]
            assert torch.equal(g.view(torch.int32 if d == 0 else torch.int16),
                               r.view(torch.int32 if d == 0 else torch.int16)), (d, m, n, k)
        # every schedule the autotuner may pick gives the same bits
        assert plan.autotune(1) == 0
        info = plan.info()
        assert info["tile"] in (32, 64) and info["tiles"] > 0
        assert plan.run() == 0
        torch.cuda.synchronize()
        for (d, o, m, n, k, ao, bo, co, lda, ldb, ldc) in meta:
            assert torch.equal(arenas[d][co:co + m * n], ref[d][co:co + m * n])
    finally:
        plan.destroy()


def test_graph_replays_with_eager_work_between(hc, handle, torch_cuda):
    torch = torch_cuda
    problems, meta, arenas = bench.grouped_problem_table(torch, hc, 96, seed=11)
    plan = hc.GroupedPlan(handle, problems)
    try:
        assert plan.run() == 0
        torch.cuda.synchronize()
        first = [t.clone() for t in arenas]
        junk = torch.rand(1 << 20, device="cuda")
        for i in range(4):
            junk = junk * 1.0001 + 0.5  # eager work on the same stream
            torch.cuda.synchronize()
            # scribble over the outputs so a stale replay cannot pass
            for (d, o, m, n, k, ao, bo, co, *_r) in meta:
                arenas[d][co:co + m * n].fill_(7.0)
            assert plan.run() == 0
            torch.cuda.synchronize()
            for (d, o, m, n, k, ao, bo, co, *_r) in meta:
                assert torch.equal(arenas[d][co:co + m * n],
                                   first[d][co:co + m * n]), (i, d, m, n, k)
    finally:
        plan.destroy()


def test_grouped_against_oracle(hc, handle, torch_cuda, orc):
    torch = torch_cuda
    problems, meta, arenas = bench.grouped_problem_table(torch, hc, 48, seed=5)
    plan = hc.GroupedPlan(handle, problems)
    try:
        assert plan.run() == 0
        torch.cuda.synchronize()
        a32 = arenas[0].cpu().numpy()
        a16 = arenas[1].cpu().numpy()
        rng = np.random.default_rng(0)
        for (d, o, m, n, k, ao, bo, co, lda, ldb, ldc) in meta:
            A = (a32 if d == 0 else a16)[ao:ao + m * k]
            B = (a32 if d == 0 else a16)[bo:bo + k * n]
            C = (a32 if d == 0 else a16)[co:co + m * n]
            rows = rng.integers(0, m, 32)
            cols = rng.integers(0, n, 32)
            idx = rows + cols * ldc if o == 1 else rows * ldc + cols
            if d == 0:
                ref = orc.gemm_sample_exact(o, 0, 0, m, n, k, 1.0, A.copy(), lda,
                                            B.copy(), ldb, 0.0, np.zeros_like(C),
                                            ldc, rows, cols)
                tol = 2 * (k + 2) * 2.0 ** -23 * (k + 1)
                assert np.max(np.abs(C[idx] - ref)) <= tol
            else:
                full = orc.hgemm(o, 0, 0, m, n, k, 1.0, A.copy(), lda, B.copy(),
                                 ldb, 0.0, np.zeros_like(C), ldc)
                ulp = np.spacing(np.abs(full[idx]).astype(np.float16)).astype(np.float64)
                err = np.abs(C[idx].astype(np.float64) - full[idx].astype(np.float64))
                assert np.all(err <= ulp + k * 2.0 ** -23 * 2)
    finally:
        plan.destroy()


def test_grouped_rejects_invalid(hc, handle, torch_cuda):
    import ctypes
    lib = hc.load()
    x = torch_cuda.zeros(64, device="cuda")
    good = hc.GemmProblem(hc.R_32F, hc.COL_MAJOR, 0, 0, 4, 4, 4, 0, 1.0, 0.0,
                          x.data_ptr(), 4, x.data_ptr(), 4, x.data_ptr(), 4)
    for field, val in (("dtype", 9), ("lda", 2), ("M", -1), ("transA", 2)):
        bad = hc.GemmProblem.from_buffer_copy(good)
        setattr(bad, field, val)
        table = (hc.GemmProblem * 1)(bad)
        p = ctypes.c_void_p()
        assert lib.hcblasGroupedPlanCreate(handle.value, table, 1,
                                           ctypes.byref(p)) == 3, field
    assert lib.hcblasGroupedGemm(handle.value, None) == 3
    assert lib.hcblasGroupedPlanDestroy(None) == 3
    # an empty group is a valid no-op
    p = ctypes.c_void_p()
    assert lib.hcblasGroupedPlanCreate(handle.value, None, 0, ctypes.byref(p)) == 0
    assert lib.hcblasGroupedGemm(handle.value, p) == 0
    assert lib.hcblasGroupedPlanDestroy(p) == 0


def _irregular_table(torch, hc, seed, aligned):
    """Problems with odd sizes, both transposes, padded leading dimensions,
    alpha != 1 and beta != 0 (C is read), in both orders and dtypes. aligned:
    leading dimensions rounded up to multiples of 8 and 256-byte aligned
    matrices; otherwise odd padding and matrices at odd element offsets."""
    rng = np.random.default_rng(seed)
    sizes = [1, 5, 8, 17, 33, 64, 70, 130]
    meta, used = [], [0, 0]
    for i in range(64):
        d = int(rng.integers(0, 2))
        o = int(rng.integers(0, 2))
        ta, tb = int(rng.integers(0, 2)), int(rng.integers(0, 2))
        m, n, k = (int(rng.choice(sizes)) for _ in range(3))
        alpha = float(rng.choice([1.0, -0.75]))
        beta = float(rng.choice([0.0, 0.5]))
        col = o == 1

        def stored(r, c):
            ld = r if col else c
            # aligned: ld a multiple of 8 (vector accesses of C with a
            # partial last vector when the size is odd); else odd padding
            ld = (ld + 7) // 8 * 8 if aligned else ld + int(rng.choice([0, 1, 3]))
            return ld, ld * (c if col else r)

        lda, na = stored(m, k) if ta == 0 else stored(k, m)
        ldb, nb = stored(k, n) if tb == 0 else stored(n, k)
        ldc, nc = stored(m, n)
        offs = []
        for n_el in (na, nb, nc):
            skew = int(rng.integers(0, 3)) if not aligned else 0
            offs.append(used[d] + skew)
            used[d] += (n_el + skew + 63) // 64 * 64
        meta.append((d, o, ta, tb, m, n, k, alpha, beta, offs, lda, ldb, ldc))
    g = torch.Generator(device="cuda").manual_seed(seed)
    ar32 = torch.rand(used[0], generator=g, device="cuda") * 2 - 1
    ar16 = (torch.rand(used[1], generator=g, device="cuda") * 2 - 1).half()
    arenas = (ar32, ar16)
    problems = []
    for (d, o, ta, tb, m, n, k, alpha, beta, offs, lda, ldb, ldc) in meta:
        base, es = arenas[d].data_ptr(), (4 if d == 0 else 2)
        problems.append(hc.GemmProblem(
            hc.R_32F if d == 0 else hc.R_16F, o, ta, tb, m, n, k, 0, alpha,
            beta, base + offs[0] * es, lda, base + offs[1] * es, ldb,
            base + offs[2] * es, ldc))
    return problems, meta, arenas


@pytest.mark.parametrize("aligned", [True, False])
def test_grouped_irregular_problems_bitwise(hc, handle, torch_cuda, aligned):
    """Odd sizes, transposes, beta != 0 and unaligned C take the epilogue's
    partial-vector and element paths: still bit-identical to per-problem
    plain launches."""
    torch = torch_cuda
    api = hc.api
    problems, meta, arenas = _irregular_table(torch, hc, 11 + aligned, aligned)
    init = [t.clone() for t in arenas]
    plan = hc.GroupedPlan(handle, problems)
    try:
        for tune in (False, True):
            for d in (0, 1):
                arenas[d].copy_(init[d])
            if tune:
                assert plan.autotune(1) == 0
                for d in (0, 1):
                    arenas[d].copy_(init[d])
            assert plan.run() == 0
            torch.cuda.synchronize()
            got = [t.clone() for t in arenas]
            ref = [t.clone() for t in init]
            for (d, o, ta, tb, m, n, k, alpha, beta, offs, lda, ldb, ldc) in meta:
                es = 4 if d == 0 else 2
                fn = api.hcblas_sgemm if d == 0 else api.hcblas_hgemm
                st = fn(handle, o, ta, tb, m, n, k, alpha,
                        ref[d].data_ptr() + offs[0] * es, lda,
                        ref[d].data_ptr() + offs[1] * es, ldb, beta,
                        ref[d].data_ptr() + offs[2] * es, ldc, 0, 0, 0)
                assert st == 0
            torch.cuda.synchronize()
            for d in (0, 1):
                iv = torch.int32 if d == 0 else torch.int16
                assert torch.equal(got[d].view(iv), ref[d].view(iv)), (d, tune)
    finally:
        plan.destroy()
