"""GEMM parity on the GPU through the C ABI (S/D/C/Z/H, plain, pointer-array
and strided batches, explicit order and offsets).

Reference tests mirrored: test/unit/hcblas_{s,d,c,z}gemm.cpp (the 189x9x19
order x trans sweep on integer data, exact vs CBLAS), test/unit-api/
test-blas.cpp (123x78x23 closed form, pointer-array batches of 32), plus
seeded real-valued problems across tile boundaries. Real-valued results must
lie within c*K*eps*(|alpha|*sum|a||b| + |beta||c|) of the FP64 /
double-double value (c = 2).
"""
import numpy as np
import pytest

from gemm_bounds import sample_bound
from gemm_bounds import check_sampled
from gpu_util import bitwise_equal, to_dev, to_host

pytestmark = pytest.mark.gpu
ROW, COL = 0, 1
N_, T_, C_ = 0, 1, 2
DT = {"s": np.float32, "d": np.float64, "c": np.complex64, "z": np.complex128}


def _run(hc, handle, p, order, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta,
         C, ldc, aoff=0, boff=0, coff=0):
    api = hc.api
    dA, dB, dC = to_dev(A), to_dev(B), to_dev(C)
    if p in "sd":
        st = getattr(api, f"hcblas_{p}gemm")(handle, order, ta, tb, M, N, K,
                                             alpha, dA, lda, dB, ldb, beta, dC,
                                             ldc, aoff, boff, coff)
    else:
        st = getattr(api, f"hcblas_{p}gemm")(handle, order, ta, tb, M, N, K,
                                             alpha, dA, aoff, lda, dB, boff,
                                             ldb, beta, dC, coff, ldc)
    assert st == 0
    return to_host(dC)


@pytest.mark.parametrize("p", ["s", "d", "c", "z"])
def test_reference_recipe_bit_exact(hc, handle, golden, p):
    g = golden("gemm_ref_int.npz")
    M, N, K = (int(v) for v in g["dims"])
    ncases = 8 if p in "sd" else 11
    ab = 1.0 if p in "sd" else 1 + 1j
    for i in range(ncases):
        o, ta, tb, lda, ldb, ldc = (int(v) for v in g[f"{p}_case{i}"])
        got = _run(hc, handle, p, o, ta, tb, M, N, K, ab, g[f"{p}_A"], lda,
                   g[f"{p}_B"], ldb, ab, g[f"{p}_C"], ldc)
        assert bitwise_equal(got, g[f"{p}_expect{i}"]), f"case {i}"


@pytest.mark.parametrize("p", ["s", "d", "c", "z"])
def test_real_valued_golden(hc, handle, golden, orc, p):
    g = golden("real_valued.npz")
    M, N, K = (int(v) for v in g["dims"])
    a, b = g[f"{p}_ab"]
    if p in "sd":
        a, b = float(a.real), float(b.real)
    else:
        a, b = complex(a), complex(b)
    eps = np.finfo(DT[p]).eps
    for i in range(8):
        o, ta, tb, lda, ldb, ldc = (int(v) for v in g[f"{p}_case{i}"])
        got = _run(hc, handle, p, o, ta, tb, M, N, K, a, g[f"{p}_A"], lda,
                   g[f"{p}_B"], ldb, b, g[f"{p}_C"], ldc)
        exp = g[f"{p}_expect{i}"]
        assert np.max(np.abs(got - exp)) <= 4 * K * eps * 4, f"case {i}"


def test_closed_forms(hc, handle):
    # test-blas.cpp:1861-1920 (A=B=C=1, alpha=beta=2, 123x78x23 -> 2K+2)
    M, N, K = 123, 78, 23
    for p in "sdcz":
        dt = DT[p]
        for o in (COL, ROW):
            lda, ldb, ldc = (M, K, M) if o == COL else (K, N, N)
            got = _run(hc, handle, p, o, N_, N_, M, N, K, 2, np.ones(M * K, dt),
                       lda, np.ones(K * N, dt), ldb, 2, np.ones(M * N, dt), ldc)
            assert np.all(got == 2 * K + 2), (p, o)
    # test-hipblas.cpp:1815-1853 (A=1, B=2, C=3, alpha=beta=1 -> 2K+3)
    api = hc.api
    for o in (COL, ROW):
        lda, ldb, ldc = (M, K, M) if o == COL else (K, N, N)
        dC = to_dev(np.full(M * N, 3, np.uint16) * 0 + hc.half_bits(3.0))
        st = api.hcblas_hgemm(handle, o, N_, N_, M, N, K, 1.0,
                              to_dev(np.full(M * K, hc.half_bits(1.0), np.uint16)),
                              lda,
                              to_dev(np.full(K * N, hc.half_bits(2.0), np.uint16)),
                              ldb, 1.0, dC, ldc, 0, 0, 0)
        assert st == 0
        got = to_host(dC).view(np.uint16).view(np.float16)
        assert np.all(got == 2 * K + 3)


def _rand(orc, p, n, seed):
    if p == "s":
        return orc.uniform_f32(n, seed)
    if p == "d":
        return orc.uniform_f64(n, seed)
    gen = orc.uniform_f32 if p == "c" else orc.uniform_f64
    v = gen(2 * n, seed)
    return (v[0::2] + 1j * v[1::2]).astype(DT[p])


def _check_against_exact(orc, p, o, ta, tb, M, N, K, alpha, A, lda, B, ldb,
                         beta, C, ldc, got, nsamp=512):
    """Componentwise bound c*sqrt(K)*u*(|alpha| |op(A)||op(B)| + |beta||C|)
    against the double-double value (tests/gemm_bounds.py; c = 4 real,
    8 complex)."""
    rng = np.random.default_rng(M * 131 + N * 7 + K)
    ns = min(nsamp, M * N)
    rows = rng.integers(0, M, ns)
    cols = rng.integers(0, N, ns)
    u = np.finfo(DT[p]).eps
    check_sampled(orc, o, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C,
                  ldc, got, rows, cols, u, c=4.0 if p in "sd" else 8.0)


SHAPES = [(1, 1, 1), (5, 3, 7), (16, 16, 16), (64, 64, 64), (65, 63, 67),
          (128, 128, 128), (129, 250, 33), (300, 17, 515), (512, 384, 96),
          (1000, 1000, 8)]


@pytest.mark.parametrize("p", ["s", "d", "c", "z"])
@pytest.mark.parametrize("M,N,K", SHAPES)
def test_seeded_shapes_all_cases(hc, handle, orc, p, M, N, K):
    transes = (N_, T_) if p in "sd" else (N_, T_, C_)
    alpha = 0.75 if p in "sd" else 0.75 - 0.5j
    beta = -1.25 if p in "sd" else -1.25 + 0.25j
    seed = M * 1000 + N * 10 + K
    for o in (COL, ROW):
        for ta in transes:
            for tb in transes:
                if p in "cz" and (ta, tb) not in ((N_, N_), (T_, C_), (C_, N_),
                                                  (C_, C_), (N_, T_)):
                    continue
                ar, ac = (M, K) if ta == N_ else (K, M)
                br, bc = (K, N) if tb == N_ else (N, K)
                # stored rows (leading dimension covers them), plus padding
                lda = (ar if o == COL else ac) + 2
                ldb = (br if o == COL else bc) + 1
                ldc = (M if o == COL else N) + 3
                A = _rand(orc, p, lda * (ac if o == COL else ar), seed + 1)
                B = _rand(orc, p, ldb * (bc if o == COL else br), seed + 2)
                C = _rand(orc, p, ldc * (N if o == COL else M), seed + 3)
                got = _run(hc, handle, p, o, ta, tb, M, N, K, alpha, A, lda, B,
                           ldb, beta, C, ldc)
                _check_against_exact(orc, p, o, ta, tb, M, N, K, alpha, A, lda,
                                     B, ldb, beta, C, ldc, got)


@pytest.mark.parametrize("p", ["s", "d", "c", "z"])
def test_offsets_and_special_scalars(hc, handle, orc, p):
    M, N, K = 70, 45, 33
    dt = DT[p]
    A = _rand(orc, p, M * K + 5, 1)
    B = _rand(orc, p, K * N + 7, 2)
    C = _rand(orc, p, M * N + 3, 3)
    # offsets
    got = _run(hc, handle, p, COL, N_, N_, M, N, K, 1.0, A, M, B, K, 0.5, C, M,
               aoff=5, boff=7, coff=3)
    _check_against_exact(orc, p, COL, N_, N_, M, N, K, 1.0, A[5:], M, B[7:], K,
                         0.5, C[3:], M, got[3:])
    assert bitwise_equal(got[:3], C[:3])
    # beta == 0: NaN in C must not propagate
    Cn = np.full(M * N, np.nan, dt)
    got = _run(hc, handle, p, COL, N_, N_, M, N, K, 1.0, A, M, B, K, 0.0, Cn, M)
    assert np.all(np.isfinite(got))
    # alpha == 0: C = beta*C exactly (A, B not read)
    got = _run(hc, handle, p, COL, N_, N_, M, N, K, 0.0, A, M, B, K, 2.0,
               C[:M * N], M)
    assert bitwise_equal(got, (C[:M * N] * dt(2)).astype(dt))
    # K == 0: C = beta*C
    got = _run(hc, handle, p, COL, N_, N_, M, N, 0, 1.0, A, M, B, 1, 0.5,
               C[:M * N], M)
    assert bitwise_equal(got, (C[:M * N] * dt(0.5)).astype(dt))
    # M == 0 / N == 0: quick return, C untouched
    got = _run(hc, handle, p, COL, N_, N_, 0, N, K, 1.0, A, 1, B, K, 0.5,
               C[:M * N], 1)
    assert bitwise_equal(got, C[:M * N])


def test_sgemm_config1_rowmajor_1024(hc, handle, orc):
    """Row-major NN 1024^3, A and B uniform[-1,1] from seed 0, C = 0,
    alpha = 1, beta = 0 (the plain-SGEMM configuration)."""
    M = N = K = 1024
    A = orc.uniform_f32(M * K, 0, 0)
    B = orc.uniform_f32(K * N, 0, M * K)
    C = np.zeros(M * N, np.float32)
    got = _run(hc, handle, "s", ROW, N_, N_, M, N, K, 1.0, A, K, B, N, 0.0, C,
               N)
    _check_against_exact(orc, "s", ROW, N_, N_, M, N, K, 1.0, A, K, B, N, 0.0,
                         C, N, got, nsamp=4096)
    # relative L2 (test/include/helper_functions.h:10-53, threshold 1e-5)
    rows = np.arange(0, M, 7)
    ref = orc.sgemm_ref64(ROW, N_, N_, 64, N, K, 1.0, A, K, B, N, 0.0, C, N)
    g64 = got[:64 * N].astype(np.float64)
    assert np.linalg.norm(g64 - ref) / np.linalg.norm(ref) <= 1e-5
    assert rows.size


@pytest.mark.parametrize("order", [COL, ROW])
def test_hgemm_fp32_accumulate(hc, handle, orc, order):
    api = hc.api
    M, N, K = 131, 77, 300
    rng = np.random.default_rng(5)
    A = rng.uniform(-1, 1, M * K).astype(np.float16)
    B = rng.uniform(-1, 1, K * N).astype(np.float16)
    C = rng.uniform(-1, 1, M * N).astype(np.float16)
    for ta in (N_, T_):
        for tb in (N_, T_):
            lda = (M if ta == N_ else K) if order == COL else (K if ta == N_ else M)
            ldb = (K if tb == N_ else N) if order == COL else (N if tb == N_ else K)
            ldc = M if order == COL else N
            dC = to_dev(C.view(np.uint16))
            st = api.hcblas_hgemm(handle, order, ta, tb, M, N, K, 1.0,
                                  to_dev(A.view(np.uint16)), lda,
                                  to_dev(B.view(np.uint16)), ldb, 0.5, dC, ldc,
                                  0, 0, 0)
            assert st == 0
            got = to_host(dC).view(np.uint16).view(np.float16).astype(np.float64)
            ref = orc.hgemm(order, ta, tb, M, N, K, 1.0, A, lda, B, ldb, 0.5, C,
                            ldc).astype(np.float64)
            # fp32 accumulation then one rounding to half: within one half ulp
            # of the exactly rounded value plus the fp32 accumulation error.
            ulp = np.spacing(np.abs(ref).astype(np.float16)).astype(np.float64)
            assert np.all(np.abs(got - ref) <= ulp + K * 2 ** -23 * 2), (ta, tb)


def test_hgemm_c_api(hc, handle):
    api = hc.api
    M, N, K = 64, 32, 48
    A = np.full(M * K, hc.half_bits(0.5), np.uint16)
    B = np.full(K * N, hc.half_bits(2.0), np.uint16)
    C = np.full(M * N, hc.half_bits(1.0), np.uint16)
    dC = to_dev(C)
    st = api.hcblasHgemm(handle, N_, N_, M, N, K, 1.0, to_dev(A), M, to_dev(B),
                         K, 2.0, dC, M)
    assert st == 0
    got = to_host(dC).view(np.uint16).view(np.float16)
    assert np.all(got == K + 2)


@pytest.mark.parametrize("p", ["s", "d", "c", "z"])
def test_pointer_array_batched(hc, handle, orc, torch_cuda, p):
    """test-blas.cpp:1922-2060: device array of device pointers, batch 32."""
    torch = torch_cuda
    api = hc.api
    M, N, K, b = 123, 78, 23, 32
    As = [_rand(orc, p, M * K, 10 + i) for i in range(b)]
    Bs = [_rand(orc, p, K * N, 100 + i) for i in range(b)]
    Cs = [_rand(orc, p, M * N, 1000 + i) for i in range(b)]
    dA = [to_dev(a) for a in As]
    dB = [to_dev(x) for x in Bs]
    dC = [to_dev(c) for c in Cs]
    pa = torch.tensor([t.data_ptr() for t in dA], dtype=torch.int64, device="cuda")
    pb = torch.tensor([t.data_ptr() for t in dB], dtype=torch.int64, device="cuda")
    pc = torch.tensor([t.data_ptr() for t in dC], dtype=torch.int64, device="cuda")
    alpha = 1.0 if p in "sd" else 1 - 0.5j
    beta = 1.0 if p in "sd" else 0.5 + 0j
    fn = getattr(api, f"hcblas{p.upper()}gemmBatched")
    st = fn(handle, N_, T_, M, N, K, alpha, pa, M, pb, N, beta, pc, M, b)
    assert st == 0
    for i in range(b):
        got = to_host(dC[i])
        _check_against_exact(orc, p, COL, N_, T_, M, N, K, alpha, As[i], M,
                             Bs[i], N, beta, Cs[i], M, got, nsamp=128)


@pytest.mark.parametrize("p", ["s", "d", "c", "z", "h"])
def test_pointer_array_batch_element_offsets(hc, handle, orc, torch_cuda, p):
    """Pointer arrays whose entries point at odd element offsets inside one
    arena (bases not 16-byte aligned, leading dimensions and the common
    offset aligned): the kernel checks each problem's bases and takes the
    element-wise loaders for the misaligned ones. Sizes large enough for the
    16-byte path (every tile shape's main loop) in both trans cases."""
    torch = torch_cuda
    api = hc.api
    M, N, K, b = 192, 160, 256, 6
    dt = {"s": np.float32, "d": np.float64, "c": np.complex64,
          "z": np.complex128, "h": np.float16}[p]
    esz = np.dtype(dt).itemsize
    for ta in (N_, T_):
        lda = M if ta == N_ else K
        sa, sb, sc = lda * (K if ta == N_ else M), K * N, M * N
        # problem i starts at element offset off[i] of its arena (0 keeps
        # one aligned entry beside the misaligned ones)
        off = [0, 1, 3, 5, 2, 7][:b]
        span_a = sum(sa + o for o in off) + 8
        span_b = sum(sb + o for o in off) + 8
        span_c = sum(sc + o for o in off) + 8
        if p == "h":
            rng = np.random.default_rng(9 + ta)
            Ah = rng.uniform(-1, 1, span_a).astype(np.float16)
            Bh = rng.uniform(-1, 1, span_b).astype(np.float16)
            Ch = rng.uniform(-1, 1, span_c).astype(np.float16)
        else:
            Ah = _rand(orc, p, span_a, 21 + ta)
            Bh = _rand(orc, p, span_b, 22 + ta)
            Ch = _rand(orc, p, span_c, 23 + ta)
        dA = to_dev(Ah.view(np.uint16) if p == "h" else Ah)
        dB = to_dev(Bh.view(np.uint16) if p == "h" else Bh)
        dC = to_dev(Ch.view(np.uint16) if p == "h" else Ch)
        starts_a, starts_b, starts_c = [], [], []
        ca = cb = cc = 0
        for o in off:
            starts_a.append(ca + o)
            starts_b.append(cb + o)
            starts_c.append(cc + o)
            ca += sa + o
            cb += sb + o
            cc += sc + o
        pa = torch.tensor([dA.data_ptr() + s_ * esz for s_ in starts_a],
                          dtype=torch.int64, device="cuda")
        pb = torch.tensor([dB.data_ptr() + s_ * esz for s_ in starts_b],
                          dtype=torch.int64, device="cuda")
        pc = torch.tensor([dC.data_ptr() + s_ * esz for s_ in starts_c],
                          dtype=torch.int64, device="cuda")
        alpha = 0.75 if p in "sdh" else 0.75 - 0.5j
        beta = 0.5 if p in "sdh" else 0.5 + 0.25j
        fn = getattr(api, f"hcblas{p.upper()}gemmBatched")
        assert fn(handle, ta, N_, M, N, K, alpha, pa, lda, pb, K, beta, pc, M,
                  b) == 0
        got = to_host(dC)
        if p == "h":
            got = got.view(np.uint16).view(np.float16)
        for i in range(b):
            Ai = Ah[starts_a[i]:starts_a[i] + sa]
            Bi = Bh[starts_b[i]:starts_b[i] + sb]
            Ci = Ch[starts_c[i]:starts_c[i] + sc]
            gi = got[starts_c[i]:starts_c[i] + sc]
            if p == "h":
                ref = orc.hgemm(COL, ta, N_, M, N, K, alpha, Ai, lda, Bi, K,
                                beta, Ci, M).astype(np.float64)
                ulp = np.spacing(np.abs(ref).astype(np.float16)).astype(
                    np.float64)
                assert np.all(np.abs(gi.astype(np.float64) - ref) <=
                              ulp + K * 2.0 ** -23 * 2), (ta, i)
            else:
                _check_against_exact(orc, p, COL, ta, N_, M, N, K, alpha, Ai,
                                     lda, Bi, K, beta, Ci, M, gi, nsamp=256)
        # untouched outside the M x N blocks (gaps between the problems)
        mask = np.ones(span_c, bool)
        for st_ in starts_c:
            mask[st_:st_ + sc] = False
        assert bitwise_equal(got[mask], Ch[mask])


@pytest.mark.parametrize("p", ["s", "d", "c", "z"])
def test_strided_batched_row_major(hc, handle, orc, p):
    """Row-major strided batches (the reference silently skipped row-major
    batched GEMM, sgemm_array_wrapper.cpp:113-151)."""
    api = hc.api
    M, N, K, b = 40, 36, 50, 5
    dtc = {"s": 0, "d": 1, "c": 2, "z": 3}[p]
    sA, sB, sC = M * K + 4, K * N + 8, M * N + 12
    A = _rand(orc, p, sA * b, 1)
    B = _rand(orc, p, sB * b, 2)
    C = _rand(orc, p, sC * b, 3)
    dC = to_dev(C)
    alpha = 1.25 if p in "sd" else 1.25 + 1j
    st = api.hcblas_gemm_strided_batched(handle, dtc, ROW, T_, N_, M, N, K,
                                         alpha, to_dev(A), M, 2, sA, to_dev(B),
                                         N, 4, sB, 0.0, dC, N, 6, sC, b)
    assert st == 0
    got = to_host(dC)
    for i in range(b):
        _check_against_exact(orc, p, ROW, T_, N_, M, N, K, alpha,
                             A[i * sA + 2:], M, B[i * sB + 4:], N, 0.0,
                             C[i * sC + 6:], N, got[i * sC + 6:], nsamp=128)


def test_bf16_strided_batched_fp32_out(hc, handle, torch_cuda):
    """bf16 inputs, fp32 accumulation and fp32 output (the differentiation
    configuration's forward shape, reduced batch)."""
    torch = torch_cuda
    api = hc.api
    b, M, N, K = 4, 256, 1024, 512
    g = torch.Generator(device="cuda").manual_seed(3)
    A = (torch.rand(b, M, K, generator=g, device="cuda") * 2 - 1).bfloat16()
    B = (torch.rand(b, K, N, generator=g, device="cuda") * 2 - 1).bfloat16()
    C = torch.empty(b, M, N, device="cuda", dtype=torch.float32)
    st = api.hcblas_gemm_strided_batched(handle, 5, ROW, N_, N_, M, N, K, 1.0,
                                         A, K, 0, M * K, B, N, 0, K * N, 0.0,
                                         C, N, 0, M * N, b)
    assert st == 0
    torch.cuda.synchronize()
    ref = torch.bmm(A.double(), B.double().contiguous())
    err = (C.double() - ref).abs().max().item()
    assert err <= K * 2 ** -23 * 4, err


# Sizes of the reference's generators (test/unit/test_constants.h:18-64,
# glibc srand(1)): vsmall = 48, large = 8378, vlarge = 14522. Shapes of
# test/unit/hcblas_sgemm.cpp: slim A (M = N = vlarge, K = vsmall,
# :1190-1313), slim C (M = K = vlarge, N = vsmall, :1315-1437) and the large
# square sizes of :962-1188, column-major, all four trans cases. The
# comparison is the reference's relative L2 error over sampled entries
# (test/include/helper_functions.h:10-53, 1e-5 in FP32) against the
# double-double value, plus the element-wise c*K*eps bound.
def _sampled_check(hc, handle, orc, torch, p, ta, tb, M, N, K, seed):
    api = hc.api
    dt = DT[p]
    ar, ac = (M, K) if ta == N_ else (K, M)
    br, bc = (K, N) if tb == N_ else (N, K)
    A = _rand(orc, p, ar * ac, seed + 1)
    B = _rand(orc, p, br * bc, seed + 2)
    big_c = M * N > 50_000_000
    alpha = 0.75
    beta = 0.0 if big_c else -1.25
    C = np.zeros(1, dt) if big_c else _rand(orc, p, M * N, seed + 3)
    dC = (torch.empty(M * N, dtype=torch.float32 if p == "s" else torch.float64,
                      device="cuda") if big_c else to_dev(C))
    st = getattr(api, f"hcblas_{p}gemm")(handle, COL, ta, tb, M, N, K, alpha,
                                         to_dev(A), ar, to_dev(B), br, beta,
                                         dC, M, 0, 0, 0)
    assert st == 0
    rng = np.random.default_rng(seed)
    ns = 512
    rows = rng.integers(0, M, ns)
    cols = rng.integers(0, N, ns)
    idx = torch.as_tensor(rows + cols * M, device="cuda")
    got = to_host(dC[idx]).astype(np.float64)
    ref = orc.gemm_sample_exact(COL, ta, tb, M, N, K, alpha, A, ar, B, br,
                                beta, C, M, rows, cols)
    rel_l2 = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert rel_l2 <= (1e-5 if p == "s" else 1e-13), rel_l2
    bound = sample_bound(orc, COL, ta, tb, M, N, K, alpha, A, ar, B, br, beta,
                         C, M, rows, cols, ref, np.finfo(dt).eps)
    assert np.all(np.abs(got - ref) <= bound)


@pytest.mark.parametrize("p", ["s", "d"])
@pytest.mark.parametrize("M,N,K", [(14522, 14522, 48), (14522, 48, 14522)])
def test_reference_slim_shapes(hc, handle, orc, torch_cuda, p, M, N, K):
    for ta in (N_, T_):
        for tb in (N_, T_):
            _sampled_check(hc, handle, orc, torch_cuda, p, ta, tb, M, N, K,
                           seed=M + 3 * N + 7 * K + 11 * ta + 13 * tb)
    torch_cuda.cuda.empty_cache()


@pytest.mark.parametrize("ta,tb", [(N_, N_), (N_, T_), (T_, N_), (T_, T_)])
def test_reference_large_square(hc, handle, orc, torch_cuda, ta, tb):
    n = 8378
    _sampled_check(hc, handle, orc, torch_cuda, "s", ta, tb, n, n, n,
                   seed=n + 11 * ta + 13 * tb)
    torch_cuda.cuda.empty_cache()


# Long-K problems with few tiles take the split-K form (partial tiles over K
# slices, then an in-order reduction with alpha / beta): all trans cases, S
# and D, an odd shape and one whose K is not a multiple of the slice.
@pytest.mark.parametrize("p", ["s", "d"])
@pytest.mark.parametrize("M,N,K", [(256, 128, 16384), (100, 72, 9000)])
def test_split_k_shapes(hc, handle, orc, torch_cuda, p, M, N, K):
    for ta in (N_, T_):
        for tb in (N_, T_):
            _sampled_check(hc, handle, orc, torch_cuda, p, ta, tb, M, N, K,
                           seed=M + 5 * N + 3 * K + 17 * ta + 19 * tb)


@pytest.mark.parametrize("order", [COL, ROW])
def test_hgemm_split_k(hc, handle, orc, order):
    """HGEMM with K = 8192 on one tile takes the split-K form (FP32
    partials, then one rounding to half in the reduction)."""
    api = hc.api
    M, N, K = 72, 40, 8200
    rng = np.random.default_rng(8)
    A = rng.uniform(-1, 1, M * K).astype(np.float16)
    B = rng.uniform(-1, 1, K * N).astype(np.float16)
    C = rng.uniform(-1, 1, M * N).astype(np.float16)
    for ta in (N_, T_):
        for tb in (N_, T_):
            lda = (M if ta == N_ else K) if order == COL else (K if ta == N_ else M)
            ldb = (K if tb == N_ else N) if order == COL else (N if tb == N_ else K)
            ldc = M if order == COL else N
            dC = to_dev(C.view(np.uint16))
            st = api.hcblas_hgemm(handle, order, ta, tb, M, N, K, 0.75,
                                  to_dev(A.view(np.uint16)), lda,
                                  to_dev(B.view(np.uint16)), ldb, -0.5, dC,
                                  ldc, 0, 0, 0)
            assert st == 0
            got = to_host(dC).view(np.uint16).view(np.float16).astype(np.float64)
            ref = orc.hgemm(order, ta, tb, M, N, K, 0.75, A, lda, B, ldb, -0.5,
                            C, ldc).astype(np.float64)
            ulp = np.spacing(np.abs(ref).astype(np.float16)).astype(np.float64)
            assert np.all(np.abs(got - ref) <= ulp + K * 2 ** -23 * 2), (ta, tb)


def test_bf16_split_k_single_problem(hc, handle, torch_cuda):
    """A single BF16 problem with a long K and few tiles: split-K partials in
    FP32, reduced in slice order (alpha, beta applied once)."""
    torch = torch_cuda
    api = hc.api
    M, N, K = 192, 128, 16384
    g = torch.Generator(device="cuda").manual_seed(4)
    A = (torch.rand(M, K, generator=g, device="cuda") * 2 - 1).bfloat16()
    B = (torch.rand(K, N, generator=g, device="cuda") * 2 - 1).bfloat16()
    C0 = torch.rand(M, N, generator=g, device="cuda") * 2 - 1
    C = C0.clone()
    st = api.hcblas_gemm_strided_batched(handle, 5, ROW, N_, N_, M, N, K, 1.5,
                                         A, K, 0, M * K, B, N, 0, K * N, 0.25,
                                         C, N, 0, M * N, 1)
    assert st == 0
    torch.cuda.synchronize()
    ref = 1.5 * (A.double() @ B.double()) + 0.25 * C0.double()
    err = (C.double() - ref).abs().max().item()
    assert err <= K * 2 ** -23 * 4, err


# Shapes large enough for the biggest tiles (256x256 FP32, 256x128 FP64),
# and 2048^2 shapes for the next ones (128x128 FP32 on 4 waves, 128x64 FP64
# on 8 waves; K = 200 ends in a partial K step), all of which issue a
# step's LDS writes and loads among its MFMAs; beta != 0, so the
# LDS-staged epilogue reads C.
@pytest.mark.parametrize("p,M,N,K", [("s", 4096, 4096, 256), ("d", 4096, 2048, 64),
                                     ("s", 2048, 2048, 200), ("d", 2048, 2048, 200)])
def test_large_tiles_with_beta(hc, handle, orc, torch_cuda, p, M, N, K):
    for ta in (N_, T_):
        for tb in (N_, T_):
            _sampled_check(hc, handle, orc, torch_cuda, p, ta, tb, M, N, K,
                           seed=M + 7 * N + K + 23 * ta + 29 * tb)


@pytest.mark.parametrize("M,N,K", [(4096, 4096, 512), (2048, 2048, 328)])
@pytest.mark.parametrize("ta,tb", [(N_, N_), (N_, T_), (T_, N_), (T_, T_)])
def test_hgemm_large_tiles_with_beta(hc, handle, torch_cuda, M, N, K, ta, tb):
    """HGEMM on 256x256 (4096^2) and 128x128 (2048^2, K with a partial last
    step) tiles, FP16 in and out
    with FP32 accumulation, every transpose form (the r-contiguous operands
    are transposed in registers on their way to LDS), beta != 0: sampled
    entries against the FP64 value rounded to half."""
    torch = torch_cuda
    api = hc.api
    g = torch.Generator(device="cuda").manual_seed(12 + 2 * ta + tb)
    A = (torch.rand(M * K, generator=g, device="cuda") * 2 - 1).half()
    B = (torch.rand(K * N, generator=g, device="cuda") * 2 - 1).half()
    C0 = (torch.rand(M * N, generator=g, device="cuda") * 2 - 1).half()
    C = C0.clone()
    lda = M if ta == N_ else K
    ldb = K if tb == N_ else N
    st = api.hcblas_hgemm(handle, COL, ta, tb, M, N, K, 0.75, A.data_ptr(),
                          lda, B.data_ptr(), ldb, 0.5, C.data_ptr(), M, 0, 0,
                          0)
    assert st == 0
    torch.cuda.synchronize()
    rng = np.random.default_rng(3)
    rows = torch.as_tensor(rng.integers(0, M, 1024), device="cuda")
    cols = torch.as_tensor(rng.integers(0, N, 1024), device="cuda")
    # op(A)(m, k) and op(B)(k, n) as dense FP64 views of the stored arrays
    Am = (A.view(K, M).t() if ta == N_ else A.view(M, K)).double()
    Bm = (B.view(N, K).t() if tb == N_ else B.view(K, N)).double()
    ref = 0.75 * (Am[rows] * Bm[:, cols].t()).sum(1) + \
        0.5 * C0.view(N, M).t().double()[rows, cols]
    got = C.view(N, M).t()[rows, cols].double().cpu().numpy()
    ref = ref.cpu().numpy()
    ulp = np.spacing(np.abs(ref).astype(np.float16)).astype(np.float64)
    assert np.all(np.abs(got - ref) <= ulp + K * 2 ** -23 * 2)
