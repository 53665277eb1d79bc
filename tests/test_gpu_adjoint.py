"""Adjoint kernels on the GPU (no reference counterpart).

For the loss L = sum(C o W) of C = alpha*op(A)*op(B) + beta*C_in the
cotangent is dC = W, and the adjoints must equal
    dA = alpha * W * op(B)^T,  dB = alpha * op(A)^T * W,  dC_in = beta * W
(and for A_out = alpha*x*y^T + A_in: dx = alpha*W*y, dy = alpha*W^T*x).

Checks, through the C ABI and through the libtorch autograd layer:
  - every order x trans case against the FP64 gradient written from the
    definition above (numpy matrix products of op views);
  - FP64 central finite differences of L on sampled entries;
  - torch.autograd.gradcheck (FP64 finite differences) on the op;
  - the differentiation configuration: 256 x (256 x 1024 x 512) bf16
    batched GEMM with fp32 accumulation and a 4096^2 SGER.
"""
import ctypes

import numpy as np
import pytest

from gpu_util import bitwise_equal, to_dev, to_host

pytestmark = pytest.mark.gpu
ROW, COL = 0, 1
N_, T_ = 0, 1


def _storage(order, rows, cols, ld, batch, stride):
    """Flat size of a strided batch of rows x cols matrices."""
    foot = ld * (cols - 1) + rows if order == COL else ld * (rows - 1) + cols
    return (batch - 1) * stride + foot


def _mat(orc, X, order, trans, rows, cols, ld):
    return orc._op_view(X, trans, rows, cols, ld, order)


def _ref_adjoints(orc, order, ta, tb, M, N, K, alpha, beta, A, lda, sA, B,
                  ldb, sB, W, ldw, sW, batch):
    """FP64 adjoints per batch entry from the definition, as logical
    matrices: dA in A's storage shape (rows_a x cols_a), dB likewise."""
    out = []
    ra, ca = (K, M) if ta else (M, K)
    rb, cb = (N, K) if tb else (K, N)
    for b in range(batch):
        opA = _mat(orc, A[b * sA:].astype(np.float64), order, ta, M, K, lda)
        opB = _mat(orc, B[b * sB:].astype(np.float64), order, tb, K, N, ldb)
        Wm = _mat(orc, W[b * sW:].astype(np.float64), order, N_, M, N, ldw)
        dopA = alpha * Wm @ opB.T        # M x K
        dopB = alpha * opA.T @ Wm        # K x N
        dA = dopA.T if ta else dopA      # storage shape ra x ca
        dB = dopB.T if tb else dopB
        assert dA.shape == (ra, ca) and dB.shape == (rb, cb)
        out.append((dA, dB, beta * Wm))
    return out


def _run_backward(hc, handle, dtype, order, ta, tb, M, N, K, alpha, beta, A,
                  lda, sA, B, ldb, sB, W, ldw, sW, batch, ldda, sdA, lddb,
                  sdB, ldc_in, sC_in, out_dtype, grad_dtype=None):
    """dA, dB (grad_dtype, default out_dtype; BF16 bits as uint16) and
    dC_in (out_dtype) from the C ABI."""
    lib = hc.load()
    gd = out_dtype if grad_dtype is None else grad_dtype
    dA = np.zeros(_storage(order, *((K, M) if ta else (M, K)), ldda, batch, sdA), gd)
    dB = np.zeros(_storage(order, *((N, K) if tb else (K, N)), lddb, batch, sdB), gd)
    dCin = np.zeros(_storage(order, M, N, ldc_in, batch, sC_in), out_dtype)
    gA, gB, gC = to_dev(dA), to_dev(dB), to_dev(dCin)
    tA, tB, tW = to_dev(A), to_dev(B), to_dev(W)
    ct = ctypes.c_double if out_dtype == np.float64 else ctypes.c_float
    a, bt = ct(alpha), ct(beta)
    st = lib.hcblas_gemm_strided_batched_backward(
        handle.value, dtype, order, ta, tb, M, N, K, ctypes.byref(a),
        ctypes.byref(bt), ctypes.c_void_p(tA.data_ptr()), ctypes.c_int64(lda),
        ctypes.c_int64(sA), ctypes.c_void_p(tB.data_ptr()), ctypes.c_int64(ldb),
        ctypes.c_int64(sB), ctypes.c_void_p(tW.data_ptr()), ctypes.c_int64(ldw),
        ctypes.c_int64(sW), ctypes.c_void_p(gA.data_ptr()),
        ctypes.c_int64(ldda), ctypes.c_int64(sdA),
        ctypes.c_void_p(gB.data_ptr()), ctypes.c_int64(lddb),
        ctypes.c_int64(sdB), ctypes.c_void_p(gC.data_ptr()),
        ctypes.c_int64(ldc_in), ctypes.c_int64(sC_in), batch)
    assert st == 0
    return to_host(gA), to_host(gB), to_host(gC)


def _bf16_bits(x):
    """Round-to-nearest-even BF16 bit patterns (uint16) of float32 x."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def _bf16_value(b):
    return (b.astype(np.uint32) << 16).view(np.float32)


@pytest.mark.parametrize("p,shape", [("s", "odd"), ("d", "odd"),
                                     ("bf", "odd"), ("bf", "aligned")])
@pytest.mark.parametrize("order", [COL, ROW])
@pytest.mark.parametrize("ta", [N_, T_])
@pytest.mark.parametrize("tb", [N_, T_])
def test_gemm_adjoints_all_cases(hc, handle, orc, p, shape, order, ta, tb):
    """dA, dB, dC_in of the strided batched GEMM in every order x trans case
    with padded leading dimensions and batch strides; "bf": BF16 A and B
    (the mixed kernel: FP32 dC split into BF16 planes, operands stored the
    transposed way copied k-contiguous first), BF16 adjoints. Componentwise
    bound 4 sqrt(n) u |alpha| (|dC| |op(B)^T|) + 2 u |ref| (n the inner
    length, u the unit roundoff of the FP32 / FP64 result), plus for BF16
    adjoints one rounding to nearest: 2^-8 |ref| (a truncating conversion,
    up to 2^-7 |ref|, fails it)."""
    if shape == "odd":
        M, N, K, batch, pad = 37, 29, 45, 3, (3, 1, 2)
    else:
        M, N, K, batch, pad = 96, 64, 160, 2, (8, 8, 8)
    alpha, beta = 0.75, -1.25
    dt = np.float64 if p == "d" else np.float32
    gen = orc.uniform_f64 if p == "d" else orc.uniform_f32
    ra, ca = (K, M) if ta else (M, K)
    rb, cb = (N, K) if tb else (K, N)
    lda = (ra if order == COL else ca) + pad[0]
    ldb = (rb if order == COL else cb) + pad[1]
    ldw = (M if order == COL else N) + pad[2]
    al = 8 if shape == "aligned" else 1
    sA = _storage(order, ra, ca, lda, 1, 0) + (5 if al == 1 else 8)
    sB = _storage(order, rb, cb, ldb, 1, 0) + (7 if al == 1 else 16)
    sW = _storage(order, M, N, ldw, 1, 0) + (1 if al == 1 else 8)
    A = gen(_storage(order, ra, ca, lda, batch, sA), 1)
    B = gen(_storage(order, rb, cb, ldb, batch, sB), 2)
    W = gen(_storage(order, M, N, ldw, batch, sW), 3)
    dtype = {"s": 0, "d": 1, "bf": 5}[p]
    if p == "bf":
        Ab, Bb = _bf16_bits(A), _bf16_bits(B)
        A, B = _bf16_value(Ab), _bf16_value(Bb)  # the values the GPU sees
        dev_A, dev_B = Ab, Bb
    else:
        dev_A, dev_B = A, B
    ldda, lddb, ldci = lda + al, ldb + 2 * al, ldw + al
    sdA = _storage(order, ra, ca, ldda, 1, 0) + 2 * al
    sdB = _storage(order, rb, cb, lddb, 1, 0)
    sCi = _storage(order, M, N, ldci, 1, 0) + 3 * al
    gA, gB, gC = _run_backward(hc, handle, dtype, order, ta, tb, M, N, K,
                               alpha, beta, dev_A, lda, sA, dev_B, ldb, sB, W,
                               ldw, sW, batch, ldda, sdA, lddb, sdB, ldci, sCi,
                               dt, np.uint16 if p == "bf" else None)
    if p == "bf":
        gA, gB = _bf16_value(gA), _bf16_value(gB)
    ur = 2.0 ** -8 if p == "bf" else 0.0  # output rounding of the adjoints
    refs = _ref_adjoints(orc, order, ta, tb, M, N, K, alpha, beta, A, lda, sA,
                         B, ldb, sB, W, ldw, sW, batch)
    absr = _ref_adjoints(orc, order, ta, tb, M, N, K, abs(alpha), 0.0,
                         np.abs(A), lda, sA, np.abs(B), ldb, sB, np.abs(W),
                         ldw, sW, batch)
    u = np.finfo(dt).eps / 2
    for b, ((rA, rB, rC), (pA, pB, _)) in enumerate(zip(refs, absr)):
        gotA = _mat(orc, gA[b * sdA:], order, N_, ra, ca, ldda)
        gotB = _mat(orc, gB[b * sdB:], order, N_, rb, cb, lddb)
        gotC = _mat(orc, gC[b * sCi:], order, N_, M, N, ldci)
        bA = 4 * np.sqrt(N) * u * pA + 2 * u * np.abs(rA)
        bB = 4 * np.sqrt(M) * u * pB + 2 * u * np.abs(rB)
        assert np.all(np.abs(gotA - rA) <= bA * (1 + ur) + ur * np.abs(rA)), b
        assert np.all(np.abs(gotB - rB) <= bB * (1 + ur) + ur * np.abs(rB)), b
        # dC_in = beta * W is one rounding of the product: exact match
        Wm = _mat(orc, W[b * sW:], order, N_, M, N, ldw)
        assert bitwise_equal(gotC.astype(dt), (dt(beta) * Wm).astype(dt)), b


def test_gemm_adjoint_central_differences(hc, handle, orc):
    """FP64 central differences of L(A) and L(B) on sampled entries agree
    with the FP64 adjoints (L is bilinear, so the difference quotient is
    exact up to rounding)."""
    M, N, K, batch = 24, 31, 17, 2
    alpha, beta = 1.5, 0.5
    order, ta, tb = ROW, T_, N_
    A = orc.uniform_f64(batch * K * M, 11)
    B = orc.uniform_f64(batch * K * N, 12)
    W = orc.uniform_f64(batch * M * N, 13)
    C0 = orc.uniform_f64(batch * M * N, 14)
    gA, gB, gC = _run_backward(hc, handle, 1, order, ta, tb, M, N, K, alpha,
                               beta, A, M, K * M, B, N, K * N, W, N, M * N,
                               batch, M, K * M, N, K * N, N, M * N,
                               np.float64)

    def loss(Av, Bv, Cv):
        tot = 0.0
        for b in range(batch):
            a = Av[b * K * M:(b + 1) * K * M].reshape(K, M)  # op(A) = a^T
            bb = Bv[b * K * N:(b + 1) * K * N].reshape(K, N)
            c = alpha * a.T @ bb + beta * Cv[b * M * N:(b + 1) * M * N].reshape(M, N)
            tot += float(np.sum(c * W[b * M * N:(b + 1) * M * N].reshape(M, N)))
        return tot

    rng = np.random.default_rng(0)
    h = 2.0 ** -12
    for X, G in ((A, gA), (B, gB), (C0, gC)):
        for i in rng.integers(0, X.size, 12):
            Xp, Xm = X.copy(), X.copy()
            Xp[i] += h
            Xm[i] -= h
            args_p = [A, B, C0]
            args_m = [A, B, C0]
            k = [id(A), id(B), id(C0)].index(id(X))
            args_p[k], args_m[k] = Xp, Xm
            fd = (loss(*args_p) - loss(*args_m)) / (2 * h)
            assert abs(fd - G[i]) <= 1e-9 * max(1.0, abs(fd)), (k, i, fd, G[i])


def test_bf16_config_adjoints(hc, handle, orc, torch_cuda):
    """Differentiation configuration: 256 x (M=256, N=1024, K=512) bf16
    batched GEMM (row-major NN, fp32 accumulation), W uniform[-1,1]. Sampled
    rows of dA / dB (bf16, rounded once) against the FP64 gradient of the
    bf16 inputs; dC_in exact."""
    torch = torch_cuda
    batch, M, N, K = 256, 256, 1024, 512
    alpha, beta = 1.0, 0.5
    A32 = torch.from_numpy(orc.uniform_f32(batch * M * K, 21)).cuda()
    B32 = torch.from_numpy(orc.uniform_f32(batch * K * N, 22)).cuda()
    W = torch.from_numpy(orc.uniform_f32(batch * M * N, 23)).cuda()
    A = A32.to(torch.bfloat16)
    B = B32.to(torch.bfloat16)
    dA = torch.empty(batch * M * K, device="cuda", dtype=torch.bfloat16)
    dB = torch.empty(batch * K * N, device="cuda", dtype=torch.bfloat16)
    dCi = torch.empty(batch * M * N, device="cuda")
    lib = hc.load()
    a, bt = ctypes.c_float(alpha), ctypes.c_float(beta)
    st = lib.hcblas_gemm_strided_batched_backward(
        handle.value, 5, ROW, N_, N_, M, N, K, ctypes.byref(a), ctypes.byref(bt),
        ctypes.c_void_p(A.data_ptr()), ctypes.c_int64(K), ctypes.c_int64(M * K),
        ctypes.c_void_p(B.data_ptr()), ctypes.c_int64(N), ctypes.c_int64(K * N),
        ctypes.c_void_p(W.data_ptr()), ctypes.c_int64(N), ctypes.c_int64(M * N),
        ctypes.c_void_p(dA.data_ptr()), ctypes.c_int64(K), ctypes.c_int64(M * K),
        ctypes.c_void_p(dB.data_ptr()), ctypes.c_int64(N), ctypes.c_int64(K * N),
        ctypes.c_void_p(dCi.data_ptr()), ctypes.c_int64(N), ctypes.c_int64(M * N),
        batch)
    assert st == 0
    torch.cuda.synchronize()
    rng = np.random.default_rng(1)
    Ah = A.float().cpu().numpy().astype(np.float64).reshape(batch, M, K)
    Bh = B.float().cpu().numpy().astype(np.float64).reshape(batch, K, N)
    Wh = W.cpu().numpy().astype(np.float64).reshape(batch, M, N)
    gA = dA.float().cpu().numpy().astype(np.float64).reshape(batch, M, K)
    gB = dB.float().cpu().numpy().astype(np.float64).reshape(batch, K, N)
    # componentwise FP32 bound 4 sqrt(n) u |alpha| (|W| |B^T|) + 2 u |ref|,
    # then one rounding to BF16 (2^-8 relative)
    u, ub = 2.0 ** -24, 2.0 ** -8
    for b in rng.integers(0, batch, 8):
        i = rng.integers(0, M)
        ref = alpha * Wh[b, i, :] @ Bh[b].T          # row i of dA[b]
        absp = abs(alpha) * np.abs(Wh[b, i, :]) @ np.abs(Bh[b]).T
        bound = 4 * np.sqrt(N) * u * absp + 2 * u * np.abs(ref)
        assert np.all(np.abs(gA[b, i] - ref) <= bound * (1 + ub) + ub * np.abs(ref))
        j = rng.integers(0, N)
        ref = alpha * Ah[b].T @ Wh[b, :, j]          # column j of dB[b]
        absp = abs(alpha) * np.abs(Ah[b]).T @ np.abs(Wh[b, :, j])
        bound = 4 * np.sqrt(M) * u * absp + 2 * u * np.abs(ref)
        assert np.all(np.abs(gB[b, :, j] - ref) <= bound * (1 + ub) + ub * np.abs(ref))
    assert bitwise_equal(dCi.cpu().numpy(), (np.float32(beta) * W.cpu().numpy()))


@pytest.mark.parametrize("p", ["s", "d"])
@pytest.mark.parametrize("order", [COL, ROW])
def test_ger_adjoints(hc, handle, orc, p, order):
    M, N = 300, 211
    alpha = -0.75
    dt = np.float32 if p == "s" else np.float64
    gen = orc.uniform_f32 if p == "s" else orc.uniform_f64
    x = gen(2 * M, 1)   # incx = 2
    y = gen(3 * N, 2)   # incy = 3
    ldw = (M if order == COL else N) + 5
    W = gen(_storage(order, M, N, ldw, 1, 0), 3)
    lib = hc.load()
    dx, dy = to_dev(np.zeros(M, dt)), to_dev(np.zeros(N, dt))
    fn = lib.hcblas_sger_backward if p == "s" else lib.hcblas_dger_backward
    ct = ctypes.c_float if p == "s" else ctypes.c_double
    a = ct(alpha)
    tx, ty, tW = to_dev(x), to_dev(y), to_dev(W)
    st = fn(handle.value, order, M, N, ctypes.byref(a),
            ctypes.c_void_p(tx.data_ptr()), ctypes.c_int64(2),
            ctypes.c_void_p(ty.data_ptr()), ctypes.c_int64(3),
            ctypes.c_void_p(tW.data_ptr()), ctypes.c_int64(ldw),
            ctypes.c_void_p(dx.data_ptr()), ctypes.c_void_p(dy.data_ptr()))
    assert st == 0
    Wm = _mat(orc, W.astype(np.float64), order, N_, M, N, ldw)
    rx = alpha * Wm @ y[::3].astype(np.float64)
    ry = alpha * Wm.T @ x[::2].astype(np.float64)
    eps = np.finfo(dt).eps
    assert np.max(np.abs(to_host(dx) - rx)) <= 2 * (N + 2) * eps * N
    assert np.max(np.abs(to_host(dy) - ry)) <= 2 * (M + 2) * eps * M


def test_torch_autograd_gradcheck(torch_cuda):
    """torch.autograd.gradcheck (FP64 central differences) on both ops."""
    torch = torch_cuda
    from hcblas_amd import torch_ext
    g = torch.Generator(device="cuda").manual_seed(5)
    for ta, tb in ((False, False), (True, False), (False, True), (True, True)):
        A = torch.rand((2, 7, 5) if ta else (2, 5, 7), generator=g, device="cuda",
                       dtype=torch.float64, requires_grad=True)
        B = torch.rand((2, 6, 7) if tb else (2, 7, 6), generator=g, device="cuda",
                       dtype=torch.float64, requires_grad=True)
        C = torch.rand((2, 5, 6), generator=g, device="cuda", dtype=torch.float64,
                       requires_grad=True)
        assert torch.autograd.gradcheck(
            lambda a, b, c: torch_ext.gemm_batched(a, b, c, 0.75, -0.5, ta, tb),
            (A, B, C), eps=1e-6, atol=1e-9)
    x = torch.rand(9, generator=g, device="cuda", dtype=torch.float64, requires_grad=True)
    y = torch.rand(4, generator=g, device="cuda", dtype=torch.float64, requires_grad=True)
    A = torch.rand((9, 4), generator=g, device="cuda", dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(lambda a, b, c: torch_ext.ger(a, b, c, 1.25),
                                    (x, y, A), eps=1e-6, atol=1e-9)


def test_torch_autograd_config(torch_cuda, orc):
    """The differentiation configuration through torch.autograd: bf16 batched
    GEMM (gradients returned in bf16, the inputs' type) and SGER 4096^2, loss
    sum(C o W)."""
    torch = torch_cuda
    from hcblas_amd import torch_ext
    batch, M, N, K = 256, 256, 1024, 512
    A = torch.from_numpy(orc.uniform_f32(batch * M * K, 31)).cuda().to(torch.bfloat16)
    B = torch.from_numpy(orc.uniform_f32(batch * K * N, 32)).cuda().to(torch.bfloat16)
    W = torch.from_numpy(orc.uniform_f32(batch * M * N, 33)).cuda().view(batch, M, N)
    A = A.view(batch, M, K).requires_grad_()
    B = B.view(batch, K, N).requires_grad_()
    C = torch_ext.gemm_batched(A, B)
    assert C.dtype == torch.float32
    ref = torch.bmm(A.detach().float(), B.detach().float())
    assert torch.max(torch.abs(C.detach() - ref)).item() <= 2 * (K + 2) * 2.0 ** -23 * K
    (C * W).sum().backward()
    assert A.grad.dtype == torch.bfloat16 and B.grad.dtype == torch.bfloat16
    refA = torch.bmm(W, B.detach().float().transpose(1, 2).contiguous())
    refB = torch.bmm(A.detach().float().transpose(1, 2).contiguous(), W)
    # one bf16 rounding of an fp32-accurate adjoint: relative 2^-8
    assert torch.all(torch.abs(A.grad.float() - refA) <= 2.0 ** -8 * torch.abs(refA) + 1e-3)
    assert torch.all(torch.abs(B.grad.float() - refB) <= 2.0 ** -8 * torch.abs(refB) + 1e-3)

    n = 4096
    x = torch.from_numpy(orc.uniform_f32(n, 34)).cuda().requires_grad_()
    y = torch.from_numpy(orc.uniform_f32(n, 35)).cuda().requires_grad_()
    A0 = torch.from_numpy(orc.uniform_f32(n * n, 36)).cuda().view(n, n).requires_grad_()
    W2 = torch.from_numpy(orc.uniform_f32(n * n, 37)).cuda().view(n, n)
    out = torch_ext.ger(x, y, A0, 1.5)
    assert torch.equal(out.detach(), torch.addr(A0.detach(), x.detach(), y.detach(), alpha=1.5)) or \
        torch.max(torch.abs(out.detach() - torch.addr(A0.detach(), x.detach(), y.detach(), alpha=1.5))).item() <= 4 * 2.0 ** -23
    (out * W2).sum().backward()
    Wd = W2.double()
    rx = 1.5 * Wd @ y.detach().double()
    ry = 1.5 * Wd.T @ x.detach().double()
    assert torch.max(torch.abs(x.grad.double() - rx)).item() <= 2 * (n + 2) * 2.0 ** -23 * n
    assert torch.max(torch.abs(y.grad.double() - ry)).item() <= 2 * (n + 2) * 2.0 ** -23 * n
    assert torch.equal(A0.grad, W2)
