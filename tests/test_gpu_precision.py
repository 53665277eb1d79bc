"""The precision configuration at its real size: S / D GEMM at 8192^3 (NN and
TN, uniform[-1,1] and sign * 2^u with u ~ U[-20, 20]) and C / Z GEMM at
4096^3 (NN, both distributions). Each split / Ozaki form must be no less
accurate than the plain FP32 / FP64 MFMA kernel on the same inputs, in both
the maximum and the mean over 4096 sampled entries of

    e_ij = |C_ij - C_ref_ij| / (|op(A)| |op(B)|)_ij

with C_ref the double-double value. Dividing by the absolute product, not by
|C_ref_ij|, keeps entries whose sum cancels from dominating the maximum (the
relative error of such an entry measures the cancellation, not the kernel).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NS = 4096


def _gen(torch, g, n, dt, wide):
    if not wide:
        return torch.rand(n, generator=g, device="cuda", dtype=dt) * 2 - 1
    u = torch.rand(n, generator=g, device="cuda", dtype=torch.float64) * 40 - 20
    sg = torch.where(torch.rand(n, generator=g, device="cuda") < 0.5, -1.0, 1.0)
    return (sg.double() * torch.exp2(u)).to(dt)


def _operands(torch, g, m, dt, wide):
    if dt in (torch.complex64, torch.complex128):
        rdt = torch.float32 if dt == torch.complex64 else torch.float64
        return [torch.complex(_gen(torch, g, m * m, rdt, wide),
                              _gen(torch, g, m * m, rdt, wide)) for _ in "AB"]
    return [_gen(torch, g, m * m, dt, wide) for _ in "AB"]


def _sample_errors(orc, torch, m, ta, A, B, C, rows, cols):
    """e_ij on the sampled entries of column-major C = op(A) B (beta = 0)."""
    Ah = A.view(m, m)   # Ah[j, i] = A(i, j)
    Bh = B.view(m, m)   # Bh[c, k] = B(k, c)
    Ch = C.view(m, m)   # Ch[c, r] = C(r, c)
    tr = torch.as_tensor(rows, device="cuda")
    tc = torch.as_tensor(cols, device="cuda")
    a_rows = (Ah[tr, :] if ta else Ah[:, tr].t()).contiguous().cpu().numpy()
    b_cols = Bh[tc, :].contiguous().cpu().numpy()
    got = Ch[tc, tr].cpu().numpy()
    d = np.arange(rows.size)
    ref = orc.gemm_sample_exact(1, 1, 0, d.size, d.size, m, 1.0,
                                a_rows.reshape(-1), m, b_cols.reshape(-1), m,
                                0.0, np.zeros(1, a_rows.dtype), d.size, d, d)
    absp = orc.gemm_sample_exact(1, 1, 0, d.size, d.size, m, 1.0,
                                 np.abs(a_rows).astype(np.float64).reshape(-1),
                                 m, np.abs(b_cols).astype(np.float64).reshape(-1),
                                 m, 0.0, np.zeros(1), d.size, d, d)
    assert np.all(np.isfinite(got))
    return np.abs(got.astype(ref.dtype) - ref) / absp


# (dtype tag, size, plain entry, split forms: name -> (explicit entry or
# math mode for the complex types))
REAL = [("s", 8192, {"split_bf16": "hcblas_sgemm_split_bf16",
                     "ozaki_crt": "hcblas_sgemm_ozaki_crt"}),
        ("d", 8192, {"ozaki_crt": "hcblas_dgemm_ozaki_crt",
                     "ozaki_int8": "hcblas_dgemm_ozaki_int8"})]
CPLX = [("c", 4096, {"split_bf16": 1, "ozaki_crt": 8}),
        ("z", 4096, {"ozaki_crt": 4, "ozaki_int8": 2})]


@pytest.mark.parametrize("tag,m,splits", REAL + CPLX,
                         ids=[r[0] for r in REAL + CPLX])
@pytest.mark.parametrize("wide", [False, True], ids=["uniform", "wide"])
def test_split_error_no_worse_than_plain(hc, handle, orc, torch_cuda, tag, m,
                                         splits, wide):
    torch = torch_cuda
    api, lib = hc.api, hc.load()
    dt = {"s": torch.float32, "d": torch.float64, "c": torch.complex64,
          "z": torch.complex128}[tag]
    g = torch.Generator(device="cuda").manual_seed(7 + wide)
    A, B = _operands(torch, g, m, dt, wide)
    C = torch.zeros(m * m, device="cuda", dtype=dt)
    rng = np.random.default_rng(3)
    rows = rng.integers(0, m, NS)
    cols = rng.integers(0, m, NS)
    plain = getattr(api, f"hcblas{tag.upper()}gemm")
    one, zero = (1.0, 0.0) if tag in "sd" else (1 + 0j, 0j)
    for ta in ((0, 1) if tag in "sd" else (0,)):
        assert plain(handle, ta, 0, m, m, m, one, A, m, B, m, zero, C, m) == 0
        torch.cuda.synchronize()
        ep = _sample_errors(orc, torch, m, ta, A, B, C, rows, cols)
        for name, entry in splits.items():
            C.zero_()
            if tag in "sd":
                st = getattr(api, entry)(handle, 1, ta, 0, m, m, m, one, A, m,
                                         B, m, zero, C, m, 0, 0, 0)
            else:
                assert lib.hcblasSetMathMode(handle.value, entry) == 0
                st = plain(handle, ta, 0, m, m, m, one, A, m, B, m, zero, C, m)
                assert lib.hcblasSetMathMode(handle.value, 0) == 0
            assert st == 0
            torch.cuda.synchronize()
            es = _sample_errors(orc, torch, m, ta, A, B, C, rows, cols)
            print(f"{tag} {m} ta={ta} {'wide' if wide else 'uniform'} {name}: "
                  f"plain max {ep.max():.3e} mean {ep.mean():.3e} | split max "
                  f"{es.max():.3e} mean {es.mean():.3e}")
            assert es.max() <= ep.max(), (name, ta, es.max(), ep.max())
            assert es.mean() <= ep.mean(), (name, ta, es.mean(), ep.mean())
