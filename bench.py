#!/usr/bin/env python3
"""Benchmark of the hcBLAS hot path on one MI355X.

Workload (one "step"): the plain bandwidth set --
  SGEMV and DGEMV at M = N = 16384, trans N and T, alpha = 1.5, beta = 0.5;
  SGER and DGER at 8192 x 8192;
  SAXPY, SSCAL, SCOPY, SDOT, SASUM at n = 2^28 (fp32);
  DDOT, DASUM at n = 2^27 (fp64);
all entries uniform[-1,1] (synthetic, seeded, resident in HBM before timing),
unit stride, called through the C ABI (libhcblas.so).

value = total algorithmic bytes of the step / step time (GB/s); bytes are
compulsory traffic (each operand read once, outputs written once, y read only
when beta != 0). The JSON line also carries
  roofline      -- the dominant kernel (largest share of the step) measured
                   with HIP events on the stream it is launched on;
  cpu_baseline  -- CBLAS (OpenBLAS through scipy, the job's host cores) on
                   the full bandwidth set with the GPU's own inputs (the
                   oracle's single-thread C restatement on a bounded sample
                   when no BLAS library is importable); cpu_gemm gives the
                   host SGEMM / DGEMM rates of the GEMM configurations;
  ops           -- per-routine GB/s;
  gemm          -- the GEMM half of the metric: plain SGEMM/DGEMM 8192^3 and
                   split / Ozaki forms in TFLOP/s with max/mean relative error
                   against a double-double reference on sampled entries, plus
                   the 1024^3 row-major SGEMM configuration.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"fp32": 157.3, "fp64": 78.6, "bf16": 2516.6, "fp16": 2516.6,
               "int8": 5033.0}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-gemm", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--gemm-n", type=int, default=8192)
    ap.add_argument("--case", default=None, choices=sorted(CASES),
                    help="run only this GEMM case `--reps` times (for "
                         "profiler counter passes) and print its median ms")
    ap.add_argument("--reps", type=int, default=5)
    return ap.parse_args()


# Single GEMM cases of the GEMM configurations, for profiler passes:
# name -> (dtype tag, size, entry, row-major?)
CASES = {
    "sgemm_8192": ("s", 8192, "hcblasSgemm", False),
    "dgemm_8192": ("d", 8192, "hcblasDgemm", False),
    "sgemm_rowmajor_1024": ("s", 1024, "hcblas_sgemm", True),
    "hgemm_8192": ("h", 8192, "hcblas_hgemm", False),
    "sgemm_2048": ("s", 2048, "hcblasSgemm", False),
    "dgemm_2048": ("d", 2048, "hcblasDgemm", False),
    "dgemm_1024": ("d", 1024, "hcblasDgemm", False),
    "sgemm_split_bf16_8192": ("s", 8192, "hcblas_sgemm_split_bf16", False),
    "sgemm_ozaki_crt_8192": ("s", 8192, "hcblas_sgemm_ozaki_crt", False),
    "dgemm_ozaki_crt_8192": ("d", 8192, "hcblas_dgemm_ozaki_crt", False),
    "grouped_4096": (None, 0, None, False),
    "adjoint_bf16": (None, 0, None, False),
}


def run_case(hc, h, torch, name, reps):
    """Runs one entry of CASES `reps` times after one warm-up call; returns
    the median milliseconds per call (HIP events on the handle's stream)."""
    api = hc.api
    tag, n, entry, row = CASES[name]
    g = torch.Generator(device="cuda").manual_seed(7)
    if name == "grouped_4096":
        problems, meta, arenas = grouped_problem_table(torch, hc, 4096)
        plan = hc.GroupedPlan(h, problems)
        if plan.autotune(3) != 0:
            raise RuntimeError("autotune failed")
        fn = plan.run
    elif name == "adjoint_bf16":
        from hcblas_amd import torch_ext
        b, M, N, K = 256, 256, 1024, 512
        A = (torch.rand(b, M, K, generator=g, device="cuda") * 2 - 1).bfloat16()
        B = (torch.rand(b, K, N, generator=g, device="cuda") * 2 - 1).bfloat16()
        W = torch.rand(b, M, N, generator=g, device="cuda") * 2 - 1
        A.requires_grad_(True)
        B.requires_grad_(True)

        def fn():
            A.grad = None
            B.grad = None
            (torch_ext.gemm_batched(A, B) * W).sum().backward()
            return 0
    else:
        dt = {"s": torch.float32, "d": torch.float64, "h": torch.float16}[tag]
        A = (torch.rand(n * n, generator=g, device="cuda") * 2 - 1).to(dt)
        B = (torch.rand(n * n, generator=g, device="cuda") * 2 - 1).to(dt)
        C = torch.zeros(n * n, device="cuda", dtype=dt)
        f = getattr(api, entry)
        if entry.startswith("hcblas_"):
            order = 0 if row else 1
            fn = lambda: f(h, order, 0, 0, n, n, n, 1.0, A, n, B, n, 0.0, C, n,
                           0, 0, 0)
        else:
            fn = lambda: f(h, 0, 0, n, n, n, 1.0, A, n, B, n, 0.0, C, n)
    if fn() != 0:
        raise RuntimeError(f"{name} failed")
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ts = []
    for _ in range(reps):
        s.record()
        fn()
        e.record()
        torch.cuda.synchronize()
        ts.append(s.elapsed_time(e))
    return sorted(ts)[len(ts) // 2]


def bandwidth_ops(hc, h, torch):
    """(name, algorithmic bytes, callable) for the bandwidth set."""
    api = hc.api
    dev = "cuda"
    g = torch.Generator(device=dev).manual_seed(1234)

    def U(n, dt):
        return torch.rand(n, generator=g, device=dev, dtype=dt) * 2 - 1

    f32, f64 = torch.float32, torch.float64
    n_mv = 16384
    n_r1 = 8192
    n_s = 1 << 28
    n_d = 1 << 27
    A_s = U(n_mv * n_mv, f32)
    A_d = U(n_mv * n_mv, f64)
    xs, ys = U(n_mv, f32), U(n_mv, f32)
    xd, yd = U(n_mv, f64), U(n_mv, f64)
    G_s = U(n_r1 * n_r1, f32)
    G_d = U(n_r1 * n_r1, f64)
    gxs, gys = U(n_r1, f32), U(n_r1, f32)
    gxd, gyd = U(n_r1, f64), U(n_r1, f64)
    vx, vy = U(n_s, f32), U(n_s, f32)
    vz = torch.empty_like(vx)
    dx, dy = U(n_d, f64), U(n_d, f64)

    def chk(st):
        if st != 0:
            raise RuntimeError(f"status {st}")

    mv_s = 4 * (n_mv * n_mv + 3 * n_mv)
    mv_d = 8 * (n_mv * n_mv + 3 * n_mv)
    r1_s = 4 * (2 * n_r1 * n_r1 + 2 * n_r1)
    r1_d = 8 * (2 * n_r1 * n_r1 + 2 * n_r1)
    ops = [
        ("sgemv_n", mv_s, lambda: chk(api.hcblasSgemv(h, 0, n_mv, n_mv, 1.5, A_s, n_mv, xs, 1, 0.5, ys, 1))),
        ("sgemv_t", mv_s, lambda: chk(api.hcblasSgemv(h, 1, n_mv, n_mv, 1.5, A_s, n_mv, xs, 1, 0.5, ys, 1))),
        ("dgemv_n", mv_d, lambda: chk(api.hcblasDgemv(h, 0, n_mv, n_mv, 1.5, A_d, n_mv, xd, 1, 0.5, yd, 1))),
        ("dgemv_t", mv_d, lambda: chk(api.hcblasDgemv(h, 1, n_mv, n_mv, 1.5, A_d, n_mv, xd, 1, 0.5, yd, 1))),
        ("sger", r1_s, lambda: chk(api.hcblasSger(h, n_r1, n_r1, 1.5, gxs, 1, gys, 1, G_s, n_r1))),
        ("dger", r1_d, lambda: chk(api.hcblasDger(h, n_r1, n_r1, 1.5, gxd, 1, gyd, 1, G_d, n_r1))),
        ("saxpy", 3 * 4 * n_s, lambda: chk(api.hcblasSaxpy(h, n_s, 1.5, vx, 1, vy, 1))),
        ("sscal", 2 * 4 * n_s, lambda: chk(api.hcblasSscal(h, n_s, -1.0, vy, 1))),
        ("scopy", 2 * 4 * n_s, lambda: chk(api.hcblasScopy(h, n_s, vx, 1, vz, 1))),
        ("sdot", 2 * 4 * n_s, lambda: chk(api.hcblasSdot(h, n_s, vx, 1, vy, 1, None)[0])),
        ("sasum", 4 * n_s, lambda: chk(api.hcblasSasum(h, n_s, vx, 1, None)[0])),
        ("ddot", 2 * 8 * n_d, lambda: chk(api.hcblasDdot(h, n_d, dx, 1, dy, 1, None)[0])),
        ("dasum", 8 * n_d, lambda: chk(api.hcblasDasum(h, n_d, dx, 1, None)[0])),
    ]
    tensors = {"A_s": A_s, "A_d": A_d, "xs": xs, "ys": ys, "xd": xd, "yd": yd,
               "G_s": G_s, "G_d": G_d, "gxs": gxs, "gys": gys, "gxd": gxd,
               "gyd": gyd, "vx": vx, "vy": vy, "vz": vz, "dx": dx, "dy": dy}
    return ops, tensors


def measured_traffic(op):
    """HBM bytes per launch of `op`'s kernel from profiles/traffic.json, which
    holds rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE per dispatch of the bench
    kernels, each scaled by what the counter reads for scopy (a plain copy of
    known size). None when no such measurement is committed."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)["ops"].get(op)
    except (OSError, ValueError, KeyError):
        return None
    if not rec:
        return None
    return int(rec["fetch_bytes"] + rec["write_bytes"])


def time_bandwidth(torch, ops, steps, warmup):
    """Whole-step time (every op once per step, including the host round trip
    of the synchronous dot / asum results), then each op's own duration from
    a burst of `steps` back-to-back calls bracketed by events on the stream
    the library launches on (torch's current stream, set on the handle), so
    host launch gaps do not count against a kernel."""
    for _ in range(warmup):
        for _, _, fn in ops:
            fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(steps):
        for _, _, fn in ops:
            fn()
    end.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_ms = start.elapsed_time(end) / steps
    per_op = []
    for _, _, fn in ops:
        fn()
        s = torch.cuda.Event(enable_timing=True)
        e = torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(steps):
            fn()
        e.record()
        torch.cuda.synchronize()
        per_op.append(s.elapsed_time(e) / steps)
    return step_ms, per_op, wall


def gemm_section(hc, h, torch, oracle, n, reps=20, nc=4096):
    """Plain vs split-precision GEMM (the precision configuration):
    S / D GEMM at n^3, NN and TN (column-major), uniform[-1,1] and
    sign * 2^u (u ~ U[-20, 20]); C / Z GEMM at nc^3, NN, uniform. Every
    result carries its median time over `reps` calls, TFLOP/s at the
    plain-equivalent 2MNK (8MNK complex), and max / mean relative error vs
    the double-double oracle on 4096 sampled entries; split results also
    carry their speed ratio to the plain kernel on the same inputs."""
    import numpy as np
    api = hc.api
    lib = hc.load()
    out = {}
    g = torch.Generator(device="cuda").manual_seed(7)
    rng = np.random.default_rng(0)
    ns = 4096

    def timed(fn):
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        times = []
        for _ in range(reps):
            s.record()
            fn()
            e.record()
            torch.cuda.synchronize()
            times.append(s.elapsed_time(e))
        return sorted(times)[len(times) // 2]

    def errors(m, ta, A, B, C, dt_np):
        # op(A)(r, :) and B(:, c) of the sampled entries (column-major
        # storage), then one oracle call on the ns x m / m x ns slices whose
        # diagonal holds the sampled entries (beta = 0: C is not read)
        rows = rng.integers(0, m, ns)
        cols = rng.integers(0, m, ns)
        Ah = A.view(m, m)   # Ah[j, i] = A_stored(i, j)
        Bh = B.view(m, m)   # Bh[c, k] = B(k, c)
        Ch = C.view(m, m)   # Ch[c, r] = C(r, c)
        tr = torch.as_tensor(rows, device="cuda")
        tc = torch.as_tensor(cols, device="cuda")
        a_rows = (Ah[tr, :] if ta else Ah[:, tr].t()).contiguous().cpu().numpy()
        b_cols = Bh[tc, :].contiguous().cpu().numpy()
        got = Ch[tc, tr].cpu().numpy()
        diag = np.arange(ns)
        ref = oracle.gemm_sample_exact(1, 1, 0, ns, ns, m, 1.0, a_rows.reshape(-1), m,
                                       b_cols.reshape(-1), m, 0.0,
                                       np.zeros(1, a_rows.dtype), ns, diag, diag)
        absp = oracle.gemm_sample_exact(
            1, 1, 0, ns, ns, m, 1.0, np.abs(a_rows).astype(np.float64).reshape(-1), m,
            np.abs(b_cols).astype(np.float64).reshape(-1), m, 0.0, np.zeros(1), ns,
            diag, diag)
        err = np.abs(got.astype(ref.dtype) - ref)
        rel = err / np.maximum(np.abs(ref), 1e-300)
        nrm = err / absp
        return (float(rel.max()), float(rel.mean()), float(nrm.max()),
                float(nrm.mean()))

    def result(ms, flop, peak_key, err):
        tf = flop / (ms * 1e-3) / 1e12
        return {"tflops": round(tf, 2), "ms": round(ms, 3),
                "peak_frac": round(tf / PEAK_TFLOPS[peak_key], 4),
                "max_rel_err": err[0], "mean_rel_err": err[1],
                "max_err_vs_absprod": err[2], "mean_err_vs_absprod": err[3]}

    def gen(m, dt, wide):
        if not wide:
            return torch.rand(m * m, generator=g, device="cuda", dtype=dt) * 2 - 1
        u = torch.rand(m * m, generator=g, device="cuda", dtype=torch.float64) * 40 - 20
        sg = torch.where(torch.rand(m * m, generator=g, device="cuda") < 0.5, -1.0, 1.0)
        return (sg.double() * torch.exp2(u)).to(dt)

    real = (("s", torch.float32, "fp32",
             (("split_bf16", api.hcblas_sgemm_split_bf16),
              ("ozaki_crt", api.hcblas_sgemm_ozaki_crt))),
            ("d", torch.float64, "fp64",
             (("ozaki_crt", api.hcblas_dgemm_ozaki_crt),
              ("ozaki_int8", api.hcblas_dgemm_ozaki_int8))))
    for tag, dt, pk, splits in real:
        plain_fn = getattr(api, f"hcblas{tag.upper()}gemm")
        for wide in (False, True):
            A = gen(n, dt, wide)
            B = gen(n, dt, wide)
            C = torch.zeros(n * n, device="cuda", dtype=dt)
            for ta in (0, 1):
                key = f"{tag}gemm_{'tn' if ta else 'nn'}_{'wide' if wide else 'uniform'}"
                ent = {}
                fns = [("plain", lambda: plain_fn(h, ta, 0, n, n, n, 1.0, A, n, B, n, 0.0, C, n))]
                for name, f in splits:
                    fns.append((name, lambda f=f: f(h, 1, ta, 0, n, n, n, 1.0, A, n, B, n,
                                                    0.0, C, n, 0, 0, 0)))
                for name, fn in fns:
                    st = fn()
                    torch.cuda.synchronize()
                    if st != 0:
                        ent[name] = {"status": int(st)}
                        continue
                    ms = timed(fn)
                    ent[name] = result(ms, 2.0 * n ** 3, pk, errors(n, ta, A, B, C, dt))
                for name, _ in splits:
                    if "ms" in ent.get(name, {}) and "ms" in ent["plain"]:
                        ent[name]["vs_plain"] = round(ent["plain"]["ms"] / ent[name]["ms"], 3)
                out[key] = ent
            del A, B, C
            torch.cuda.empty_cache()
    # complex at nc^3 through the handle's math modes
    for tag, dt, pk, modes in (("c", torch.complex64, "fp32", (("split_bf16", 1), ("ozaki_crt", 8))),
                               ("z", torch.complex128, "fp64", (("ozaki_crt", 4), ("ozaki_int8", 2)))):
        rdt = torch.float32 if tag == "c" else torch.float64
        A = torch.complex(gen(nc, rdt, False), gen(nc, rdt, False))
        B = torch.complex(gen(nc, rdt, False), gen(nc, rdt, False))
        C = torch.zeros(nc * nc, device="cuda", dtype=dt)
        fn = getattr(api, f"hcblas{tag.upper()}gemm")
        call = lambda: fn(h, 0, 0, nc, nc, nc, 1.0 + 0j, A, nc, B, nc, 0j, C, nc)
        ent = {}
        for name, mode in (("plain", 0),) + modes:
            lib.hcblasSetMathMode(h.value, mode)
            st = call()
            torch.cuda.synchronize()
            if st != 0:
                ent[name] = {"status": int(st)}
                continue
            ms = timed(call)
            ent[name] = result(ms, 8.0 * nc ** 3, pk, errors(nc, 0, A, B, C, dt))
        lib.hcblasSetMathMode(h.value, 0)
        for name, _ in modes:
            if "ms" in ent.get(name, {}) and "ms" in ent["plain"]:
                ent[name]["vs_plain"] = round(ent["plain"]["ms"] / ent[name]["ms"], 3)
        out[f"{tag}gemm_{nc}_nn_uniform"] = ent
        del A, B, C
        torch.cuda.empty_cache()
    # headline ratios: plain-equivalent speed-up on NN uniform at n^3, and the
    # worst case over the four S / D cases
    def ratios(tag, name):
        v = [out[k][name]["vs_plain"] for k in out
             if k.startswith(f"{tag}gemm_") and name in out[k] and "vs_plain" in out[k][name]
             and not k.startswith(f"{tag}gemm_{nc}")]
        return {"nn_uniform": out[f"{tag}gemm_nn_uniform"][name].get("vs_plain"),
                "min": min(v) if v else None}
    out["speedup_vs_plain"] = {"sgemm_split_bf16": ratios("s", "split_bf16"),
                               "sgemm_ozaki_crt": ratios("s", "ozaki_crt"),
                               "dgemm_ozaki_crt": ratios("d", "ozaki_crt"),
                               "dgemm_ozaki_int8": ratios("d", "ozaki_int8")}
    # 1024^3 row-major NN SGEMM (plain-SGEMM configuration). Throughput is
    # the mean over 50 back-to-back calls between two events (the device
    # time per call); latency_ms is the median of single calls bracketed by
    # events, which adds the host's launch time (about 6 us) to the kernel.
    m = 1024
    A = torch.rand(m * m, generator=g, device="cuda") * 2 - 1
    B = torch.rand(m * m, generator=g, device="cuda") * 2 - 1
    C = torch.zeros(m * m, device="cuda")
    fn = lambda: api.hcblas_sgemm(h, 0, 0, 0, m, m, m, 1.0, A, m, B, m, 0.0, C, m, 0, 0, 0)
    fn()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ts = []
    for _ in range(20):
        s.record()
        fn()
        e.record()
        torch.cuda.synchronize()
        ts.append(s.elapsed_time(e))
    lat = sorted(ts)[len(ts) // 2]
    runs = []
    for _ in range(5):
        s.record()
        for _ in range(50):
            fn()
        e.record()
        torch.cuda.synchronize()
        runs.append(s.elapsed_time(e) / 50)
    ms = sorted(runs)[len(runs) // 2]
    out["sgemm_rowmajor_1024"] = {"tflops": round(2.0 * m ** 3 / (ms * 1e-3) / 1e12, 2),
                                  "ms": round(ms, 4), "calls_timed": 50,
                                  "latency_ms": round(lat, 4),
                                  "peak_frac": round(2.0 * m ** 3 / (ms * 1e-3) / 1e12
                                                     / PEAK_TFLOPS["fp32"], 4)}
    # For scale: torch's FP32 matmul (hipBLASLt, TF32 off) on the same
    # row-major operands, timed the same way (not part of the product).
    torch.backends.cuda.matmul.allow_tf32 = False
    Am, Bm = A.view(m, m), B.view(m, m)
    torch.mm(Am, Bm)
    torch.cuda.synchronize()
    runs = []
    for _ in range(5):
        s.record()
        for _ in range(50):
            torch.mm(Am, Bm)
        e.record()
        torch.cuda.synchronize()
        runs.append(s.elapsed_time(e) / 50)
    tms = sorted(runs)[len(runs) // 2]
    out["sgemm_rowmajor_1024"]["torch_matmul_ms"] = round(tms, 4)
    out["sgemm_rowmajor_1024"]["torch_matmul_peak_frac"] = round(
        2.0 * m ** 3 / (tms * 1e-3) / 1e12 / PEAK_TFLOPS["fp32"], 4)
    del A, B, C
    # HGEMM (binary16 in / out, FP32 accumulation) at 8192^3, column-major
    # NN and TN: median of 5 single calls, with torch's half matmul on the
    # same NN operands for scale.
    n = 8192
    Ah = (torch.rand(n * n, generator=g, device="cuda") * 2 - 1).half()
    Bh = (torch.rand(n * n, generator=g, device="cuda") * 2 - 1).half()
    Ch = torch.zeros(n * n, device="cuda").half()
    one = 1.0

    def med(f):
        f()
        torch.cuda.synchronize()
        ts = []
        for _ in range(5):
            s.record()
            f()
            e.record()
            torch.cuda.synchronize()
            ts.append(s.elapsed_time(e))
        return sorted(ts)[2]
    hg = {}
    for ta in (0, 1):
        ms = med(lambda: api.hcblas_hgemm(h, 1, ta, 0, n, n, n, one, Ah, n,
                                          Bh, n, 0.0, Ch, n, 0, 0, 0))
        tf = 2.0 * n ** 3 / (ms * 1e-3) / 1e12
        hg["nt"[ta] + "n"] = {"ms": round(ms, 4), "tflops": round(tf, 1),
                              "peak_frac": round(tf / PEAK_TFLOPS["fp16"], 4)}
    Am, Bm = Ah.view(n, n), Bh.view(n, n)
    tms = med(lambda: torch.mm(Bm, Am))
    hg["torch_matmul_ms"] = round(tms, 4)
    out["hgemm_8192"] = hg
    return out


def convnet_section(hc, h, torch, reps=20):
    """Column-major SGEMM over the 155 distinct shapes of the reference's
    convolution-network lists (AlexNet, GoogLeNet, OverFeat, Oxford;
    tests/golden/convnet_shapes.json): median time of `reps` calls per shape,
    aggregate TF/s = sum 2MNK / sum of the medians, per list and overall."""
    import json
    api = hc.api
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests",
                        "golden", "convnet_shapes.json")
    with open(path) as f:
        rows = [r for r in json.load(f)["shapes"] if r[13].endswith("BlasDimensions")]
    per = {}
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for r in rows:
        M, N, K, ta, tb, lda, ldb, ldc = r[:8]
        alpha, beta = r[8], r[9]
        A = torch.rand(lda * (M if ta else K), device="cuda")
        B = torch.rand(ldb * (K if tb else N), device="cuda")
        C = torch.rand(ldc * N, device="cuda")
        fn = lambda: api.hcblas_sgemm(h, 1, ta, tb, M, N, K, alpha, A, lda, B, ldb,
                                      beta, C, ldc, 0, 0, 0)
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            s.record()
            fn()
            e.record()
            torch.cuda.synchronize()
            ts.append(s.elapsed_time(e))
        ms = sorted(ts)[len(ts) // 2]
        agg = per.setdefault(r[13].replace("BlasDimensions", "").lower(),
                             {"shapes": 0, "flop": 0.0, "ms": 0.0})
        agg["shapes"] += 1
        agg["flop"] += 2.0 * M * N * K
        agg["ms"] += ms
        del A, B, C
    tot_f = sum(v["flop"] for v in per.values())
    tot_ms = sum(v["ms"] for v in per.values())
    out = {"shapes": sum(v["shapes"] for v in per.values()),
           "tflops": round(tot_f / (tot_ms * 1e-3) / 1e12, 2),
           "ms_sum": round(tot_ms, 3), "per_list": {}}
    for k, v in per.items():
        out["per_list"][k] = {"shapes": v["shapes"], "ms_sum": round(v["ms"], 3),
                              "tflops": round(v["flop"] / (v["ms"] * 1e-3) / 1e12, 2)}
    return out


def grouped_problem_table(torch, hc, count=4096, seed=2024):
    """The grouped-GEMM configuration: `count` independent NN problems with
    M, N, K each log-uniform on [16, 512] rounded to a multiple of 8, half
    fp32 SGEMM and half fp16 HGEMM (fp32 accumulate), row- or column-major at
    random, entries uniform[-1,1] (seeded), alpha = 1, beta = 0.
    Returns (problems, meta, arenas); `meta` rows are
    (dtype, order, M, N, K, a_off, b_off, c_off, lda, ldb, ldc) in elements of
    the dtype's arena."""
    import numpy as np
    rng = np.random.default_rng(seed)

    def logu(n):
        v = np.exp(rng.uniform(np.log(16), np.log(512), n))
        return (np.round(v / 8) * 8).astype(np.int64)

    M, N, K = logu(count), logu(count), logu(count)
    dt = rng.integers(0, 2, count)
    order = rng.integers(0, 2, count)
    sizes = [0, 0]
    meta = []
    for i in range(count):
        d = int(dt[i])
        al = 64  # elements; keeps every matrix 256-byte aligned
        offs = []
        for n_el in (M[i] * K[i], K[i] * N[i], M[i] * N[i]):
            offs.append(sizes[d])
            sizes[d] += int((n_el + al - 1) // al * al)
        row = order[i] == 0
        lda = K[i] if row else M[i]
        ldb = N[i] if row else K[i]
        ldc = N[i] if row else M[i]
        meta.append((d, int(order[i]), int(M[i]), int(N[i]), int(K[i]),
                     offs[0], offs[1], offs[2], int(lda), int(ldb), int(ldc)))
    g = torch.Generator(device="cuda").manual_seed(seed)
    ar32 = torch.rand(max(1, sizes[0]), generator=g, device="cuda") * 2 - 1
    ar16 = (torch.rand(max(1, sizes[1]), generator=g, device="cuda") * 2 - 1).half()
    arenas = (ar32, ar16)
    problems = []
    for (d, o, m, n, k, ao, bo, co, lda, ldb, ldc) in meta:
        base = arenas[d].data_ptr()
        es = 4 if d == 0 else 2
        problems.append(hc.GemmProblem(
            hc.R_32F if d == 0 else hc.R_16F, o, hc.OP_N, hc.OP_N, m, n, k, 0,
            1.0, 0.0, base + ao * es, lda, base + bo * es, ldb,
            base + co * es, ldc))
    return problems, meta, arenas


def run_plain_group(hc, h, meta, arenas, c_arenas):
    """The same problems as per-problem plain launches (hcblas_sgemm /
    hcblas_hgemm), writing C into `c_arenas`."""
    api = hc.api
    for (d, o, m, n, k, ao, bo, co, lda, ldb, ldc) in meta:
        es = 4 if d == 0 else 2
        A = arenas[d].data_ptr() + ao * es
        B = arenas[d].data_ptr() + bo * es
        C = c_arenas[d].data_ptr() + co * es
        fn = api.hcblas_sgemm if d == 0 else api.hcblas_hgemm
        st = fn(h, o, 0, 0, m, n, k, 1.0, A, lda, B, ldb, 0.0, C, ldc, 0, 0, 0)
        if st != 0:
            raise RuntimeError(f"plain launch status {st}")


def grouped_section(hc, h, torch, count=4096, reps=10):
    """Grouped GEMM: one autotuned, graph-captured persistent launch versus
    per-problem plain launches on the same 4096-problem table."""
    problems, meta, arenas = grouped_problem_table(torch, hc, count)
    flops32 = sum(2.0 * m * n * k for (d, o, m, n, k, *_) in meta if d == 0)
    flops16 = sum(2.0 * m * n * k for (d, o, m, n, k, *_) in meta if d == 1)
    plan = hc.GroupedPlan(h, problems)
    if plan.autotune(3) != 0:
        raise RuntimeError("autotune failed")
    info = plan.info()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    plan.run()
    torch.cuda.synchronize()
    s.record()
    for _ in range(reps):
        plan.run()
    e.record()
    torch.cuda.synchronize()
    ms_g = s.elapsed_time(e) / reps
    c_ar = (torch.empty_like(arenas[0]), torch.empty_like(arenas[1]))
    run_plain_group(hc, h, meta, arenas, c_ar)
    torch.cuda.synchronize()
    s.record()
    for _ in range(3):
        run_plain_group(hc, h, meta, arenas, c_ar)
    e.record()
    torch.cuda.synchronize()
    ms_p = s.elapsed_time(e) / 3
    plan.destroy()
    bound_ms = (flops32 / (PEAK_TFLOPS["fp32"] * 1e12) +
                flops16 / (PEAK_TFLOPS["bf16"] * 1e12)) * 1e3
    # compulsory HBM traffic: A and B read once, C written once (beta = 0)
    hbm_bytes = sum((m * k + k * n + m * n) * (4 if d == 0 else 2)
                    for (d, o, m, n, k, *_) in meta)
    hbm_ms = hbm_bytes / (HBM_PEAK_GBS * 1e9) * 1e3
    tot = flops32 + flops16
    return {"problems": count, "flop": tot, "fp32_flop": flops32,
            "grouped_ms": round(ms_g, 4), "grouped_tflops": round(tot / ms_g / 1e9, 2),
            "plain_launches_ms": round(ms_p, 3),
            "plain_launches_tflops": round(tot / ms_p / 1e9, 2),
            "speedup": round(ms_p / ms_g, 2),
            "mfma_bound_ms": round(bound_ms, 4),
            "hbm_bytes": hbm_bytes, "hbm_bound_ms": round(hbm_ms, 4),
            "roofline_frac": round(max(bound_ms, hbm_ms) / ms_g, 4),
            "plan": info}


def adjoint_section(torch, reps=5):
    """Differentiation configuration through the libtorch layer: forward and
    backward of C = A @ B over 256 problems of 256 x 1024 x 512 in bf16 with
    fp32 accumulation (loss sum(C o W)), and of a 4096^2 SGER."""
    from hcblas_amd import torch_ext
    torch_ext.load()
    g = torch.Generator(device="cuda").manual_seed(9)
    batch, M, N, K = 256, 256, 1024, 512
    A = (torch.rand((batch, M, K), generator=g, device="cuda") * 2 - 1).to(torch.bfloat16)
    B = (torch.rand((batch, K, N), generator=g, device="cuda") * 2 - 1).to(torch.bfloat16)
    W = torch.rand((batch, M, N), generator=g, device="cuda") * 2 - 1
    A.requires_grad_()
    B.requires_grad_()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def fwd():
        return torch_ext.gemm_batched(A, B)

    def fwd_bwd():
        # gradients set to None as a training step's zero_grad does, so the
        # backward writes them instead of adding into the previous ones
        A.grad = None
        B.grad = None
        C = torch_ext.gemm_batched(A, B)
        C.backward(W)

    out = {}
    for name, fn in (("forward", fwd), ("forward_backward", fwd_bwd)):
        fn()
        torch.cuda.synchronize()
        s.record()
        for _ in range(reps):
            fn()
        e.record()
        torch.cuda.synchronize()
        out[name + "_ms"] = round(s.elapsed_time(e) / reps, 4)
    flop = 2.0 * batch * M * N * K
    out["forward_tflops"] = round(flop / out["forward_ms"] / 1e9, 1)
    out["backward_tflops"] = round(
        2 * flop / (out["forward_backward_ms"] - out["forward_ms"]) / 1e9, 1)
    n = 4096
    x = torch.rand(n, generator=g, device="cuda").requires_grad_()
    y = torch.rand(n, generator=g, device="cuda").requires_grad_()
    A0 = torch.rand((n, n), generator=g, device="cuda").requires_grad_()
    W2 = torch.rand((n, n), generator=g, device="cuda")

    def ger_fb():
        torch_ext.ger(x, y, A0, 1.5).backward(W2)

    ger_fb()
    torch.cuda.synchronize()
    s.record()
    for _ in range(reps):
        ger_fb()
    e.record()
    torch.cuda.synchronize()
    out["sger_forward_backward_ms"] = round(s.elapsed_time(e) / reps, 4)
    return out


def cpu_threads():
    """Host threads for the CPU baseline: the cores this process may run on,
    capped by OMP_NUM_THREADS when the environment sets it (the share of a
    shared host that belongs to this job)."""
    n = len(os.sched_getaffinity(0))
    try:
        cap = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        cap = 0
    return min(n, cap) if cap > 0 else n


def host_copies(tensors):
    """Host (numpy) copies of the bandwidth set's device inputs, so that the
    CPU baseline runs on the same seeded values as the GPU."""
    return {k: t.cpu().numpy() for k, t in tensors.items()}


def cpu_baseline_blas(host, reps=5):
    """CBLAS (OpenBLAS through scipy.linalg.blas, all the job's host cores)
    on the full bandwidth set with the GPU's inputs: 1 warm-up, then the
    median of `reps` runs per routine; value = total algorithmic bytes /
    sum of the medians. None when scipy's BLAS is not importable."""
    try:
        import numpy as np
        import threadpoolctl
        from scipy.linalg import blas
        import scipy
    except ImportError:
        return None
    threads = cpu_threads()
    info = [i for i in threadpoolctl.threadpool_info()
            if i.get("user_api") == "blas" and "scipy.libs" in i.get("filepath", "")]
    lib = (f"{info[0]['internal_api']} {info[0]['version']}" if info
           else "unknown BLAS")
    n_mv, n_r1 = 16384, 8192
    n_s, n_d = host["vx"].size, host["dx"].size

    def colmajor(flat, m, n):  # m x n column-major view, Fortran-contiguous
        return flat.reshape(n, m).T

    A_s = colmajor(host["A_s"], n_mv, n_mv)
    A_d = colmajor(host["A_d"], n_mv, n_mv)
    G_s = colmajor(host["G_s"], n_r1, n_r1)
    G_d = colmajor(host["G_d"], n_r1, n_r1)
    xs, ys, xd, yd = host["xs"], host["ys"], host["xd"], host["yd"]
    gxs, gys, gxd, gyd = host["gxs"], host["gys"], host["gxd"], host["gyd"]
    vx, vy, vz = host["vx"], host["vy"], host["vz"]
    dx, dy = host["dx"], host["dy"]
    mv_s = 4 * (n_mv * n_mv + 3 * n_mv)
    mv_d = 8 * (n_mv * n_mv + 3 * n_mv)
    r1_s = 4 * (2 * n_r1 * n_r1 + 2 * n_r1)
    r1_d = 8 * (2 * n_r1 * n_r1 + 2 * n_r1)
    jobs = [
        ("sgemv_n", mv_s, lambda: blas.sgemv(1.5, A_s, xs, 0.5, ys, trans=0, overwrite_y=1)),
        ("sgemv_t", mv_s, lambda: blas.sgemv(1.5, A_s, xs, 0.5, ys, trans=1, overwrite_y=1)),
        ("dgemv_n", mv_d, lambda: blas.dgemv(1.5, A_d, xd, 0.5, yd, trans=0, overwrite_y=1)),
        ("dgemv_t", mv_d, lambda: blas.dgemv(1.5, A_d, xd, 0.5, yd, trans=1, overwrite_y=1)),
        ("sger", r1_s, lambda: blas.sger(1.5, gxs, gys, a=G_s, overwrite_a=1)),
        ("dger", r1_d, lambda: blas.dger(1.5, gxd, gyd, a=G_d, overwrite_a=1)),
        ("saxpy", 12 * n_s, lambda: blas.saxpy(vx, vy, a=1.5)),
        ("sscal", 8 * n_s, lambda: blas.sscal(-1.0, vy)),
        ("scopy", 8 * n_s, lambda: blas.scopy(vx, vz)),
        ("sdot", 8 * n_s, lambda: blas.sdot(vx, vy)),
        ("sasum", 4 * n_s, lambda: blas.sasum(vx)),
        ("ddot", 16 * n_d, lambda: blas.ddot(dx, dy)),
        ("dasum", 8 * n_d, lambda: blas.dasum(dx)),
    ]
    per_op = {}
    total_b, total_t = 0, 0.0
    with threadpoolctl.threadpool_limits(limits=threads, user_api="blas"):
        for name, b, fn in jobs:
            fn()
            ts = []
            for _ in range(reps):
                t0 = time.perf_counter()
                fn()
                ts.append(time.perf_counter() - t0)
            t = sorted(ts)[len(ts) // 2]
            per_op[name] = round(b / t / 1e9, 2)
            total_b += b
            total_t += t
    del np
    return {"value": round(total_b / total_t / 1e9, 3), "unit": "GB/s",
            "cores": threads, "kind": "reference",
            "library": f"{lib} (scipy {scipy.__version__} scipy.linalg.blas), "
                       "the CBLAS the reference's tests compare against",
            "sample": "the full bandwidth set on the GPU run's own inputs: "
                      "S/DGEMV 16384^2 N+T, S/DGER 8192^2, SAXPY/SSCAL/SCOPY/"
                      "SDOT/SASUM n=2^28, DDOT/DASUM n=2^27; median of "
                      f"{reps} after 1 warm-up per routine",
            "seconds": round(total_t, 3), "ops": per_op}


def cpu_gemm_blas(n_small=1024, n_big=8192, reps=5):
    """CBLAS SGEMM / DGEMM on the host (OpenBLAS through scipy, all the job's
    host cores): the 1024^3 row-major NN SGEMM configuration and the plain
    8192^3 NN S/DGEMM of the precision configuration, uniform[-1,1] inputs;
    1 warm-up, median of `reps` (2 for 8192^3)."""
    try:
        import numpy as np
        import threadpoolctl
        from scipy.linalg import blas
    except ImportError:
        return None
    threads = cpu_threads()
    rng = np.random.default_rng(0)
    out = {"cores": threads}
    with threadpoolctl.threadpool_limits(limits=threads, user_api="blas"):
        for key, n, dt, fn, r in (("sgemm_rowmajor_1024", n_small, np.float32,
                                   blas.sgemm, reps),
                                  ("sgemm_8192", n_big, np.float32, blas.sgemm, 2),
                                  ("dgemm_8192", n_big, np.float64, blas.dgemm, 2)):
            A = rng.uniform(-1, 1, (n, n)).astype(dt)  # row-major M x K
            B = rng.uniform(-1, 1, (n, n)).astype(dt)  # row-major K x N
            # row-major C = A B is column-major C^T = B^T A^T; .T of a
            # C-ordered array is the Fortran-ordered transpose (no copy)
            call = (lambda A=A, B=B, fn=fn: fn(1.0, B.T, A.T))
            call()
            ts = []
            for _ in range(r):
                t0 = time.perf_counter()
                call()
                ts.append(time.perf_counter() - t0)
            t = sorted(ts)[len(ts) // 2]
            out[key] = {"gflops": round(2.0 * n ** 3 / t / 1e9, 1),
                        "ms": round(t * 1e3, 3)}
            del A, B
    del np
    return out


def cpu_baseline_port(oracle):
    """Fallback when no BLAS library is importable: the oracle's C
    restatement (single thread) on a bounded sample of the bandwidth set:
    gemv/ger at 8192^2, level 1 at n = 2^25 / 2^24."""
    import numpy as np
    n_mv, n_s, n_d = 8192, 1 << 25, 1 << 24
    A_s = oracle.uniform_f32(n_mv * n_mv, 1)
    A_d = oracle.uniform_f64(n_mv * n_mv, 2)
    xs, ys = oracle.uniform_f32(n_mv, 3), oracle.uniform_f32(n_mv, 4)
    xd, yd = oracle.uniform_f64(n_mv, 5), oracle.uniform_f64(n_mv, 6)
    vx, vy = oracle.uniform_f32(n_s, 7), oracle.uniform_f32(n_s, 8)
    dx, dy = oracle.uniform_f64(n_d, 9), oracle.uniform_f64(n_d, 10)
    L = oracle.lib()
    P = lambda a: a.ctypes.data
    mv_s = 4 * (n_mv * n_mv + 3 * n_mv)
    mv_d = 8 * (n_mv * n_mv + 3 * n_mv)
    r1_s = 4 * (2 * n_mv * n_mv + 2 * n_mv)
    r1_d = 8 * (2 * n_mv * n_mv + 2 * n_mv)
    jobs = [
        (mv_s, lambda: L.orc_sgemv(1, 0, n_mv, n_mv, 1.5, P(A_s), n_mv, P(xs), 1, 0.5, P(ys), 1)),
        (mv_s, lambda: L.orc_sgemv(1, 1, n_mv, n_mv, 1.5, P(A_s), n_mv, P(xs), 1, 0.5, P(ys), 1)),
        (mv_d, lambda: L.orc_dgemv(1, 0, n_mv, n_mv, 1.5, P(A_d), n_mv, P(xd), 1, 0.5, P(yd), 1)),
        (mv_d, lambda: L.orc_dgemv(1, 1, n_mv, n_mv, 1.5, P(A_d), n_mv, P(xd), 1, 0.5, P(yd), 1)),
        (r1_s, lambda: L.orc_sger(1, n_mv, n_mv, 1.5, P(xs), 1, P(ys), 1, P(A_s), n_mv)),
        (r1_d, lambda: L.orc_dger(1, n_mv, n_mv, 1.5, P(xd), 1, P(yd), 1, P(A_d), n_mv)),
        (12 * n_s, lambda: L.orc_saxpy(n_s, 1.5, P(vx), 1, P(vy), 1)),
        (8 * n_s, lambda: L.orc_sscal(n_s, 1.0, P(vy), 1)),
        (8 * n_s, lambda: L.orc_scopy(n_s, P(vx), 1, P(vy), 1)),
        (8 * n_s, lambda: L.orc_sdot(n_s, P(vx), 1, P(vy), 1)),
        (4 * n_s, lambda: L.orc_sasum(n_s, P(vx), 1)),
        (16 * n_d, lambda: L.orc_ddot(n_d, P(dx), 1, P(dy), 1)),
        (8 * n_d, lambda: L.orc_dasum(n_d, P(dx), 1)),
    ]
    total_b, total_t, passes = 0, 0.0, 0
    # whole passes over the sample until about 12 s of CPU work
    while passes < 2 or total_t < 12.0:
        for b, fn in jobs:
            t0 = time.perf_counter()
            fn()
            total_t += time.perf_counter() - t0
            total_b += b
        passes += 1
    del np
    return {"value": round(total_b / total_t / 1e9, 3), "unit": "GB/s", "cores": 1,
            "kind": "port",
            "sample": f"bandwidth set at reduced size, {passes} passes: S/DGEMV N+T "
                      "and S/DGER at 8192^2, SAXPY/SSCAL/SCOPY/SDOT/SASUM n=2^25, "
                      "DDOT/DASUM n=2^24 (oracle/refblas.c, netlib loop order)",
            "seconds": round(total_t, 2)}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    import torch
    import hcblas_amd as hc

    torch.cuda.set_device(0)
    h = hc.Handle()
    h.set_stream(torch.cuda.current_stream())
    if args.case:
        ms = run_case(hc, h, torch, args.case, args.reps)
        h.destroy()
        print(json.dumps({"case": args.case, "reps": args.reps,
                          "median_ms": round(ms, 4)}))
        return
    ops, tensors = bandwidth_ops(hc, h, torch)
    step_ms, per_op, wall = time_bandwidth(torch, ops, args.steps, args.warmup)
    # the CPU baseline runs later on host copies of the same inputs
    host = None if args.no_cpu else host_copies(tensors)
    del tensors
    total_bytes = sum(b for _, b, _ in ops)
    value = total_bytes / (step_ms * 1e-3) / 1e9
    op_gbs = {name: round(b / (t * 1e-3) / 1e9, 1) for (name, b, _), t in zip(ops, per_op)}
    dom = max(range(len(ops)), key=lambda i: per_op[i])
    dname, dbytes, _ = ops[dom]
    achieved = dbytes / (per_op[dom] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": dname, "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(dname),
                "avg_ms": round(per_op[dom], 4)}
    result = {
        "metric": "GEMM TFLOP/s split vs plain at equal err vs FP64; GEMV/dot HBM GB/s",
        "value": round(value, 1), "unit": "GB/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(step_ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32+fp64",
        "data": "synthetic uniform[-1,1], seeded, resident in HBM",
        "config": {"workload": "plain bandwidth set: S/DGEMV 16384^2 N+T a=1.5 b=0.5, "
                               "S/DGER 8192^2, SAXPY/SSCAL/SCOPY/SDOT/SASUM n=2^28, "
                               "DDOT/DASUM n=2^27", "bytes_per_step": total_bytes},
        "roofline": roofline, "ops": op_gbs,
    }
    if not args.no_gemm:
        import oracle
        del ops
        torch.cuda.empty_cache()
        result["gemm"] = gemm_section(hc, h, torch, oracle, args.gemm_n,
                                      nc=args.gemm_n // 2)
        result["convnet"] = convnet_section(hc, h, torch)
    if not args.no_gemm:
        result["grouped"] = grouped_section(hc, h, torch)
        result["adjoint"] = adjoint_section(torch)
    if not args.no_cpu:
        base = cpu_baseline_blas(host)
        del host
        if base is None:
            import oracle
            base = cpu_baseline_port(oracle)
        result["cpu_baseline"] = base
        result["cpu_gemm"] = cpu_gemm_blas()
    h.destroy()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
