"""Optional libtorch layer: builds and loads libhcblas_torch.so, which
registers torch.ops.hcblas.gemm_batched and torch.ops.hcblas.ger as
torch::autograd::Function wrappers over the C ABI (source:
hcblas_amd/csrc/torch_autograd.cpp).

    from hcblas_amd import torch_ext
    torch_ext.load()
    C = torch_ext.gemm_batched(A, B, alpha=1.0)   # differentiable
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "torch_autograd.cpp")
LIB_DIR = os.path.join(HERE, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libhcblas_torch.so")

_loaded = False


def build(verbose=False):
    """Compiles the extension with g++ against the installed torch (host code
    only: every GPU kernel lives in libhcblas.so)."""
    import torch
    from torch.utils import cpp_extension as ce
    tdir = os.path.dirname(torch.__file__)
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared",
           "-D__HIP_PLATFORM_AMD__=1", "-DUSE_ROCM=1",
           f"-D_GLIBCXX_USE_CXX11_ABI={abi}"]
    for inc in ce.include_paths():
        cmd += ["-isystem", inc]
    cmd += ["-isystem", "/opt/rocm/include", SRC,
            "-L" + os.path.join(tdir, "lib"), "-ltorch", "-ltorch_cpu",
            "-ltorch_hip", "-lc10", "-lc10_hip",
            "-L" + LIB_DIR, "-lhcblas",
            "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + os.path.join(tdir, "lib"),
            "-o", LIB_PATH + ".tmp"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(LIB_PATH + ".tmp", LIB_PATH)


def load():
    """Loads the prebuilt extension; raises if it has not been built."""
    global _loaded
    if _loaded:
        return
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with hcblas_amd.torch_ext.build()")
    import torch
    from . import _lib
    _lib.load()  # libhcblas.so first, so the extension binds to it
    torch.ops.load_library(LIB_PATH)
    _loaded = True


def gemm_batched(A, B, C=None, alpha=1.0, beta=0.0, transA=False,
                 transB=False):
    """C_out[b] = alpha * op(A[b]) @ op(B[b]) + beta * C[b] (row-major,
    A / B float32, float64 or bfloat16; the output is float32 for bfloat16
    inputs). Differentiable in A, B and C."""
    import torch
    load()
    return torch.ops.hcblas.gemm_batched(A, B, C, float(alpha), float(beta),
                                         bool(transA), bool(transB))


def ger(x, y, A, alpha=1.0):
    """A_out = alpha * outer(x, y) + A (row-major, float32 / float64).
    Differentiable in x, y and A."""
    import torch
    load()
    return torch.ops.hcblas.ger(x, y, A, float(alpha))
