"""Device-memory helpers for the GPU tests (torch is only the allocator)."""
import numpy as np
import torch

_TORCH = {
    np.dtype(np.float32): torch.float32,
    np.dtype(np.float64): torch.float64,
    np.dtype(np.complex64): torch.complex64,
    np.dtype(np.complex128): torch.complex128,
    np.dtype(np.float16): torch.float16,
    np.dtype(np.uint16): torch.int16,
    np.dtype(np.int8): torch.int8,
}


def to_dev(a):
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint16:
        t = torch.from_numpy(a.view(np.int16))
    else:
        t = torch.from_numpy(a)
    return t.to("cuda")


def to_host(t, dtype=None):
    torch.cuda.synchronize()
    a = t.cpu().numpy()
    if dtype is not None and a.dtype != dtype:
        a = a.view(dtype)
    return a


def empty(n, dtype):
    return torch.empty(int(n), dtype=_TORCH[np.dtype(dtype)], device="cuda")


def bitwise_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(
        a.view(np.uint8), b.view(np.uint8))
