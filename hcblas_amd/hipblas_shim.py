"""ctypes binding of the hipBLAS-compatible library
(``hcblas_amd/lib/libhcblas_hipblas.so``, declared in
``include/hipblas_hcblas.h``). Loads libhcblas first so both share one HIP
runtime and one copy of the library; fails loudly if it is not built."""
from __future__ import annotations

import ctypes
import os

from . import _abi, _lib

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib",
                        "libhcblas_hipblas.so")
OP_N, OP_T, OP_C = 111, 112, 113

_shim = None


def load():
    global _shim
    if _shim is not None:
        return _shim
    _lib.load()
    if not os.path.exists(LIB_PATH):
        raise _lib.HcblasLibraryMissing(
            f"{LIB_PATH} is not built; run `make -C hcblas_amd/csrc`")
    lib = ctypes.CDLL(LIB_PATH)
    for name, proto in _abi.parse_header(_abi.SHIM_HEADER,
                                         "hipblasStatus_t").items():
        fn = getattr(lib, name)
        fn.argtypes = _abi.argtypes(proto)
        fn.restype = ctypes.c_int
    _shim = lib
    return lib
