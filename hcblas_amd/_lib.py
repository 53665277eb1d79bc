"""Loader for the gfx950 library ``hcblas_amd/lib/libhcblas.so``.

There is no fallback: if the shared library is missing or does not export an
entry point of ``include/hcblas.h``, importing fails loudly.
"""
from __future__ import annotations

import ctypes
import os

from . import _abi

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib",
                        "libhcblas.so")

_lib = None
_protos = None


class HcblasLibraryMissing(ImportError):
    pass


def load():
    """Returns the ctypes handle of libhcblas with every prototype bound."""
    global _lib, _protos
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HcblasLibraryMissing(
            f"{LIB_PATH} is not built; run `make -C hcblas_amd/csrc` or "
            "`python -c 'import __graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    protos = _abi.parse_header()
    for name, proto in protos.items():
        fn = getattr(lib, name)  # AttributeError = missing export
        fn.argtypes = _abi.argtypes(proto)
        fn.restype = ctypes.c_int
    _lib, _protos = lib, protos
    return lib


def prototypes():
    load()
    return _protos
