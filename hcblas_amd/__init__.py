"""hcblas_amd: MI355X-native (gfx950) drop-in for the rocmarchive/hcBLAS API.

The product is the C-ABI shared library ``hcblas_amd/lib/libhcblas.so``
(declared in ``include/hcblas.h``); this package only binds it for Python
callers and tests. Importing it loads the library and fails loudly if it is
not built -- there is no CPU fallback.
"""
from ._lib import LIB_PATH, HcblasLibraryMissing, load
from .api import (C_32F, C_64F, COL_MAJOR, DEFAULT_MATH, OP_C, OP_N, OP_T,
                  OZAKI_INT8_MATH, R_16BF, R_16F, R_32F, R_64F, ROW_MAJOR,
                  SPLIT_BF16_MATH, GemmProblem, GroupedPlan, Handle,
                  HcblasError, Status, api, check, half_bits)

__all__ = [
    "LIB_PATH", "HcblasLibraryMissing", "load", "Handle", "HcblasError",
    "GemmProblem", "GroupedPlan",
    "Status", "api", "check", "half_bits", "OP_N", "OP_T", "OP_C",
    "ROW_MAJOR", "COL_MAJOR", "DEFAULT_MATH", "SPLIT_BF16_MATH",
    "OZAKI_INT8_MATH", "R_32F", "R_64F", "C_32F", "C_64F", "R_16F", "R_16BF",
]

load()
