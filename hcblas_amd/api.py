"""Python mirror of the hcBLAS C API over torch device tensors.

Every entry point of ``include/hcblas.h`` is reachable as an attribute of this
module with the reference's name and argument order (lib/include/hcblas.h,
lib/include/hcblaslib.h), e.g.::

    h = Handle()
    st = api.hcblasSgemm(h, OP_N, OP_N, m, n, k, 1.0, A, m, B, k, 0.0, C, m)
    st, r = api.hcblasSdot(h, n, x, 1, y, 1, None)

Conversions (driven by the C prototypes parsed from the header):
  - a torch tensor for a pointer parameter passes its device address;
  - a Python number for ``alpha`` / ``beta`` is boxed as the host scalar of the
    prototype's type (float, double, complex, binary16, or the type named by a
    preceding ``dtype`` argument for the ``void *`` forms);
  - ``None`` for ``result`` allocates the host output scalar, and the call then
    returns ``(status, value)``;
  - a :class:`Handle` passes its context pointer.
The return value is the integer hcblasStatus_t (it is not raised), so status
contracts can be tested exactly as the reference's unit-api tests do.
"""
from __future__ import annotations

import ctypes
import enum
import struct

from . import _lib


class Status(enum.IntEnum):
    SUCCESS = 0
    NOT_INITIALIZED = 1
    ALLOC_FAILED = 2
    INVALID_VALUE = 3
    MAPPING_ERROR = 4
    EXECUTION_FAILED = 5
    INTERNAL_ERROR = 6


OP_N, OP_T, OP_C = 0, 1, 2
ROW_MAJOR, COL_MAJOR = 0, 1
DEFAULT_MATH, SPLIT_BF16_MATH, OZAKI_INT8_MATH = 0, 1, 2
R_32F, R_64F, C_32F, C_64F, R_16F, R_16BF = 0, 1, 2, 3, 4, 5


class GemmProblem(ctypes.Structure):
    """hcblasGemmProblem of include/hcblas.h (one entry of a grouped GEMM)."""
    _fields_ = [("dtype", ctypes.c_int32), ("order", ctypes.c_int32),
                ("transA", ctypes.c_int32), ("transB", ctypes.c_int32),
                ("M", ctypes.c_int32), ("N", ctypes.c_int32),
                ("K", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("alpha", ctypes.c_float), ("beta", ctypes.c_float),
                ("A", ctypes.c_void_p), ("lda", ctypes.c_int64),
                ("B", ctypes.c_void_p), ("ldb", ctypes.c_int64),
                ("C", ctypes.c_void_p), ("ldc", ctypes.c_int64)]


class GroupedPlan:
    """Owns an hcblasGroupedPlan_t built from a list of GemmProblem."""

    def __init__(self, handle, problems):
        lib = _lib.load()
        self._table = (GemmProblem * len(problems))(*problems)
        self._p = ctypes.c_void_p()
        st = lib.hcblasGroupedPlanCreate(handle.value, self._table,
                                         len(problems), ctypes.byref(self._p))
        if st != Status.SUCCESS:
            raise HcblasError("hcblasGroupedPlanCreate", st)
        self.handle = handle

    @property
    def value(self):
        return self._p.value

    def run(self):
        return _lib.load().hcblasGroupedGemm(self.handle.value, self._p)

    def autotune(self, reps=3):
        return _lib.load().hcblasGroupedPlanAutotune(self.handle.value,
                                                     self._p, reps)

    def info(self):
        vals = [ctypes.c_int() for _ in range(4)]
        st = _lib.load().hcblasGroupedPlanInfo(self._p, *[ctypes.byref(v) for v in vals])
        if st != Status.SUCCESS:
            raise HcblasError("hcblasGroupedPlanInfo", st)
        return dict(zip(("tile", "order", "tiles", "workgroups"),
                        (v.value for v in vals)))

    def destroy(self):
        if self._p.value:
            _lib.load().hcblasGroupedPlanDestroy(self._p)
            self._p = ctypes.c_void_p()


class HcblasError(RuntimeError):
    def __init__(self, name, status):
        super().__init__(f"{name} returned {Status(status).name}")
        self.status = status


class Handle:
    """Owns an hcblasHandle_t (hcblasCreate / hcblasDestroy)."""

    def __init__(self):
        lib = _lib.load()
        self._h = ctypes.c_void_p()
        st = lib.hcblasCreate(ctypes.byref(self._h))
        if st != Status.SUCCESS:
            raise HcblasError("hcblasCreate", st)

    @property
    def value(self):
        return self._h.value

    def ref(self):
        """Pointer to the handle variable (for hcblasDestroy)."""
        return ctypes.byref(self._h)

    def destroy(self):
        return _lib.load().hcblasDestroy(ctypes.byref(self._h))

    def set_stream(self, stream):
        s = stream if isinstance(stream, int) else getattr(stream, "cuda_stream", 0)
        st = _lib.load().hcblasSetStream(self._h, ctypes.c_void_p(s))
        if st != Status.SUCCESS:
            raise HcblasError("hcblasSetStream", st)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        if self._h.value:
            self.destroy()


def half_bits(x: float) -> int:
    return struct.unpack("<H", struct.pack("<e", float(x)))[0]


_SCALAR_BY_TYPE = {
    "float": "f",
    "double": "d",
    "hcComplex": "c",
    "hcDoubleComplex": "z",
    "hcblasHalf": "h",
}
_SCALAR_BY_DTYPE = {R_32F: "f", R_64F: "d", C_32F: "c", C_64F: "z",
                    R_16F: "f", R_16BF: "f"}


def _box(kind, value):
    if kind == "f":
        return ctypes.c_float(float(value))
    if kind == "d":
        return ctypes.c_double(float(value))
    if kind == "c":
        v = complex(value)
        return (ctypes.c_float * 2)(v.real, v.imag)
    if kind == "z":
        v = complex(value)
        return (ctypes.c_double * 2)(v.real, v.imag)
    if kind == "h":
        return ctypes.c_uint16(half_bits(value))
    raise ValueError(kind)


def _unbox(kind, box):
    if kind in ("f", "d"):
        return box.value
    if kind in ("c", "z"):
        return complex(box[0], box[1])
    raise ValueError(kind)


def _base_type(ctype: str) -> str:
    return ctype.replace("const", "").replace("*", "").replace("[]", "").strip()


def _convert(name, proto, args):
    if len(args) != len(proto.args):
        raise TypeError(f"{name} takes {len(proto.args)} arguments, got {len(args)}")
    out = []
    keep = []
    result_box = None
    dtype = None
    for (ctype, pname), a in zip(proto.args, args):
        if pname == "dtype":
            dtype = int(a)
        if isinstance(a, Handle):
            out.append(a.value)
            continue
        if pname in ("alpha", "beta") and isinstance(a, (int, float, complex)):
            bt = _base_type(ctype)
            kind = _SCALAR_BY_DTYPE[dtype] if bt == "void" else _SCALAR_BY_TYPE[bt]
            box = _box(kind, a)
            keep.append(box)
            out.append(ctypes.addressof(box))
            continue
        if pname == "result" and a is None:
            kind = _SCALAR_BY_TYPE[_base_type(ctype)]
            box = _box(kind, 0)
            result_box = (kind, box)
            out.append(ctypes.addressof(box))
            continue
        if hasattr(a, "data_ptr"):
            out.append(a.data_ptr())
            continue
        out.append(a)
    return out, keep, result_box


class _Api:
    def __getattr__(self, name):
        protos = _lib.prototypes()
        if name not in protos:
            raise AttributeError(name)
        proto = protos[name]
        fn = getattr(_lib.load(), name)

        def call(*args):
            cargs, keep, result_box = _convert(name, proto, args)
            st = fn(*cargs)
            del keep
            if result_box is not None:
                return st, _unbox(*result_box)
            return st

        call.__name__ = name
        return call


api = _Api()


def check(name, status):
    if status != Status.SUCCESS:
        raise HcblasError(name, status)
