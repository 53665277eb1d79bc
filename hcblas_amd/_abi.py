"""C-ABI description of libhcblas, read from ``include/hcblas.h``.

The header is the single source of truth for the exported entry points: the
ctypes binding (``hcblas_amd._lib``) and the ABI test both take the list of
functions and their argument types from it, so the two cannot drift apart.
"""
from __future__ import annotations

import ctypes
import os
import re
from dataclasses import dataclass
from typing import Dict, List, Tuple

REPO_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO_ROOT, "include", "hcblas.h")

_ENUMS = {
    "hcblasStatus_t",
    "hcblasOperation_t",
    "hcblasOrder_t",
    "hcblasMathMode_t",
    "hcblasDatatype_t",
    "hipblasOperation_t",
}
SHIM_HEADER = os.path.join(REPO_ROOT, "include", "hipblas_hcblas.h")


@dataclass(frozen=True)
class Prototype:
    name: str
    args: Tuple[Tuple[str, str], ...]  # (C type, parameter name)


def _ctype(ctype: str):
    t = ctype.replace("const", " ").strip()
    t = re.sub(r"\s+", " ", t)
    if "*" in t or "[]" in t:
        return ctypes.c_void_p
    t = t.strip()
    if t in ("int",) or t in _ENUMS:
        return ctypes.c_int
    if t == "int64_t":
        return ctypes.c_int64
    if t in ("hcblasHandle_t", "hipStream_t", "hcblasGroupedPlan_t",
             "hipblasHandle_t"):
        return ctypes.c_void_p
    raise ValueError(f"unmapped C type: {ctype!r}")


def parse_header(path: str = HEADER,
                 ret: str = "hcblasStatus_t") -> Dict[str, Prototype]:
    """Returns every ``<ret> name(args);`` declared in the header
    (``hipblasStatus_t`` for the hipBLAS-compatible header)."""
    with open(path, "r", encoding="utf-8") as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    protos: Dict[str, Prototype] = {}
    for m in re.finditer(ret + r"\s+(\w+)\s*\(([^;{]*?)\)\s*;", text):
        name, argtext = m.group(1), m.group(2)
        args: List[Tuple[str, str]] = []
        for raw in argtext.split(","):
            raw = " ".join(raw.split())
            if not raw or raw == "void":
                continue
            am = re.match(r"(.*?)(\w+)\s*(\[\])?$", raw)
            if am is None:
                raise ValueError(f"cannot parse argument {raw!r} of {name}")
            ctype = am.group(1).strip() + ("[]" if am.group(3) else "")
            args.append((ctype, am.group(2)))
        protos[name] = Prototype(name, tuple(args))
    return protos


def argtypes(proto: Prototype):
    return [_ctype(t) for t, _ in proto.args]
