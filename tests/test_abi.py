"""C-ABI boundary checks that need no GPU (reference contract:
lib/include/hcblas.h, lib/src/hcblas.cpp; status codes hcblas.h:56-65).

- libhcblas.so loads and exports every entry point declared in
  include/hcblas.h, with C linkage;
- every entry point that takes a handle returns NOT_INITIALIZED for a null
  handle before touching any other argument (hcblas.cpp:1308 and every other
  routine), which is what the reference's unit-api tests assert after
  hcblasDestroy (e.g. test/unit-api/test-blas.cpp:76-79);
- the status / operation / order enum values match the reference.
"""
import ctypes
import os
import re
import subprocess

import pytest

from hcblas_amd import _abi, _lib


def test_header_parses_and_is_complete():
    protos = _abi.parse_header()
    # The reference API surface (hcblas.h) plus the explicit-order entries.
    for name in [
        "hcblasCreate", "hcblasDestroy", "hcblasSetStream", "hcblasGetStream",
        "hcblasSetVector", "hcblasGetVector", "hcblasSetMatrix",
        "hcblasGetMatrix", "hcblasSasum", "hcblasDasum", "hcblasSasumBatched",
        "hcblasDasumBatched", "hcblasSaxpy", "hcblasDaxpy",
        "hcblasSaxpyBatched", "hcblasScopy", "hcblasDcopy",
        "hcblasScopyBatched", "hcblasDcopyBatched", "hcblasSdot",
        "hcblasDdot", "hcblasSdotBatched", "hcblasDdotBatched", "hcblasSscal",
        "hcblasDscal", "hcblasCscal", "hcblasZscal", "hcblasCsscal",
        "hcblasZdscal", "hcblasSscalBatched", "hcblasDscalBatched",
        "hcblasCscalBatched", "hcblasZscalBatched", "hcblasCsscalBatched",
        "hcblasZdscalBatched", "hcblasSgemv", "hcblasSgemvBatched",
        "hcblasDgemv", "hcblasDgemvBatched", "hcblasSger", "hcblasDger",
        "hcblasSgerBatched", "hcblasDgerBatched", "hcblasSgemm",
        "hcblasDgemm", "hcblasCgemm", "hcblasZgemm", "hcblasHgemm",
        "hcblasSgemmBatched", "hcblasCgemmBatched", "hcblasDgemmBatched",
        "hcblasZgemmBatched", "hcblas_sgemm", "hcblas_dgemm", "hcblas_hgemm",
        "hcblas_cgemm", "hcblas_zgemm", "hcblas_sgemv", "hcblas_dgemv",
        "hcblas_sger", "hcblas_dger",
    ]:
        assert name in protos, name


def test_library_exports_every_declared_symbol():
    protos = _abi.parse_header()
    lib = _lib.load()
    for name in protos:
        assert hasattr(lib, name), f"libhcblas.so does not export {name}"
    # C linkage: the dynamic symbol table holds the plain names.
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(\S+)", out))
    missing = [n for n in protos if n not in exported]
    assert not missing, missing


def test_null_handle_is_not_initialized():
    lib = _lib.load()
    protos = _abi.parse_header()
    skip = {"hcblasCreate", "hcblasDestroy", "hcblasGetVersionString",
            "hcblasGroupedPlanInfo", "hcblasGroupedPlanDestroy"}
    checked = 0
    for name, proto in protos.items():
        if name in skip:
            continue
        assert proto.args[0][0] == "hcblasHandle_t", name
        args = []
        for ctype, _ in proto.args:
            t = _abi._ctype(ctype)
            args.append(None if t is ctypes.c_void_p else 1)
        st = getattr(lib, name)(*args)
        assert st == 1, f"{name}: expected NOT_INITIALIZED, got {st}"
        checked += 1
    assert checked > 60


def test_plan_entries_reject_null_plan():
    lib = _lib.load()
    assert lib.hcblasGroupedPlanInfo(None, None, None, None, None) == 3
    assert lib.hcblasGroupedPlanDestroy(None) == 3


def test_destroy_null_and_double_destroy_contract():
    lib = _lib.load()
    assert lib.hcblasDestroy(None) == 1
    h = ctypes.c_void_p(None)
    assert lib.hcblasDestroy(ctypes.byref(h)) == 1


def test_enum_values_match_reference():
    text = open(_abi.HEADER).read()
    vals = dict(re.findall(r"(HCBLAS_\w+)\s*=\s*(\d+)", text))
    assert [int(vals[f"HCBLAS_STATUS_{s}"]) for s in (
        "SUCCESS", "NOT_INITIALIZED", "ALLOC_FAILED", "INVALID_VALUE",
        "MAPPING_ERROR", "EXECUTION_FAILED", "INTERNAL_ERROR")] == list(range(7))
    assert (int(vals["HCBLAS_OP_N"]), int(vals["HCBLAS_OP_T"]),
            int(vals["HCBLAS_OP_C"])) == (0, 1, 2)
    # hcblaslib.h:47 enum hcblasOrder { RowMajor, ColMajor }
    assert (int(vals["HCBLAS_ROW_MAJOR"]), int(vals["HCBLAS_COL_MAJOR"])) == (0, 1)


def test_version_string():
    lib = _lib.load()
    buf = ctypes.create_string_buffer(64)
    assert lib.hcblasGetVersionString(buf, 64) == 0
    assert b"gfx950" in buf.value


def test_python_api_boxes_scalars():
    from hcblas_amd import api
    # Null handle: conversion runs, the library returns NOT_INITIALIZED.
    assert api.hcblasSgemm(None, 0, 0, 4, 4, 4, 1.0, 0, 4, 0, 4, 0.0, 0, 4) == 1
    st, v = api.hcblasDdot(None, 4, 0, 1, 0, 1, None)
    assert st == 1 and v == 0.0
    assert api.hcblasZgemm(None, 0, 0, 1, 1, 1, 1 + 2j, 0, 1, 0, 1, 0j, 0, 1) == 1
    with pytest.raises(TypeError):
        api.hcblasSgemm(None, 0, 0)


def test_torch_extension_registers_ops():
    """The optional libtorch layer loads against libhcblas.so and registers
    its two differentiable ops (no GPU call)."""
    import os
    from hcblas_amd import torch_ext
    if not os.path.exists(torch_ext.LIB_PATH):
        pytest.skip("libhcblas_torch.so not built")
    import torch
    torch_ext.load()
    schema = str(torch.ops.hcblas.gemm_batched.default._schema)
    assert "Tensor? C" in schema and "bool transB" in schema
    assert "float alpha" in str(torch.ops.hcblas.ger.default._schema)


def test_hipblas_shim_exports_and_null_handle():
    """The hipBLAS-compatible library exports every routine of
    include/hipblas_hcblas.h, and a null handle maps to
    HIPBLAS_STATUS_NOT_INITIALIZED (hipblas.cpp:107-126)."""
    from hcblas_amd import hipblas_shim
    lib = hipblas_shim.load()
    protos = _abi.parse_header(_abi.SHIM_HEADER, "hipblasStatus_t")
    assert len(protos) >= 50
    out = subprocess.run(["nm", "-D", "--defined-only", hipblas_shim.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(\S+)", out))
    assert not [n for n in protos if n not in exported]
    handle_less = {"hipblasCreate", "hipblasSetVector", "hipblasGetVector",
                   "hipblasSetMatrix", "hipblasGetMatrix"}
    for name, proto in protos.items():
        if name in handle_less:
            continue
        args = []
        for ctype, _ in proto.args:
            t = _abi._ctype(ctype)
            if ctype == "hipblasOperation_t":
                args.append(hipblas_shim.OP_N)
            else:
                args.append(None if t is ctypes.c_void_p else 1)
        assert getattr(lib, name)(*args) == 1, name
    # an unknown operation code is an invalid value, not an exception
    assert lib.hipblasSgemm(None, 0, hipblas_shim.OP_N, 1, 1, 1, None, None, 1,
                            None, 1, None, None, 1) == 3
    assert lib.hipblasCreate(None) == 3


def test_hcblaslib_mirror_compiles_host_only(tmp_path):
    """include/hcblaslib.hpp is host code over the C ABI: it compiles with
    g++ alone (no HIP headers), and its overloads resolve for the
    reference's call forms (a dot with a float& result, a pointer-array
    GEMM, the 'n' / 't' transpose codes)."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    src = tmp_path / "use.cpp"
    src.write_text(
        '#include "hcblaslib.hpp"\n'
        "static_assert(NoTrans == 'n' && Trans == 't', \"codes\");\n"
        "static_assert(HCBLAS_SUCCEEDS == 0 && HCBLAS_INVALID == -1, \"st\");\n"
        "static_assert((int)HCBLAS_NO_TRANS == 'n' && (int)HCBLAS_TRANS == 't',"
        " \"alias\");\n"
        "int use(Hcblaslibrary &hc, float *x, float *y, float **pa) {\n"
        "  float r = 0; hipStream_t s = hc.currentAcclView;\n"
        "  int st = hc.hcblas_sdot(s, 4, x, 1, 0, y, 1, 0, r);\n"
        "  st += hc.hcblas_sgemm(s, ColMajor, NoTrans, Trans, 2, 2, 2, 1.f, pa,"
        " 2, 0, pa, 2, 0, 0.f, pa, 2, 0, 0, 0, 0, 1);\n"
        "  st += hc.hcblas_saxpy(s, 4, 2.f, x, 1, 4, y, 1, 4, 0, 0, 3);\n"
        "  return st + (int)r;\n"
        "}\n")
    inc = os.path.join(os.path.dirname(os.path.dirname(
        os.path.abspath(__file__))), "include")
    p = subprocess.run([gxx, "-std=c++17", "-Wall", "-Werror", "-fsyntax-only",
                        "-I", inc, str(src)], capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
