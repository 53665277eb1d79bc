"""Runs the C++ port of the reference's lib-level unit tests
(tests/cpp/test_hcblaslib.cpp) on the GPU: the Hcblaslibrary overloads of
include/hcblaslib.hpp (offsets, batch offsets, 'n'/'t' transposes, the
reference's HCBLAS_SUCCEEDS / HCBLAS_INVALID statuses) against the CPU
oracle, exact on the reference's integer recipes (test/unit/hcblas_*.cpp).
"""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin", "test_hcblaslib")


def test_reference_lib_level_suite():
    assert os.path.exists(BIN), f"{BIN} is not built (make -C tests/cpp)"
    p = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(p.stdout[-4000:])
    assert p.returncode == 0, p.stdout[-4000:] + p.stderr[-2000:]
    assert "0 failed" in p.stdout
    assert "FAIL" not in p.stdout
