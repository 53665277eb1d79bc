"""Shared test configuration.

Markers:
  gpu -- needs a real MI355X (gfx950).

CPU tests (``-m "not gpu"``) cover the oracle against the golden fixtures,
the host-side argument logic and that the C-ABI library loads and exports
every entry point of include/hcblas.h.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: requires an MI355X GPU")


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def golden():
    return load_golden


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def hc():
    import hcblas_amd
    return hcblas_amd


@pytest.fixture(scope="session")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test collected on a host without a GPU")
    return torch


@pytest.fixture()
def handle(hc, torch_cuda):
    h = hc.Handle()
    yield h
    torch_cuda.cuda.synchronize()
    if h.value:
        h.destroy()
