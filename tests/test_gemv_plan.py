"""The GEMV launch plan (hcblas_amd/csrc/gemv_plan.hpp, host code) on the
CPU: the kernels and split counts of the bandwidth-set shapes, and the
split-partial bytes that the allocation-free GEMV entry compares with its
fixed scratch (HCBLAS_GEMV_NOALLOC_SCRATCH_BYTES): every shape of the
bandwidth set fits, and a plan whose partials do not fit is what makes that
entry return INVALID_VALUE before any launch."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCRATCH = 16 << 20
KERNELS = ["bandn", "panel", "band16", "band8", "column"]

PROBE = r'''
#include <cstdio>
#include <cstdlib>
#include "gemv_plan.hpp"
int main(int argc, char **argv) {
  const int cus = atoi(argv[1]);
  const bool panel = atoi(argv[2]) != 0, vec = atoi(argv[5]) != 0;
  const long long R = atoll(argv[3]), Cc = atoll(argv[4]);
  const int vn = atoi(argv[6]);
  const size_t es = (size_t)atoi(argv[7]);
  const long long batch = atoll(argv[8]);
  hcb::GemvPlan p = hcb::plan_gemv(cus, panel, R, Cc, vec, vn, es, batch);
  std::printf("%d %lld %d %zu\n", (int)p.kernel, (long long)p.oblocks,
              p.splits, p.partial_bytes);
  return 0;
}
'''


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("plan")
    src, exe = d / "probe.cpp", d / "probe"
    src.write_text(PROBE)
    subprocess.run([gxx, "-std=c++17", "-O1", "-Wall", "-Werror", "-I",
                    os.path.join(ROOT, "hcblas_amd", "csrc"), str(src), "-o",
                    str(exe)], check=True)

    def run(panel, R, Cc, vec, vn, es, batch=1, cus=256):
        out = subprocess.run([str(exe), str(cus), str(int(panel)), str(R),
                              str(Cc), str(int(vec)), str(vn), str(es),
                              str(batch)], capture_output=True, text=True,
                             check=True).stdout.split()
        return KERNELS[int(out[0])], int(out[1]), int(out[2]), int(out[3])
    return run


def test_bandwidth_set_plans(probe):
    # column-major 16384^2: N -> panel (y has R entries), T -> column case
    assert probe(True, 16384, 16384, True, 4, 4) == ("bandn", 4, 64, 4 << 20)
    assert probe(True, 16384, 16384, True, 2, 8) == ("bandn", 8, 32, 4 << 20)
    k, ob, sp, pb = probe(False, 16384, 16384, True, 4, 4)
    assert k == "band16" and pb == (0 if sp == 1 else 4 * sp * 16384)
    k, ob, sp, pb = probe(False, 16384, 16384, True, 2, 8)
    assert k == "band16" and sp == 2 and pb == 2 * 8 * 16384


@pytest.mark.parametrize("panel", [True, False])
@pytest.mark.parametrize("vec,vn,es", [(True, 4, 4), (True, 2, 8),
                                       (False, 1, 4), (False, 1, 8)])
def test_single_problems_up_to_16_gib_fit_the_scratch(probe, panel, vec, vn,
                                                      es):
    """Every single (unbatched) GEMV with up to 16 GiB of A, over a sweep of
    aspect ratios, needs at most the fixed scratch."""
    total = (16 << 30) // es
    for lr in range(4, 31):
        R = 1 << lr
        Cc = max(1, total // R)
        for (r, c) in ((R, Cc), (R - 3, Cc + 5), (R + 1, max(1, Cc - 1))):
            pb = probe(panel, r, c, vec, vn, es)[3]
            assert pb <= SCRATCH, (panel, r, c, vec, pb)


def test_oversized_plan_is_rejected_size(probe):
    """A 256 GiB FP64 matrix read along its rows (band kernel, 4096 bands
    of 1024 columns): 32 MiB of partials, more than the scratch."""
    k, ob, sp, pb = probe(False, 33554432, 1024, True, 2, 8)
    assert k == "band16" and sp == 4096 and pb == 32 << 20 and pb > SCRATCH
