"""Extracts the SGEMM shape corpus of the reference's convolution-network
benchmark (rocmarchive/hcBLAS benchmark/BLAS_benchmark_Convolution_Networks/
*.txt, read by test/src/sgemm_cn_test.cpp:42-56) into
tests/golden/convnet_shapes.json: one record per distinct call shape
(M, N, K, transA, transB, lda, ldb, ldc, alpha, beta) with the largest
aOffset / bOffset / cOffset the corpus uses with it and the list it comes
from. The call is column-major (sgemm_cn_test.cpp:61) with the offsets added
to the operand base pointers (:119-123).

Usage (in a container that has the reference checkout):
    python tests/golden/gen_convnet_shapes.py [/path/to/reference]
"""
from __future__ import annotations

import glob
import json
import os
import sys

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                   "convnet_shapes.json")


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    d = os.path.join(ref, "benchmark", "BLAS_benchmark_Convolution_Networks")
    shapes = {}
    for path in sorted(glob.glob(os.path.join(d, "*.txt"))):
        name = os.path.basename(path)
        if name.startswith("Input"):
            continue
        for line in open(path):
            p = line.split()
            if len(p) != 13:
                continue
            M, N, K, ta, tb, lda, ldb, ldc = (int(x) for x in p[:8])
            alpha, beta = float(p[8]), float(p[9])
            ao, bo, co = (int(x) for x in p[10:13])
            key = (M, N, K, ta, tb, lda, ldb, ldc, alpha, beta)
            rec = shapes.get(key)
            if rec is None:
                shapes[key] = [ao, bo, co, name[:-4]]
            else:
                rec[0], rec[1], rec[2] = max(rec[0], ao), max(rec[1], bo), max(rec[2], co)
    rows = [list(k) + v for k, v in shapes.items()]
    doc = {"source": "benchmark/BLAS_benchmark_Convolution_Networks/*.txt",
           "columns": ["M", "N", "K", "transA", "transB", "lda", "ldb", "ldc",
                       "alpha", "beta", "aOffset_max", "bOffset_max",
                       "cOffset_max", "list"],
           "order": "column-major",
           "shapes": rows}
    with open(OUT, "w") as f:
        json.dump(doc, f, separators=(",", ":"))
    print(f"{len(rows)} shapes -> {OUT}")


if __name__ == "__main__":
    main()
