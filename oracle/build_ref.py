"""Build oracle/_ref/libref.so: oracle/ref_shim.cpp compiled with the three
reference source files that build without CUDA, OpenCV, Eigen or MPI
(src/backend/RNG/xorshift.cpp, Math/LeastSquareAdjust.cpp,
Laplace/src/vector_funcs.cpp), read where they lie under the reference
tree.  TEST INFRASTRUCTURE ONLY: the library is a checker for the oracle's
restatements and nothing in the product loads it.  The rest of the
reference's path (walker, bitBlock, CUDA kernels, NNLS/Tikhonov with Eigen)
does not build here, so the oracle restates it instead.

Usage: python oracle/build_ref.py [reference root, default /root/reference]
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(HERE, "_ref")
LIB = os.path.join(OUT_DIR, "libref.so")
SOURCES = ["RNG/xorshift.cpp", "Math/LeastSquareAdjust.cpp", "Laplace/src/vector_funcs.cpp"]
INCLUDES = ["RNG", "Math", "Laplace/include"]


def build(reference: str = "/root/reference") -> str | None:
    """Compile libref.so when the reference tree is present; None otherwise."""
    backend = os.path.join(reference, "src", "backend")
    srcs = [os.path.join(backend, s) for s in SOURCES]
    if not all(os.path.exists(s) for s in srcs):
        return None
    os.makedirs(OUT_DIR, exist_ok=True)
    cmd = ["g++", "-O2", "-std=c++11", "-fPIC", "-shared", "-fopenmp", "-w",
           *[f"-I{os.path.join(backend, d)}" for d in INCLUDES],
           os.path.join(HERE, "ref_shim.cpp"), *srcs, "-o", LIB + ".tmp"]
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    out = build(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    print(out if out else "reference sources not found: oracle/_ref not built")
