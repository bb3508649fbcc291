"""Build librwnmr_amd.so in-tree: hipcc for the gfx950 kernels and the C ABI,
g++ for the host-only C++ (Laplace inversion, drivers, configs).

Usage: python -m randomwalk_nmr_amd.build   (or build() from Python).
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librwnmr_amd.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["walk_kernels.hip", "zwalk_kernels.hip", "batch_kernels.hip", "backward_kernels.hip",
               "locality_kernels.hip", "capi.hip", "drivers.hip", "probe_kernels.hip"]
CXX_SOURCES = ["laplace.cpp", "host_sim.cpp", "image_io.cpp", "histogram.cpp", "config.cpp", "output.cpp"]
# the command-line program, linked against the library
APP_SOURCE = "rwnmr_app.cpp"
APP = os.path.join(HERE, "rwnmr")

HIP_FLAGS = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
             # packed FP32 instructions are kept out of every kernel
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-Wall", "-Wextra", "-Wno-unused-parameter"]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(cmd)}\n{out}")
    return out


# The walk kernels reload image chunks with the indexed buffer load, bound by
# its LLVM intrinsic name (clang has a builtin only for the raw form).  The
# probe compiles that binding; a toolchain without it builds the raw form.
_PROBE_SRC = """#include <hip/hip_runtime.h>
typedef unsigned int v4u __attribute__((ext_vector_type(4)));
typedef unsigned int v2u __attribute__((ext_vector_type(2)));
__device__ v4u load4(__amdgpu_buffer_rsrc_t r, int vindex, int voffset, int soffset, int aux)
    __asm("llvm.amdgcn.struct.ptr.buffer.load.v4i32");
__device__ v2u load2(__amdgpu_buffer_rsrc_t r, int vindex, int voffset, int soffset, int aux)
    __asm("llvm.amdgcn.struct.ptr.buffer.load.v2i32");
__global__ void probe(unsigned int *w, v4u *out, int i)
{
    __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(w, (short)16, 64, 0x00020000);
    __amdgpu_buffer_rsrc_t r8 = __builtin_amdgcn_make_buffer_rsrc(w, (short)8, 128, 0x00020000);
    v2u h = load2(r8, i, 0, 0, 0);
    out[threadIdx.x] = load4(r, i, 0, 0, 0) + (v4u){h.x, h.y, 0u, 0u};
}
"""


def _indexed_buffer_load_flags():
    """[] when the indexed buffer load compiles, else the raw-load define."""
    src = os.path.join(OUT_DIR, "probe_indexed_load.hip")
    with open(src, "w") as f:
        f.write(_PROBE_SRC)
    p = subprocess.run([HIPCC, f"--offload-arch={ARCH}", "-O3", "--cuda-device-only", "-c", src, "-o",
                        os.path.join(OUT_DIR, "probe_indexed_load.o")], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True)
    return [] if p.returncode == 0 else ["-DRWNMR_RAW_BUFFER_LOAD"]


def build(verbose: bool = False, jobs: int = 0) -> str:
    """Compile every source (always, `jobs` compilers at once; 0: up to 8) and
    link the shared library."""
    os.makedirs(OUT_DIR, exist_ok=True)
    load_flags = _indexed_buffer_load_flags()
    if verbose and load_flags:
        print("indexed buffer load unavailable: building the raw-load form")
    cmds, objs = [], []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(OUT_DIR, src + ".o")
        cmds.append([HIPCC, *HIP_FLAGS, *load_flags, "-c", path, "-o", obj])
        objs.append(obj)
    for src in CXX_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(OUT_DIR, src + ".o")
        cmds.append(["g++", *CXX_FLAGS, "-I", os.path.join(HERE, "..", "include"), "-c", path, "-o", obj])
        objs.append(obj)
    workers = jobs or max(1, min(8, os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=workers) as pool:
        for out in pool.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-fopenmp", "-lz"])
    os.replace(tmp, LIB)
    tmp = APP + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-I", os.path.join(HERE, "..", "include"),
          os.path.join(CSRC, APP_SOURCE), "-o", tmp, "-L", HERE, "-lrwnmr_amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, APP)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
