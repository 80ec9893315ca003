"""Build libkmc.so (HIP kernels + C ABI) for gfx950, in-tree.

    python -m kvs_amd.build        # or: python kvs_amd/build.py

hipcc compiles the .hip sources for --offload-arch=gfx950 with
-ffp-contract=off (no fused multiply-add anywhere: results must round exactly
like the reference's CPU code), FP32 denormals kept, and packed FP32
instructions disabled.  Host-only C++ is compiled by g++.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
OUT = os.path.join(HERE, "libkmc.so")
BUILD = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    "-I" + INCLUDE, "-I" + CSRC,
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-I" + INCLUDE, "-I" + CSRC]

HIP_SOURCES = ["kmc_kernels.hip", "kmc_api.hip", "kmc_batch.hip", "kmc_server.hip", "kmc_noalloc.hip",
               "kmc_slice.hip", "kmc_unstructured.hip",
               "kmc_kvsml.hip", "kmc_export.hip"]
CXX_SOURCES = ["kmc_color.cpp"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in CXX_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        _run(["g++"] + CXX_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = OUT + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950"] + objs + ["-o", tmp])
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
