"""Build the in-tree shared library afepack_amd/libafepack_amd.so (gfx950).

host.cpp (host algorithms, OpenMP) is compiled with g++; capi.hip (the C ABI
and the HIP kernels) with hipcc for gfx950 only.  Run as a script or call
build().  Output stays in the package directory so it travels with the tree.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libafepack_amd.so")
BUILD = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # no packed-FP32 instructions in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
CXX_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-fopenmp", "-Wall"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    host_o = os.path.join(BUILD, "host.o")
    capi_o = os.path.join(BUILD, "capi.o")
    _run(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, "host.cpp"), "-o", host_o])
    _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "capi.hip"), "-o", capi_o])
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, capi_o, host_o,
          "-lgomp", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
