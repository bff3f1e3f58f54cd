"""Build libjja.so in-tree with hipcc for gfx950.

The shared library holds the HIP kernels, the device plan and the host-side
Cholesky factorisation. It is built next to this file so that it travels with the
source tree; nothing is installed elsewhere.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libjja.so")

HIPCC_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-fPIC", "-shared", "-std=c++17",
    # the step expressions keep the reference's operation order: no FMA contraction
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def sources():
    return [os.path.join(CSRC, "jja_evolve.hip"), os.path.join(CSRC, "jja_factor.cpp")]


def build_library(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = LIB + ".tmp"
    cmd = [hipcc, *HIPCC_FLAGS, *sources(), "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_library(verbose=True))
