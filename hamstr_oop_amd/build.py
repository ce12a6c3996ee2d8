"""Builds the gfx950 shared library libhamstr_hip.so in-tree with hipcc."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRCS = [os.path.join(HERE, "csrc", "hamstr_hip.hip"), os.path.join(HERE, "csrc", "adi_long.hip"),
        os.path.join(HERE, "csrc", "mlp_hip.hip")]
LIB = os.path.join(HERE, "libhamstr_hip.so")
DRIVER_SRC = os.path.join(HERE, "csrc", "hamstr_main.cpp")
DRIVER = os.path.join(HERE, "hamstr_mi355x")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # the reference's arithmetic is x86-64 SSE2 without FMA: no contraction
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    cmd = [hipcc, *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), *SRCS, "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    # the drop-in driver: plain C++ over the C ABI
    cxx = os.environ.get("CXX", "g++")
    cmd = [cxx, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), DRIVER_SRC, "-o", DRIVER + ".tmp",
           "-L", HERE, "-lhamstr_hip", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(DRIVER + ".tmp", DRIVER)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    print(LIB)
    sys.exit(0)
