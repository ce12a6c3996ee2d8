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


# The sanitizer build instruments the host code only (table builders, readers,
# the C ABI's argument checks): ASan and UBSan with ROCm clang's shared
# runtime, each flag given to the host compilation alone.
ASAN_DIR = os.path.join(HERE, "_asan")
SAN_HOST = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer",
            "-shared-libasan"]


def build(verbose: bool = False, sanitize: bool = False) -> str:
    """Builds libhamstr_hip.so and the drop-in driver; with sanitize, the
    host-sanitized copies in _asan/ instead (used by tests/test_sanitizers.py)."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    out_dir = ASAN_DIR if sanitize else HERE
    os.makedirs(out_dir, exist_ok=True)
    lib = os.path.join(out_dir, "libhamstr_hip.so")
    driver = os.path.join(out_dir, "hamstr_mi355x")
    cmd = [hipcc, *HIPCC_FLAGS, *(SAN_HOST if sanitize else []), "-I", os.path.join(ROOT, "include"), *SRCS,
           "-o", lib + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(lib + ".tmp", lib)
    # the drop-in driver: plain C++ over the C ABI
    if sanitize:
        cxx = [os.environ.get("CLANGXX", "/opt/rocm/llvm/bin/clang++"), "-O1", "-g", "-fno-omit-frame-pointer",
               "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-shared-libasan"]
    else:
        cxx = [os.environ.get("CXX", "g++"), "-O2"]
    cmd = [*cxx, "-std=c++17", "-I", os.path.join(ROOT, "include"), DRIVER_SRC, "-o", driver + ".tmp",
           "-L", out_dir, "-lhamstr_hip", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(driver + ".tmp", driver)
    return lib


if __name__ == "__main__":
    print(build(verbose=True, sanitize="--sanitize" in sys.argv))
    sys.exit(0)
