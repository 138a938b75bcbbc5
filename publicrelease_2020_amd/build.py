"""Compile the HIP library in-tree for gfx950: publicrelease_2020_amd/libflosic_hip.so."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libflosic_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
SOURCES = ["flosic_kernels.hip", "flosic_api.hip"]


def command(out: str = LIB) -> list:
    return [os.path.join(ROCM, "bin", "hipcc"), "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
            # packed FP32 instructions are kept out of all device code
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument",
            "-I" + os.path.join(ROCM, "include"), *[os.path.join(CSRC, s) for s in SOURCES],
            "-L" + os.path.join(ROCM, "lib"), "-lrocblas", "-Wl,-rpath," + os.path.join(ROCM, "lib"), "-o", out]


def build(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd))
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libflosic_hip.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
