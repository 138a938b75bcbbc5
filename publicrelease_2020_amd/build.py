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


FORTRAN_EXE = os.path.join(HERE, "flosic_sic_drive")


def build_fortran(verbose: bool = False) -> str:
    """The Fortran host driver (include/flosic_hip.f90 + fortran/*.f90) linked against libflosic_hip.so."""
    fc = os.path.join(ROCM, "bin", "amdflang")
    mod = os.path.join(HERE, "_build", "fmod")
    os.makedirs(mod, exist_ok=True)
    inc = os.path.join(os.path.dirname(HERE), "include", "flosic_hip.f90")
    steps = [
        [fc, "-O2", "-J" + mod, "-c", inc, "-o", os.path.join(mod, "flosic_hip.o")],
        [fc, "-O2", "-J" + mod, "-I" + mod, "-c", os.path.join(HERE, "fortran", "flosic_sic_block.f90"), "-o",
         os.path.join(mod, "flosic_sic_block.o")],
        [fc, "-O2", "-I" + mod, os.path.join(HERE, "fortran", "flosic_sic_drive.f90"),
         os.path.join(mod, "flosic_sic_block.o"), os.path.join(mod, "flosic_hip.o"), "-L" + HERE, "-lflosic_hip",
         "-Wl,-rpath,$ORIGIN", "-o", FORTRAN_EXE + ".tmp"],
    ]
    for cmd in steps:
        if verbose:
            print(" ".join(cmd))
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError("amdflang failed building the Fortran host driver")
    os.replace(FORTRAN_EXE + ".tmp", FORTRAN_EXE)
    return FORTRAN_EXE


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
    if os.path.exists(os.path.join(ROCM, "bin", "amdflang")):
        print(build_fortran(verbose=True))
