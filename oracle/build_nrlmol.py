"""Build the reference program (nrlmol_exe, serial) from the reference sources where they lie, into
oracle/_ref/nrlmol/ (test infrastructure: it writes golden fixtures in this container and never runs elsewhere).

The object list is read from the reference's Makefile.fedora as data (its OBJ variable and the lists it names); the
reference's build system itself is not run.  Settings of a serial build: no -DMPI, -DNOCOMM as the Makefile's
OPTIONS, ROCm's amdflang for Fortran (gfortran is not in this image), gcc for the C files, and the BLAS/LAPACK the
Makefile links (MKL from /opt/conda/lib here).  Module-defining files are compiled first in the Makefile's order,
the rest in parallel.

    python oracle/build_nrlmol.py [-j N] [--dropin]

--dropin also links oracle/_ref/nrlmol/nrlmol_flosic_hip: the same objects with apotnl_sic.o, frmorb2.o, frmiorb.o,
frmorb_mpi.o and diagv_fo.o replaced by the library's Fortran drop-ins (publicrelease_2020_amd/fortran/flosic_dropin.F90,
compiled against the reference's module files) and libflosic_hip.so: the reference program with its per-FLO SIC
block on the GPU.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import re
import subprocess
import sys

REF = os.environ.get("FLOSIC_REF", "/root/reference/flosic")
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref", "nrlmol")
OBJ = os.path.join(OUT, "obj")
EXE = os.path.join(OUT, "nrlmol_exe")
DROPIN_EXE = os.path.join(OUT, "nrlmol_flosic_hip")
REPO = os.path.dirname(HERE)
PKG = os.path.join(REPO, "publicrelease_2020_amd")
REPLACED = ["apotnl_sic.o", "frmorb2.o", "frmiorb.o", "frmorb_mpi.o", "diagv_fo.o"]
FC = "/opt/rocm/bin/amdflang"
MKL = "/opt/conda/lib"
LISTS = ["FTNBASICS", "FBASICS", "FBASICSMPI", "LOCALSIC", "METAGGA", "SIC", "SIC2", "SIC3", "EXTRAFIL",
         "POSTACTIVE", "POSTACTIVEMPI", "BASIS", "DISP", "GIT", "WFX", "DMAT", "POPULATION", "LEVY_PERDEW"]
MODULE_FILES = ["mod_periodic_table", "modules_sic", "modules", "modules2", "new_input_reader", "fodmat_all"]


def object_list() -> list:
    txt = open(os.path.join(REF, "Makefile.fedora")).read()
    txt = re.sub(r"\\\n", " ", txt)
    var = {}
    for line in txt.splitlines():
        m = re.match(r"^([A-Z0-9_]+)\s*=\s*(.*)$", line)
        if m and m.group(1) not in var:
            var[m.group(1)] = m.group(2).split("#")[0].split()
    objs = []
    for name in LISTS:
        for o in var.get(name, []):
            if o.endswith(".o") and o not in objs:
                objs.append(o)
    return objs


def source_of(obj: str) -> str:
    stem = obj[:-2]
    for ext in (".F90", ".f90", ".F", ".f", ".ftn", ".c"):
        p = os.path.join(REF, stem + ext)
        if os.path.exists(p):
            return p
    raise FileNotFoundError(stem)


def compile_one(obj: str) -> tuple:
    src = source_of(obj)
    out = os.path.join(OBJ, obj)
    if os.path.exists(out) and os.path.getmtime(out) >= os.path.getmtime(src):
        return obj, 0, ""
    if src.endswith(".c"):
        cmd = ["gcc", "-O2", "-c", src, "-o", out]
    else:
        cmd = [FC, "-O2", "-cpp", "-DNOCOMM", "-w", "-I" + REF, "-I" + OBJ, "-J" + OBJ, "-c", src, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=OBJ)
    return obj, r.returncode, r.stderr[-2000:]


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=8)
    ap.add_argument("--dropin", action="store_true")
    a = ap.parse_args()
    if not os.path.isdir(REF):
        print(f"reference sources not present: {REF}")
        return 1
    os.makedirs(OBJ, exist_ok=True)
    objs = object_list()
    first = [o for m in MODULE_FILES for o in objs if o == m + ".o"]
    rest = [o for o in objs if o not in first]
    for o in first:
        _, rc, err = compile_one(o)
        if rc:
            print(f"{o}: {err}")
            return 1
    failed = []
    with cf.ThreadPoolExecutor(a.j) as ex:
        for o, rc, err in ex.map(compile_one, rest):
            if rc:
                failed.append(o)
                print(f"{o}: {err}")
    if failed:
        return 1
    cmd = [FC, "-O2", "-o", EXE + ".tmp"] + [os.path.join(OBJ, o) for o in objs] + \
          ["-L" + MKL, "-Wl,-rpath," + MKL, "-lmkl_intel_lp64", "-lmkl_sequential", "-lmkl_core", "-lpthread", "-lm",
           "-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        print(r.stderr[-4000:])
        return 1
    os.replace(EXE + ".tmp", EXE)
    print(f"{EXE}: {len(objs)} objects")
    if a.dropin:
        return build_dropin(objs)
    return 0


def build_dropin(objs: list) -> int:
    lib = os.path.join(PKG, "libflosic_hip.so")
    if not os.path.exists(lib):
        print(f"{lib} is missing: build the library first")
        return 1
    steps = [[FC, "-O2", "-J" + OBJ, "-I" + OBJ, "-c", os.path.join(REPO, "include", "flosic_hip.f90"), "-o",
              os.path.join(OBJ, "flosic_hip_iface.o")],
             [FC, "-O2", "-cpp", "-J" + OBJ, "-I" + OBJ, "-I" + REF, "-c",
              os.path.join(PKG, "fortran", "flosic_dropin.F90"), "-o", os.path.join(OBJ, "flosic_dropin.o")]]
    for cmd in steps:
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode:
            print(r.stderr[-4000:])
            return 1
    keep = [os.path.join(OBJ, o) for o in objs if o not in REPLACED]
    rpath = os.path.relpath(PKG, OUT)
    cmd = [FC, "-O2", "-o", DROPIN_EXE + ".tmp"] + keep + \
          [os.path.join(OBJ, "flosic_hip_iface.o"), os.path.join(OBJ, "flosic_dropin.o"), "-L" + PKG, "-lflosic_hip",
           "-Wl,-rpath,$ORIGIN/" + rpath, "-L" + MKL, "-Wl,-rpath," + MKL, "-lmkl_intel_lp64", "-lmkl_sequential",
           "-lmkl_core", "-lpthread", "-lm", "-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        print(r.stderr[-4000:])
        return 1
    os.replace(DROPIN_EXE + ".tmp", DROPIN_EXE)
    print(f"{DROPIN_EXE}: {len(keep)} reference objects + the library's drop-ins")
    return 0


if __name__ == "__main__":
    sys.exit(main())
