"""Golden vectors of the reference's DIAGV_FO (diagv_fo.F:5-156): the routine itself, compiled from the reference
sources by oracle/build_nrlmol.py (its objects in oracle/_ref/nrlmol/obj), is linked with diagv_fo_driver.f90 and
run on seeded orbital Hamiltonians shaped like SCFSICW's (scfsicw.f:353-395): a symmetric part with occupied
eigenvalues in [-20, -0.3] Ha and virtual ones in [0.05, 8] Ha, plus SIC(I,J) of order 0.05 Ha added to the
occupied rows.  Writes diagv_fo.npz: for each case m, n, ham (ham[J, I] = HAM(J,I)), phires (phires[inew, iold] =
PHIRES(INEW,IOLD) after the routine's final transpose) and the sweep count the reference printed.

    python tests/golden/diagv_fo/make_diagv_fo.py      (needs the reference objects: oracle/build_nrlmol.py)
"""
import os
import re
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(os.path.dirname(HERE)))
REF = os.environ.get("FLOSIC_REF", "/root/reference/flosic")
OBJ = os.path.join(REPO, "oracle", "_ref", "nrlmol", "obj")
FC = "/opt/rocm/bin/amdflang"
MKL = "/opt/conda/lib"
CASES = [(12, 4, 1), (30, 9, 2), (73, 5, 3), (40, 1, 4)]


def hamiltonian(m, n, seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.normal(size=(m, m)))
    eps = np.sort(np.concatenate([rng.uniform(-20.0, -0.3, n), rng.uniform(0.05, 8.0, m - n)]))
    H = Q @ np.diag(eps) @ Q.T
    S = 0.05 * rng.normal(size=(m, m))
    H[:n, :] += S[:, :n].T  # HAM(J,I) += SIC(I,J) for the occupied J
    return H


def main() -> int:
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import build_nrlmol as B
    objs = [os.path.join(OBJ, o) for o in B.object_list() if o != "main.o"]
    if not all(os.path.exists(o) for o in objs):
        print("reference objects missing: run oracle/build_nrlmol.py first")
        return 1
    work = tempfile.mkdtemp(prefix="diagv_fo_")
    try:
        drv = os.path.join(work, "drv")
        subprocess.run([FC, "-O2", "-cpp", "-DNOCOMM", "-I" + REF, "-I" + OBJ, "-c",
                        os.path.join(HERE, "diagv_fo_driver.f90"), "-o", drv + ".o"], check=True, cwd=work)
        subprocess.run([FC, "-O2", "-o", drv, drv + ".o"] + objs +
                       ["-L" + MKL, "-Wl,-rpath," + MKL, "-lmkl_intel_lp64", "-lmkl_sequential", "-lmkl_core",
                        "-lpthread", "-lm", "-ldl"], check=True, cwd=work)
        # DIAGGE reads the run options (CHECK_INPUTS): the config-1 input file of the committed fixtures
        shutil.copyfile(os.path.join(REPO, "tests", "golden", "reference_runs", "h2o_lsda", "inputs",
                                     "NRLMOL_INPUT.DAT"), os.path.join(work, "NRLMOL_INPUT.DAT"))
        out = {}
        for k, (m, n, seed) in enumerate(CASES):
            H = hamiltonian(m, n, seed)
            inp = f"{m} {n}\n" + "\n".join(" ".join(f"{H[i, j]:.17e}" for i in range(m)) for j in range(m)) + "\n"
            r = subprocess.run([drv], input=inp, capture_output=True, text=True, cwd=work, check=True)
            sweeps = re.findall(r"TRACER:ITER=\s*(\d+)", r.stdout)
            vals = np.array([float(v) for v in open(os.path.join(work, "phires.txt")).read().split()])
            phires = vals.reshape(m, m).T  # file: column by column
            out[f"c{k}_ham"] = H
            out[f"c{k}_n"] = np.int32(n)
            out[f"c{k}_phires"] = phires
            out[f"c{k}_sweeps"] = np.int32(int(sweeps[-1]) if sweeps else 0)
            print(f"case {k}: m={m} n={n} sweeps={out[f'c{k}_sweeps']}")
        np.savez_compressed(os.path.join(HERE, "diagv_fo.npz"), ncases=np.int32(len(CASES)), **out)
    finally:
        shutil.rmtree(work, ignore_errors=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
