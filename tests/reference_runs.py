"""Shared loader for the reference-run fixtures (tests/golden/reference_runs/<case>/, written by the reference program
itself through tests/golden/make_reference_runs.py): the run directory is rebuilt in a temporary folder and read with
formats.load_run, as a user would read a reference run directory."""
import gzip
import json
import os
import shutil
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
RUNS = os.path.join(HERE, "golden", "reference_runs")
CASES = ["hydrogen", "neon", "h2o_lsda", "h2o_pbe"]
IRES = [1, 2, 4, 6, 9, 10, 11, 12, 13]  # the RESULTS slots main.F:699-702 prints on each ORBITAL SIC line


def load(case: str):
    from publicrelease_2020_amd import formats as FM
    src = os.path.join(RUNS, case)
    d = tempfile.mkdtemp(prefix=f"flosic_{case}_")
    try:
        for f in os.listdir(src):
            if os.path.isdir(os.path.join(src, f)):
                continue  # inputs/: the SCF's own inputs, not part of the evaluated run directory
            if f.endswith(".gz"):
                with gzip.open(os.path.join(src, f)) as fi, open(os.path.join(d, f[:-3]), "wb") as fo:
                    fo.write(fi.read())
            elif f != "reference.json":
                shutil.copy(os.path.join(src, f), d)
        sysm = FM.load_run(d)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    ref = json.load(open(os.path.join(src, "reference.json")))
    ref["fforce"] = FM.read_fforce(open(os.path.join(src, "fforce.dat")).read())
    return sysm, ref


def check(esic, results, sic, debdax, ref, tol_e=1e-9, tol_f=1e-9):
    """Compare one evaluation (E_SIC; per spin RESULTS (n, 13), SIC (nwf, n) as sic[j, i] = SIC(j,i), DEBDAX (n, 3))
    with the reference's printed outputs.  Returns the differences found."""
    out = dict(dE=abs(esic - ref["zsic"]))
    deb = np.vstack([d for d in debdax if d.size])
    out["dF"] = float(np.abs(deb - ref["fforce"][:deb.shape[0]]).max())
    dres, dsic = 0.0, 0.0
    for s in range(2):
        rows = np.array(ref["orbital_sic"][s])
        if rows.size:
            dres = max(dres, float(np.abs(results[s][:, np.array(IRES) - 1] - rows).max()))
        m = ref["sic_matrix"].get(str(s + 1))
        if m:
            n = len(m)
            dsic = max(dsic, float(np.abs(sic[s][:n, :n].T - np.array(m)).max()))  # row i: SIC(1:n, i)
    out["dRESULTS"], out["dSIC"] = dres, dsic
    assert out["dE"] < tol_e, out
    assert out["dF"] < tol_f, out
    assert dres < 6e-6, out   # F11.6 on the reference's side
    assert dsic < 6e-6, out   # F12.5
    return out


def check_dft_energies(e, ref, gga):
    """The DFT-side energies of the evaluation's spin densities ({ex, ec, eh} from APOTNL_DFT's kernels) against the
    reference's SUMMARY OF ENERGY CONTRIBUTIONS of the same density (F-format, 6 decimals): exchange and correlation
    (local plus nonlocal parts for PBE) and the mean-field Coulomb energy."""
    r = ref["energies"]
    want = dict(ex=r["ex_local"] + r["ex_nonlocal"], ec=r["ec_local"] + r["ec_nonlocal"], eh=r["hartree"],
                exc=r["exc"])
    got = dict(ex=e["ex"], ec=e["ec"], eh=e["eh"], exc=e["ex"] + e["ec"])
    d = {k: abs(got[k] - want[k]) for k in want}
    assert max(d.values()) < 1.1e-6, (d, got, want)
    return d
