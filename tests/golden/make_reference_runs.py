"""Golden fixtures from the reference program itself: run nrlmol_exe (built from the reference sources by
oracle/build_nrlmol.py, serial) on small systems and keep its input and output files for the parity tests.

Each case is two runs of the reference in a scratch directory under oracle/_ref/runs/<case>/:
  scf/  the reference's own SCF (SCF-ONLY) from its inputs: mesh (VMOLD), basis (ISYMGEN), converged orbitals (WFOUT)
  sic/  a copy of scf/ with the FODs in FRMORB and NONSCFV = NONSCFFORCESV = 'Y' (main.F:570,616-760): one FLO-SIC
        evaluation with FOD forces on the converged orbitals (E_SIC, RESULTS, SIC matrix, fforce.dat)

Cases:
  hydrogen    the reference's own hydrogen example (utep-jakar/hydrogen: FLO-SIC LDA SCF with its FOD); the SCF's
              total energy is the published FLO-SIC LDA anchor -0.499922 Ha (flosic/repository_test.txt:11)
  neon        the reference's own neon example (utep-jakar/neon, 5+5 FODs); anchor -129.280706 Ha
              (repository_test.txt:20)
  h2o_lsda    H2O with O at the origin, O-H 1.809 bohr, H-O-H 104.5 deg, NRLMOL default basis and mesh, spin
              polarised LDA-PW91 DFT SCF; then FLO-SIC with 5+5 FODs (core at O, two bonds at 0.6 of O->H, two lone
              pairs 0.9 bohr from O perpendicular to the plane, seeded jitter sigma 0.02 bohr: system.make_config(1))
  h2o_pbe     the same with GGA-PBE*GGA-PBE
  h2o_lsda_scf  the reference's FLO-SIC SCF (SIC block every iteration, SCFSICW / DIAGGE_FO consuming the SIC
              matrix) from the converged LDA state of h2o_lsda with its FODs: the SUMMARY trace of every iteration,
              the last ZSIC and the final fforce.dat
  h2o_lsda_fodopt  the reference's FOD optimisation as its users run it: repeated non-SCF evaluations with forces
              in one directory, each ending in its FOD optimiser (ELECTRONIC_GEOMETRY, electronic_geometry_lbfgs.f),
              which writes the next FRMORB; per step the evaluated FODs, ZSIC, fforce.dat and the fande.out line

Kept per case in tests/golden/reference_runs/<case>/: SYMBOL, ISYMGEN, FRMORB, WFOUT.gz, VMOLD.gz (the reference's
files as written), fforce.dat, for the H2O cases the SCF's inputs (inputs/: CLUSTER, GRPMAT, NRLMOL_INPUT.DAT), and
reference.json with the numbers read from the reference's output: ZSIC (full
precision, 'MAIN : ZSIC ='), the ORBITAL SIC lines of each spin (RESULTS(1:9) in F10.6), the printed SIC(j,i)
matrix per spin (F12.5), the SCF's total energy and the counts.

    python tests/golden/make_reference_runs.py [case ...]
"""
from __future__ import annotations

import gzip
import json
import os
import re
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
REF = "/root/reference"
EXE = os.path.join(ROOT, "oracle", "_ref", "nrlmol", "nrlmol_exe")
RUNS = os.path.join(ROOT, "oracle", "_ref", "runs")
OUT = os.path.join(ROOT, "tests", "golden", "reference_runs")

ANCHORS = {"hydrogen": -0.499922, "neon": -129.280706}  # flosic/repository_test.txt:11,20


def run(directory: str) -> str:
    with open(os.path.join(directory, "out.txt"), "w") as f:
        r = subprocess.run([EXE], cwd=directory, stdout=f, stderr=subprocess.STDOUT, timeout=3600)
    text = open(os.path.join(directory, "out.txt")).read()
    # the FOD optimiser after the evaluation may stop on its own files; the evaluation's outputs come before it
    if r.returncode != 0 and "HOPE TO SEE YOU AGAIN" not in text:
        print(f"{directory}: reference exited with {r.returncode} after its evaluation")
    return text


def set_option(path: str, key: str, value: str):
    txt = open(path).read()
    txt, n = re.subn(rf"^({key}\s*=\s*)('[^']*'|\S+)", rf"\g<1>{value}", txt, flags=re.M)
    if n != 1:
        raise RuntimeError(f"{key} not found in {path}")
    open(path, "w").write(txt)


def example_inputs(case: str, d: str):
    src = os.path.join(REF, "utep-jakar", case)
    for f in os.listdir(src):
        shutil.copy(os.path.join(src, f), d)


def h2o_inputs(functional: str, d: str):
    from publicrelease_2020_amd import system as S
    h = S.make_config(1, nrad=4, nang=6)  # the geometry and FODs of the workload (its mesh is not used)
    with open(os.path.join(d, "CLUSTER"), "w") as f:
        f.write(f"{functional}\nGRP\n3\n")
        for (x, y, z), Z in zip(h.basis.xyz, h.basis.Z):
            f.write(f"{x:.10f} {y:.10f} {z:.10f} {int(Z)} ALL\n")
        f.write("0.0 0.0\n")
    shutil.copy(os.path.join(REF, "utep-jakar", "hydrogen", "GRPMAT"), d)  # the identity: point group NONE
    shutil.copy(os.path.join(REF, "utep-jakar", "hydrogen", "NRLMOL_INPUT.DAT"), d)
    set_option(os.path.join(d, "NRLMOL_INPUT.DAT"), "SPNPOLV", "'Y'")
    set_option(os.path.join(d, "NRLMOL_INPUT.DAT"), "FIXMV", "'N'")
    return h


def parse(text: str, nup: int, ndn: int) -> dict:
    z = re.findall(r"MAIN : ZSIC =\s*(\S+)", text)
    rows = [[float(v) for v in ln.split()[2:]] for ln in re.findall(r"ORBITAL SIC:.*", text)]
    sic = {}
    for s, n in ((1, nup), (2, ndn)):
        m = [x.start() for x in re.finditer(rf"SIC MATRIX FOR SPIN LSPN:\s*{s}\s*\n", text)]
        if not m or n == 0:
            continue
        lines = text[m[-1]:].splitlines()[1:]
        vals = []
        for ln in lines:
            toks = ln.split()
            try:
                vals += [float(t) for t in toks]
            except ValueError:
                break
            if len(vals) >= n * n:
                break
        sic[str(s)] = [vals[i * n:(i + 1) * n] for i in range(n)]
    # the energy summary of the evaluation's density (main.F's SUMMARY OF ENERGY CONTRIBUTIONS, F-format 6 decimals):
    # exchange-correlation (local + nonlocal parts) and the mean-field Coulomb (Hartree) energy of the spin densities
    def last(key):
        v = re.findall(rf"^{key}:\s*(\S+)", text, flags=re.M)
        return float(v[-1]) if v else None
    energies = {k: last(label) for k, label in (("exc", "EXCHANGE-CORRELATION"), ("ex_local", "LOCAL EXCHANGE"),
                                                 ("ec_local", "LOCAL CORRELATION"), ("ex_nonlocal", "NONLOCAL EXCHANGE"),
                                                 ("ec_nonlocal", "NONLOCAL CORRELATION"),
                                                 ("hartree", "MEAN-FIELD COULOMB"))}
    return dict(zsic=float(z[-1].replace("D", "E")), orbital_sic=[rows[:nup], rows[nup:nup + ndn]], sic_matrix=sic,
                energies=energies)


def scf_energy(directory: str) -> float:
    lines = [ln.split() for ln in open(os.path.join(directory, "SUMMARY")).read().splitlines()[1:] if ln.strip()]
    return float(lines[-1][5])  # EDFT+SIC of the last iteration (the DFT energy when the SCF has no FODs)


def make_case(case: str):
    from publicrelease_2020_amd import formats as FM
    base = os.path.join(RUNS, case)
    shutil.rmtree(base, ignore_errors=True)
    scf, sic = os.path.join(base, "scf"), os.path.join(base, "sic")
    os.makedirs(scf)
    fods = None
    if case in ("hydrogen", "neon"):
        example_inputs(case, scf)
    elif case == "h2o_lsda":
        fods = h2o_inputs("LDA-PW91*LDA-PW91", scf).fod
    elif case == "h2o_pbe":
        fods = h2o_inputs("GGA-PBE*GGA-PBE", scf).fod
    else:
        raise ValueError(case)
    for f in ("CLUSTER", "GRPMAT", "NRLMOL_INPUT.DAT"):
        if os.path.exists(os.path.join(scf, f)):
            shutil.copy(os.path.join(scf, f), os.path.join(base, "inputs_" + f))
    run(scf)
    etot = scf_energy(scf)
    shutil.copytree(scf, sic)
    if fods is not None:
        open(os.path.join(sic, "FRMORB"), "w").write(FM.write_frmorb(fods[0], fods[1]))
    set_option(os.path.join(sic, "NRLMOL_INPUT.DAT"), "NONSCFV", "'Y'")
    set_option(os.path.join(sic, "NRLMOL_INPUT.DAT"), "NONSCFFORCESV", "'Y'")
    frm = open(os.path.join(sic, "FRMORB")).read()  # the FODs the evaluation read (written by the SCF run if absent)
    up, dn = FM.read_frmorb(frm)
    text = run(sic)
    res = parse(text, up.shape[0], dn.shape[0])
    res.update(case=case, scf_total_energy=etot, published_total_energy=ANCHORS.get(case),
               nmsh=int(re.findall(r"TIME FOR COULOMB POTENTIAL,NMSH:\s*\S+\s+(\d+)", text)[-1]),
               nfod=[int(up.shape[0]), int(dn.shape[0])])
    dst = os.path.join(OUT, case)
    shutil.rmtree(dst, ignore_errors=True)
    os.makedirs(dst)
    for f in ("SYMBOL", "ISYMGEN", "fforce.dat"):
        shutil.copyfile(os.path.join(sic, f), os.path.join(dst, f))
    if fods is not None:  # the SCF's own inputs, so the whole run can be repeated where the reference is absent
        os.makedirs(os.path.join(dst, "inputs"))
        for f in ("CLUSTER", "GRPMAT", "NRLMOL_INPUT.DAT"):
            shutil.copyfile(os.path.join(base, "inputs_" + f), os.path.join(dst, "inputs", f))
    open(os.path.join(dst, "FRMORB"), "w").write(frm)
    for f in ("WFOUT", "VMOLD"):
        with open(os.path.join(sic, f), "rb") as fi, open(os.path.join(dst, f + ".gz"), "wb") as raw, \
                gzip.GzipFile(fileobj=raw, mode="wb", compresslevel=9, mtime=0) as fo:
            fo.write(fi.read())
    json.dump(res, open(os.path.join(dst, "reference.json"), "w"), indent=1)
    print(f"{case}: ZSIC {res['zsic']!r}, SCF total energy {etot!r}, {res['nmsh']} mesh points, FODs {res['nfod']}")


def make_scf_case(case: str):
    """FLO-SIC SCF of the config-1 H2O (the reference's SCF loop with its SIC block every iteration, started from the
    converged LDA state of h2o_lsda): the per-iteration SUMMARY trace and the final FOD forces."""
    from publicrelease_2020_amd import formats as FM
    src = os.path.join(RUNS, "h2o_lsda", "scf")
    if not os.path.isdir(src):
        make_case("h2o_lsda")
    d = os.path.join(RUNS, case)
    shutil.rmtree(d, ignore_errors=True)
    shutil.copytree(src, d)
    shutil.copy(os.path.join(OUT, "h2o_lsda", "FRMORB"), d)
    text = run(d)
    lines = open(os.path.join(d, "SUMMARY")).read().splitlines()
    last = max(i for i, ln in enumerate(lines) if ln.split()[:1] == ["IT"])  # the SCF appends a block per run
    rows = [[float(v) for v in ln.split()] for ln in lines[last + 1:] if ln.strip()]
    res = dict(case=case, summary_columns=["IT", "TRACE", "ETOT", "EKIN+ENONLOC", "CHARGE", "EDFT+SIC", "LOWEST"],
               summary=rows, zsic=float(re.findall(r"MAIN : ZSIC =\s*(\S+)", text)[-1].replace("D", "E")))
    dst = os.path.join(OUT, case)
    shutil.rmtree(dst, ignore_errors=True)
    os.makedirs(dst)
    shutil.copyfile(os.path.join(d, "fforce.dat"), os.path.join(dst, "fforce.dat"))
    json.dump(res, open(os.path.join(dst, "reference.json"), "w"), indent=1)
    print(f"{case}: {len(rows)} SCF iterations, final EDFT+SIC {rows[-1][5]!r}, last ZSIC {res['zsic']!r}")


FODOPT_STEPS = 6


def make_fodopt_case(case: str):
    """FOD optimisation of the config-1 H2O by the reference's own optimiser, FODOPT_STEPS steps: each step is one run
    of the program (NONSCF with forces on the converged LDA state of h2o_lsda) in the same directory; the run's
    ELECTRONIC_GEOMETRY (FOD_OPT, its state in the directory's optimiser files) writes the FRMORB of the next step."""
    from publicrelease_2020_amd import formats as FM
    src = os.path.join(RUNS, "h2o_lsda", "scf")
    if not os.path.isdir(src):
        make_case("h2o_lsda")
    d = os.path.join(RUNS, case)
    shutil.rmtree(d, ignore_errors=True)
    shutil.copytree(src, d)
    shutil.copyfile(os.path.join(OUT, "h2o_lsda", "FRMORB"), os.path.join(d, "FRMORB"))
    for key in ("NONSCFV", "NONSCFFORCESV"):
        set_option(os.path.join(d, "NRLMOL_INPUT.DAT"), key, "'Y'")
    steps = []
    for k in range(FODOPT_STEPS):
        up, dn = FM.read_frmorb(open(os.path.join(d, "FRMORB")).read())
        text = run(d)
        steps.append(dict(fod_up=up.tolist(), fod_dn=dn.tolist(),
                          zsic=float(re.findall(r"MAIN : ZSIC =\s*(\S+)", text)[-1].replace("D", "E")),
                          fforce=FM.read_fforce(open(os.path.join(d, "fforce.dat")).read()).tolist()))
    fande = [[float(x) for x in ln.split()] for ln in open(os.path.join(d, "fande.out")).read().splitlines()
             if ln.strip()]
    up, dn = FM.read_frmorb(open(os.path.join(d, "FRMORB")).read())
    res = dict(case=case, steps=steps, fande=fande, next_fod_up=up.tolist(), next_fod_dn=dn.tolist())
    dst = os.path.join(OUT, case)
    shutil.rmtree(dst, ignore_errors=True)
    os.makedirs(dst)
    json.dump(res, open(os.path.join(dst, "reference.json"), "w"), indent=1)
    print(f"{case}: {len(steps)} optimiser steps, E {fande[0][1]!r} -> {fande[-1][1]!r}")


def main(argv) -> int:
    if not os.path.exists(EXE):
        subprocess.check_call([sys.executable, os.path.join(ROOT, "oracle", "build_nrlmol.py")])
    for case in argv or ["hydrogen", "neon", "h2o_lsda", "h2o_pbe", "h2o_lsda_scf",
                                 "h2o_lsda_fodopt"]:
        if case == "h2o_lsda_scf":
            make_scf_case(case)
        elif case == "h2o_lsda_fodopt":
            make_fodopt_case(case)
        else:
            make_case(case)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
