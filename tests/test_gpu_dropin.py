"""The reference program with its per-FLO SIC block on the GPU: nrlmol_flosic_hip is the reference's own objects
(built from its sources by oracle/build_nrlmol.py --dropin) with APOTNL_SIC, FRMORB2, FRMIORB, SICLAGM and the SCF's
DIAGV_FO replaced by the library's Fortran drop-ins (publicrelease_2020_amd/fortran/flosic_dropin.F90), which read
the reference's module data and call the C ABI.

On the config-1 H2O inputs (tests/golden/reference_runs/h2o_lsda/inputs): the reference's own SCF runs first
(nrlmol_exe, CPU), then one FLO-SIC evaluation with forces (NONSCF) twice from the same SCF state, once by the
reference itself and once by the drop-in program.  Both print ZSIC (main.F:756) and write fforce.dat (main.F:708-715)
through the reference's own code; they agree to |dE| < 1e-10 Ha and |dF| < 1e-10 Ha/bohr (summation order only), and
the reference's SCF reproduces the golden fixture's SCF energy.

Then the reference's FLO-SIC SCF itself (SCF-ONLY with the FODs: the SIC block every iteration, SCFSICW adding the SIC
matrix to the Hamiltonian, DIAGGE_FO on the first iteration and the library's DIAGV_FO rotation on the later ones,
scfsicw.f:290-614) runs with the drop-ins from the same
converged LDA state: every iteration's EDFT+SIC in SUMMARY follows the reference's own run
(tests/golden/reference_runs/h2o_lsda_scf) to 1e-6 Ha, with the same iteration count, the final energy to 1e-7 Ha and
the final FOD forces to 1e-6 Ha/bohr.

Finally the reference's FOD optimisation (repeated non-SCF evaluations, each ending in its FOD optimiser) runs with the
drop-ins for six steps: the FODs of every step, ZSIC and forces follow the reference's own trajectory to 1e-8.
Skipped where the programs were not built (they need the reference sources at build time)."""
import os
import re
import shutil
import subprocess
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NRLMOL = os.path.join(ROOT, "oracle", "_ref", "nrlmol")
FIX = os.path.join(ROOT, "tests", "golden", "reference_runs", "h2o_lsda")


def _run(exe, d, timeout):
    t0 = time.perf_counter()
    with open(os.path.join(d, "out.txt"), "w") as f:
        subprocess.run([exe], cwd=d, stdout=f, stderr=subprocess.STDOUT, timeout=timeout)
    text = open(os.path.join(d, "out.txt")).read()
    return text, time.perf_counter() - t0


def _nonscf(d):
    p = os.path.join(d, "NRLMOL_INPUT.DAT")
    txt = open(p).read()
    txt = re.sub(r"^(NONSCFV\s*=\s*)'N'", r"\1'Y'", txt, flags=re.M)
    txt = re.sub(r"^(NONSCFFORCESV\s*=\s*)'N'", r"\1'Y'", txt, flags=re.M)
    open(p, "w").write(txt)


BUILT = os.path.exists(os.path.join(NRLMOL, "nrlmol_flosic_hip"))
SKIP = "the reference program and its drop-in build are built only where the reference sources are"


def _summary(path):
    lines = open(path).read().splitlines()
    last = max(i for i, ln in enumerate(lines) if ln.split()[:1] == ["IT"])  # the SCF appends a block per run
    return np.array([[float(v) for v in ln.split()] for ln in lines[last + 1:] if ln.strip()])


@pytest.fixture(scope="module")
def lda_state(tmp_path_factory):
    """The reference's own LDA SCF of the config-1 H2O inputs (CPU)."""
    scf = tmp_path_factory.mktemp("h2o") / "scf"
    shutil.copytree(os.path.join(FIX, "inputs"), scf, copy_function=shutil.copyfile)
    text, dt = _run(os.path.join(NRLMOL, "nrlmol_exe"), str(scf), 300)
    print(f"reference LDA SCF (CPU) {dt:.1f} s", flush=True)
    return scf


@pytest.mark.skipif(not BUILT, reason=SKIP)
def test_reference_program_with_gpu_dropins(lda_state, tmp_path):
    import json

    from publicrelease_2020_amd import formats as FM
    scf = lda_state
    ref = json.load(open(os.path.join(FIX, "reference.json")))
    assert abs(_summary(scf / "SUMMARY")[-1, 5] - ref["scf_total_energy"]) < 1e-8
    out = {}
    for name, exe in (("reference", "nrlmol_exe"), ("dropin", "nrlmol_flosic_hip")):
        d = tmp_path / name
        shutil.copytree(scf, d, copy_function=shutil.copyfile)
        shutil.copyfile(os.path.join(FIX, "FRMORB"), d / "FRMORB")
        _nonscf(str(d))
        print(f"{name}: NONSCF FLO-SIC evaluation with forces ...", flush=True)
        text, dt = _run(os.path.join(NRLMOL, exe), str(d), 300)
        z = float(re.findall(r"MAIN : ZSIC =\s*(\S+)", text)[-1])
        sic_t = float(re.findall(r"MAIN TIME FOR SIC :\s*(\S+)", text)[-1]) if "MAIN TIME FOR SIC :" in text else None
        out[name] = dict(zsic=z, fforce=FM.read_fforce(open(d / "fforce.dat").read()), wall=dt, sic_cpu_s=sic_t,
                         text=text)
        print(f"{name}: ZSIC {z!r}, wall {dt:.1f} s, reference SIC timer {sic_t}", flush=True)
    assert "FLOSIC_HIP: evaluation done" in out["dropin"]["text"]
    assert abs(out["dropin"]["zsic"] - out["reference"]["zsic"]) < 1e-10
    assert np.abs(out["dropin"]["fforce"] - out["reference"]["fforce"]).max() < 1e-10
    assert abs(out["dropin"]["zsic"] - ref["zsic"]) < 1e-9


@pytest.mark.skipif(not BUILT, reason=SKIP)
def test_reference_flosic_scf_with_gpu_dropins(lda_state, tmp_path):
    import json

    from publicrelease_2020_amd import formats as FM
    fix = os.path.join(ROOT, "tests", "golden", "reference_runs", "h2o_lsda_scf")
    ref = json.load(open(os.path.join(fix, "reference.json")))
    d = tmp_path / "flosic_scf"
    shutil.copytree(lda_state, d, copy_function=shutil.copyfile)
    shutil.copyfile(os.path.join(FIX, "FRMORB"), d / "FRMORB")
    print("drop-in FLO-SIC SCF ...", flush=True)
    text, dt = _run(os.path.join(NRLMOL, "nrlmol_flosic_hip"), str(d), 600)
    rows = _summary(d / "SUMMARY")
    want = np.array(ref["summary"])
    n = min(len(rows), len(want))
    print(f"drop-in FLO-SIC SCF: {len(rows)} iterations in {dt:.1f} s (reference {len(want)}), final EDFT+SIC "
          f"{rows[-1, 5]!r} (reference {want[-1, 5]!r}); max per-iteration |dE| "
          f"{np.abs(rows[:n, 5] - want[:n, 5]).max():.2e}", flush=True)
    assert text.count("FLOSIC_HIP: evaluation done") >= len(want)
    assert "FLOSIC_HIP: DIAGV_FO converged" in text  # the SCF's orbital rotation ran in the library
    assert len(rows) == len(want)
    assert np.abs(rows[:, 5] - want[:, 5]).max() < 1e-6
    assert abs(rows[-1, 5] - want[-1, 5]) < 1e-7
    ff = FM.read_fforce(open(d / "fforce.dat").read())
    ffr = FM.read_fforce(open(os.path.join(fix, "fforce.dat")).read())
    assert np.abs(ff - ffr).max() < 1e-6


@pytest.mark.skipif(not BUILT, reason=SKIP)
def test_reference_fod_optimisation_with_gpu_dropins(lda_state, tmp_path):
    """The reference's FOD optimisation as its users run it (repeated NONSCF evaluations with forces in one directory,
    each ending in ELECTRONIC_GEOMETRY, which writes the next FRMORB) with the drop-ins: the FODs evaluated at every
    step, ZSIC and the forces follow the reference's own trajectory (tests/golden/reference_runs/h2o_lsda_fodopt)."""
    import json

    from publicrelease_2020_amd import formats as FM
    fix = os.path.join(ROOT, "tests", "golden", "reference_runs", "h2o_lsda_fodopt")
    ref = json.load(open(os.path.join(fix, "reference.json")))
    d = tmp_path / "fodopt"
    shutil.copytree(lda_state, d, copy_function=shutil.copyfile)
    shutil.copyfile(os.path.join(FIX, "FRMORB"), d / "FRMORB")
    _nonscf(str(d))
    worst = dict(fod=0.0, zsic=0.0, force=0.0)
    for k, want in enumerate(ref["steps"]):
        up, dn = FM.read_frmorb(open(d / "FRMORB").read())
        text, dt = _run(os.path.join(NRLMOL, "nrlmol_flosic_hip"), str(d), 300)
        z = float(re.findall(r"MAIN : ZSIC =\s*(\S+)", text)[-1])
        ff = FM.read_fforce(open(d / "fforce.dat").read())
        worst["fod"] = max(worst["fod"], float(np.abs(up - np.array(want["fod_up"])).max()),
                           float(np.abs(dn - np.array(want["fod_dn"])).max()))
        worst["zsic"] = max(worst["zsic"], abs(z - want["zsic"]))
        worst["force"] = max(worst["force"], float(np.abs(ff - np.array(want["fforce"])).max()))
        print(f"FOD step {k}: ZSIC {z!r} (reference {want['zsic']!r}), {dt:.1f} s", flush=True)
    print(f"drop-in FOD optimisation, {len(ref['steps'])} steps: max |dFOD| {worst['fod']:.2e} bohr, max |dZSIC| "
          f"{worst['zsic']:.2e} Ha, max |dF| {worst['force']:.2e} Ha/bohr", flush=True)
    assert worst["fod"] < 1e-8 and worst["zsic"] < 1e-8 and worst["force"] < 1e-8
    fande = np.array([[float(x) for x in ln.split()] for ln in open(d / "fande.out").read().splitlines() if ln.strip()])
    np.testing.assert_allclose(fande, np.array(ref["fande"]), rtol=0, atol=1e-8)
