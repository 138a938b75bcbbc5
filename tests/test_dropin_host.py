"""The Fortran drop-ins' gathering of the reference's module data (publicrelease_2020_amd/fortran/flosic_dropin.F90),
checked on the CPU against the run directory read by formats.load_run.

The reference program linked with the drop-ins (oracle/_ref/nrlmol/nrlmol_flosic_hip, built where the reference
sources are) runs the config-1 H2O inputs: the reference's own SCF, then its SIC block, where FRMORB2(0,0) hands the
modules' basis (COMMON2), mesh (MESH1), orbitals (COMMON5 PSI_COEF, symmetry-adapted order) and FODs (FRM) to the
library.  With FLOSIC_DROPIN_DUMP set, those arrays are written out before any library call (without a GPU the
program then stops at flosic_create).  They must equal what formats.load_run builds from the same run directory: the
basis flattening bit for bit (COMMON2 holds what READINP computed from the same ISYMGEN), the local-order
coefficients to 1e-14 relative against the golden WFOUT the reference wrote after the same SIC block, the FODs exactly.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NRLMOL = os.path.join(ROOT, "oracle", "_ref", "nrlmol")
FIX = os.path.join(ROOT, "tests", "golden", "reference_runs", "h2o_lsda")


def _read_dump(path):
    b = open(path, "rb").read()
    k = 0

    def take(dtype, n):
        nonlocal k
        a = np.frombuffer(b, dtype=dtype, count=n, offset=k)
        k += a.nbytes
        return a.copy()

    natom, nbas, nalpha, ncoef, nmsh, nspn = take("<i4", 6)
    out = dict(natom=natom, nbas=nbas, nmsh=nmsh, nspn=nspn, xyz=take("<f8", 3 * natom),
               bare_off=take("<i4", natom + 1), alpha=take("<f8", nalpha), ncon=take("<i4", 3 * natom),
               con_off=take("<i4", 3 * natom), con=take("<f8", max(ncoef, 1)), coef=[], fod=[])
    for _ in range(nspn):
        nwf, nf = take("<i4", 2)
        out["coef"].append(take("<f8", nbas * nwf).reshape(nwf, nbas).T)
        out["fod"].append(take("<f8", 3 * nf).reshape(nf, 3))
    return out


@pytest.mark.skipif(not os.path.exists(os.path.join(NRLMOL, "nrlmol_flosic_hip")),
                    reason="the reference program and its drop-in build are built only where the reference sources are")
def test_dropin_hands_over_module_data(tmp_path):
    from reference_runs import load
    scf = tmp_path / "scf"
    shutil.copytree(os.path.join(FIX, "inputs"), scf, copy_function=shutil.copyfile)
    with open(scf / "out.txt", "w") as f:
        subprocess.run([os.path.join(NRLMOL, "nrlmol_exe")], cwd=scf, stdout=f, stderr=subprocess.STDOUT,
                       timeout=600, check=False)
    d = tmp_path / "sic"
    shutil.copytree(scf, d, copy_function=shutil.copyfile)
    shutil.copyfile(os.path.join(FIX, "FRMORB"), d / "FRMORB")
    txt = open(d / "NRLMOL_INPUT.DAT").read()
    txt = re.sub(r"^(NONSCFV\s*=\s*)'N'", r"\1'Y'", txt, flags=re.M)
    txt = re.sub(r"^(NONSCFFORCESV\s*=\s*)'N'", r"\1'Y'", txt, flags=re.M)
    open(d / "NRLMOL_INPUT.DAT", "w").write(txt)
    env = dict(os.environ, FLOSIC_DROPIN_DUMP=str(d / "dump.bin"))
    with open(d / "out.txt", "w") as f:
        subprocess.run([os.path.join(NRLMOL, "nrlmol_flosic_hip")], cwd=d, stdout=f, stderr=subprocess.STDOUT,
                       timeout=600, env=env, check=False)
    got = _read_dump(d / "dump.bin")
    sysm, ref = load("h2o_lsda")
    b = sysm.basis
    assert got["natom"] == b.natom and got["nbas"] == b.nbas and got["nmsh"] == sysm.mesh.npts and got["nspn"] == 2
    np.testing.assert_array_equal(got["xyz"].reshape(-1, 3), b.xyz)
    np.testing.assert_array_equal(got["bare_off"], b.bare_off)
    np.testing.assert_array_equal(got["alpha"], b.alpha)
    np.testing.assert_array_equal(got["ncon"], np.asarray(b.ncon).ravel())
    np.testing.assert_array_equal(got["con_off"], np.asarray(b.con_off).ravel())
    np.testing.assert_allclose(got["con"], b.con, rtol=1e-15, atol=0)
    for s in range(2):
        np.testing.assert_array_equal(got["fod"][s], sysm.fod[s])
        c = got["coef"][s]
        assert c.shape == sysm.coef[s].shape
        np.testing.assert_allclose(c, sysm.coef[s], rtol=0, atol=1e-14 * np.abs(sysm.coef[s]).max())
