"""flosic_siclag_der (SICLAG_DER, the SIC part of the atomic Pulay force) against the oracle's oracle_siclag_der
(pinned by central differences in tests/test_oracle_siclag_der.py) fed with the oracle's own FLO potentials V_i,
for LSDA and PBE, with resident planes and in mesh-block mode (several blocks)."""
import numpy as np
import pytest

from publicrelease_2020_amd import system as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("xc,block", [(S.XC_LSDA, 0), (S.XC_LSDA, 512), (S.XC_PBE, 0), (S.XC_PBE, 1024)])
def test_siclag_der_matches_oracle(xc, block):
    from oracle import oracle as O
    from publicrelease_2020_amd import flosic as F
    sysm = S.make_synthetic(3, 3, 7, xc, nrad=24, nang=50, box=5.0)
    ctx = F.Flosic(0)
    ctx.load_system(sysm)
    if block:
        ctx.set_mesh_block(block)
    ctx.evaluate()
    d = ctx.siclag_der()
    ref = O.evaluate(sysm, want_potentials=True, want_vab=False)
    assert ref["status"] == 0
    vpot = np.vstack([ref["spins"][0]["vpot"], ref["spins"][1]["vpot"]])
    d_ref = O.siclag_der(sysm, vpot)
    scale = np.abs(d_ref).max()
    err = np.abs(d - d_ref).max()
    print(f"xc {xc} block {block}: max |dDERSIC| {err:.3e} of {scale:.3e}", flush=True)
    assert d.shape == d_ref.shape
    assert err < 1e-10 * scale
    # a second call on the same evaluation returns the same values (the recomputation is deterministic)
    np.testing.assert_array_equal(ctx.siclag_der(), d)


def test_siclag_der_spin_restricted():
    """NSPN = 1: one spin channel of FLOs, the result doubled (siclag_der.F:158-166)"""
    import dataclasses

    from oracle import oracle as O
    from publicrelease_2020_amd import flosic as F
    sysm = dataclasses.replace(S.make_synthetic(3, 3, 7, S.XC_LSDA, nrad=24, nang=50, box=5.0), nspn=1)
    ctx = F.Flosic(0)
    ctx.load_system(sysm)
    ctx.evaluate()
    d = ctx.siclag_der()
    ref = O.evaluate(sysm, want_potentials=True, want_vab=False)
    assert ref["status"] == 0
    d_ref = O.siclag_der(sysm, ref["spins"][0]["vpot"], O.flo_coefs(sysm, 1))
    err = np.abs(d - d_ref).max()
    print(f"NSPN = 1: max |dDERSIC| {err:.3e} of {np.abs(d_ref).max():.3e}", flush=True)
    assert d.shape == d_ref.shape
    assert err < 1e-10 * np.abs(d_ref).max()


@pytest.mark.parametrize("xc", [S.XC_LSDA, S.XC_PBE], ids=["LSDA", "PBE"])
def test_siclag_der_several_point_blocks(xc):
    """A mesh of 46 560 points: resident LSDA rebuilds its gradient planes in more than one 32 768-point block and
    accumulates across them; the same call in mesh-block mode and the oracle give the same DERSIC."""
    from oracle import oracle as O
    from publicrelease_2020_amd import flosic as F
    sysm = S.make_synthetic(4, 4, 5, xc, nrad=60, nang=194, box=5.0)
    assert sysm.mesh.npts > 40000
    ctx = F.Flosic(0)
    ctx.load_system(sysm)
    ctx.evaluate()
    d = ctx.siclag_der()
    blk = F.Flosic(0)
    blk.load_system(sysm)
    blk.set_mesh_block(16384)
    blk.evaluate()
    d_blk = blk.siclag_der()
    ref = O.evaluate(sysm, want_potentials=True, want_vab=False)
    assert ref["status"] == 0
    vpot = np.vstack([ref["spins"][0]["vpot"], ref["spins"][1]["vpot"]])
    d_ref = O.siclag_der(sysm, vpot)
    scale = np.abs(d_ref).max()
    err, err_blk = np.abs(d - d_ref).max(), np.abs(d_blk - d).max()
    print(f"xc {xc}, M={sysm.mesh.npts}: max |dDERSIC| {err:.3e} (mesh-block vs resident {err_blk:.3e}) of "
          f"{scale:.3e}", flush=True)
    assert err < 1e-10 * scale
    assert err_blk < 1e-11 * scale
    ctx.close()
    blk.close()
