"""flosic_apotnl_dft (the DFT-side potential, APOTNL_DFT with NSPN = 2) against the oracle: spin densities
(LSDA: values; PBE: value, gradient and Hessian) from the oracle's FLO fields, the Hartree potential of the total
density from the oracle's per-FLO potentials, then oracle_dft_lsda / oracle_dft_gga (pinned in
tests/test_oracle_dft.py).  Also: the spin charges equal the sums of the per-FLO charges of the SIC path, mesh-block
mode gives the resident result, and the call is refused with one spin channel (NSPN = 1)."""
import numpy as np
import pytest

from publicrelease_2020_amd import system as S

pytestmark = pytest.mark.gpu


def test_apotnl_dft_matches_oracle():
    from oracle import oracle as O
    from publicrelease_2020_amd import flosic as F
    sysm = S.make_synthetic(3, 3, 7, S.XC_LSDA, nrad=24, nang=50, box=5.0)
    ctx = F.Flosic(0)
    ctx.load_system(sysm)
    ctx.evaluate()
    pot, en = ctx.apotnl_dft()
    ref = O.evaluate(sysm, want_potentials=True, want_vab=False)
    assert ref["status"] == 0
    vh = ref["spins"][0]["vh"].sum(0) + ref["spins"][1]["vh"].sum(0)
    rho = [(O.flo_fields(sysm, s) ** 2).sum(0) for s in (1, 2)]
    pot_ref, en_ref = O.dft_lsda(rho[0], rho[1], sysm.mesh.w, vh)
    err = np.abs(pot - pot_ref).max()
    print(f"max |dV| {err:.3e}, E_x {en['ex']:.12f} ({en_ref['ex']:.12f}), E_c {en['ec']:.12f}, E_H {en['eh']:.12f}",
          flush=True)
    assert err < 1e-10 * np.abs(pot_ref).max()
    for k in ("ex", "ec", "eh", "n_up", "n_dn"):
        assert abs(en[k] - en_ref[k]) < 1e-10 * abs(en_ref[k]) + 1e-13, k
    # the spin charges equal the sums of the FLO charges RESULTS(2) of the SIC path (another code path), and the
    # FLO counts up to the mesh quadrature
    for s, key in ((1, "n_up"), (2, "n_dn")):
        q = ctx.apotnl_sic(s)[:, 1].sum()
        assert abs(en[key] - q) < 1e-12 * q
        assert abs(en[key] - sysm.nocc[s - 1]) < 0.05 * sysm.nocc[s - 1]


def _flo_density_fields(sysm, s):
    """value, gradient and Hessian (n, 10) of the spin-s density sum phi^2 from the oracle's FLOs"""
    from oracle import oracle as O
    cf = O.flo_coefs(sysm, s)
    ph = np.einsum("gkb,bf->gkf", O.basis_eval(sysm.basis, sysm.mesh.xyz, 2), cf)  # (10, npts, n)
    p, g = ph[0], ph[1:4]
    rg = np.zeros((p.shape[0], 10))
    rg[:, 0] = (p * p).sum(1)
    rg[:, 1:4] = (2 * p[None] * g).sum(2).T
    for c, (a, b) in enumerate(((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))):
        rg[:, 4 + c] = (2 * (g[a] * g[b] + p * ph[4 + c])).sum(1)
    return rg


def test_apotnl_dft_pbe_matches_oracle():
    from oracle import oracle as O
    from publicrelease_2020_amd import flosic as F
    sysm = S.make_synthetic(3, 3, 7, S.XC_PBE, nrad=24, nang=50, box=5.0)
    ctx = F.Flosic(0)
    ctx.load_system(sysm)
    ctx.evaluate()
    pot, en = ctx.apotnl_dft()
    ref = O.evaluate(sysm, want_potentials=True, want_vab=False)
    assert ref["status"] == 0
    vh = ref["spins"][0]["vh"].sum(0) + ref["spins"][1]["vh"].sum(0)
    pot_ref, en_ref = O.dft_gga(_flo_density_fields(sysm, 1), _flo_density_fields(sysm, 2), sysm.mesh.w, vh)
    err = np.abs(pot - pot_ref).max()
    print(f"PBE: max |dV| {err:.3e}, E_x {en['ex']:.12f} ({en_ref['ex']:.12f}), E_c {en['ec']:.12f} "
          f"({en_ref['ec']:.12f})", flush=True)
    assert err < 1e-9 * np.abs(pot_ref).max()
    for k in ("ex", "ec", "eh", "n_up", "n_dn"):
        assert abs(en[k] - en_ref[k]) < 1e-10 * abs(en_ref[k]) + 1e-13, k


@pytest.mark.parametrize("xc", [S.XC_LSDA, S.XC_PBE])
def test_apotnl_dft_mesh_block_mode_matches_resident(xc):
    """mesh-block mode rebuilds the FLO fields per block; the 256-point partials are those of the resident pass"""
    from publicrelease_2020_amd import flosic as F
    sysm = S.make_synthetic(3, 3, 7, xc, nrad=24, nang=50, box=5.0)
    out = []
    for block in (0, 512):
        ctx = F.Flosic(0)
        ctx.load_system(sysm)
        if block:
            ctx.set_mesh_block(block)
        ctx.evaluate()
        out.append(ctx.apotnl_dft())
    (p0, e0), (p1, e1) = out
    err = np.abs(p1 - p0).max()
    print(f"xc {xc}: block vs resident max |dV| {err:.3e}", flush=True)
    assert err < 1e-12 * np.abs(p0).max()
    for k in e0:
        assert abs(e1[k] - e0[k]) < 1e-12 * abs(e0[k]) + 1e-14, k


def test_apotnl_dft_refuses_one_spin_channel():
    import dataclasses
    from publicrelease_2020_amd import flosic as F
    sysm = S.make_synthetic(3, 3, 7, S.XC_LSDA, nrad=16, nang=26, box=5.0)
    ctx = F.Flosic(0)
    ctx.load_system(dataclasses.replace(sysm, nspn=1))
    with pytest.raises(F.FlosicError):
        ctx.apotnl_dft()
