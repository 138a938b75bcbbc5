"""The SCF consumer of the SIC block, through the C ABI: DIAGV_FO (flosic_diagv_fo) against the reference's own
routine, and SCFSICW's FLO-SIC orbital update (flosic_scf_sic_orbitals) against the oracle's restatement fed with the
library's own SIC matrix and FLOs.

Tolerances: PHIRES against the reference's golden vectors (tests/golden/diagv_fo) up to the sign of each new
orbital to 1e-10 with the same converged sweep; the orbital update against the oracle to 1e-9 (new orbitals, up to
sign) and 1e-9 Ha (EVAL); the new orbitals S-orthonormal to 1e-10 and the occupied-virtual block of C_new^T (H + SIC)
C_new below DIAGV_FO's 1e-6 threshold."""
import os

import numpy as np
import pytest

from publicrelease_2020_amd import system as S

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "diagv_fo", "diagv_fo.npz")


def _align_rows(phi, ref):
    s = np.sign(np.sum(phi * ref, axis=1))
    s[s == 0] = 1.0
    return phi * s[:, None]


@pytest.mark.parametrize("case", range(4))
def test_diagv_fo_against_reference(case):
    from publicrelease_2020_amd import flosic as F
    g = np.load(GOLD)
    H, n, ref = g[f"c{case}_ham"], int(g[f"c{case}_n"]), g[f"c{case}_phires"]
    phi, sweeps = F.diagv_fo(H, n)
    print(f"case {case}: m={H.shape[0]} n={n} sweeps {sweeps} (reference {int(g[f'c{case}_sweeps'])})", flush=True)
    assert sweeps == int(g[f"c{case}_sweeps"])
    assert np.abs(_align_rows(phi, ref) - ref).max() < 1e-10


@pytest.mark.parametrize("xc", [S.XC_LSDA, S.XC_PBE], ids=["LSDA", "PBE"])
def test_scf_sic_orbitals_against_oracle(xc):
    from oracle import oracle as O
    from publicrelease_2020_amd import flosic as F
    sysm = S.make_synthetic(3, 3, 7, xc, nrad=24, nang=50, box=4.5)  # all nbas orbitals, as SCFSICW uses them
    ctx = F.Flosic(0)
    ctx.load_system(sysm)
    ctx.evaluate()
    N = sysm.nbas
    Sao = S.overlap_matrix(sysm.basis)
    try:
        for s in (1, 2):
            C = sysm.coef[s - 1]
            n = sysm.nocc[s - 1]
            assert C.shape == (N, N)
            # a Kohn-Sham Hamiltonian whose eigenvectors are the orbitals: H C = S C eps
            eps = np.sort(np.concatenate([np.linspace(-12.0, -0.4, n), np.linspace(0.1, 6.0, N - n)]))
            ham_ao = Sao @ C @ np.diag(eps) @ C.T @ Sao
            coef, ev, sweeps = ctx.scf_sic_orbitals(s, ham_ao)
            tmat = ctx.fermi_orbitals(s)["tmat"]
            cfv = np.concatenate([C[:, :n] @ tmat.T, C[:, n:]], axis=1)
            sic = ctx.siclagm(s)
            c_ref, ev_ref, sw_ref = O.scf_sic_orbitals(ham_ao, cfv, sic, n)
            print(f"spin {s}: {sweeps} sweeps (oracle {sw_ref}), occupied EVAL {ev[:n]}", flush=True)
            assert sweeps == sw_ref and sweeps > 0
            assert np.abs(_align_rows(coef.T, c_ref.T) - c_ref.T).max() < 1e-9
            assert np.abs(ev - ev_ref).max() < 1e-9
            assert np.abs(coef.T @ Sao @ coef - np.eye(N)).max() < 1e-10
            ham = cfv.T @ ham_ao @ cfv
            ham[:n, :] += sic[:, :n].T
            rot = np.linalg.solve(cfv, coef)  # PHIRES^T: new orbitals in the old ones
            blk = (rot.T @ ham @ rot)[:n, n:]
            assert np.abs(blk).max() < 1e-6
    finally:
        ctx.close()
