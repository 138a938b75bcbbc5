"""The oracle's SICLAG_DER (oracle_siclag_der, siclag_der.F:186-640 without point-group symmetry) pinned without the
GPU: DERSIC(x, i, A) = sum_k w dphi_i^A/dR_A,x V_i phi_i is the derivative of 1/2 sum_k w V_i phi_i^2 with respect
to the centre R_A of atom A's basis functions, with the mesh, the weights, V_i and the FLO coefficients held fixed.
Central differences of that sum, taken by moving one atom's basis functions, check the sign (GTGRAD supplies the
negative point gradient), the atom a basis function belongs to, and the coordinate index; the NSPN = 1 result is
twice the NSPN = 2 one (siclag_der.F:158-166)."""
import dataclasses

import numpy as np

from oracle import oracle as O
from publicrelease_2020_amd import system as S


def _system():
    return S.make_synthetic(3, 3, 7, S.XC_LSDA, nrad=16, nang=26, box=5.0)


def _potentials(sysm, n):
    # any fixed potentials will do: the identity holds for each V_i; distinct smooth fields per FLO
    r = sysm.mesh.xyz
    return np.stack([(1.0 + 0.3 * i) / (1.0 + np.sum((r - 0.2 * i) ** 2, axis=1)) for i in range(n)])


def _half_energy(sysm, basis, cflo, vpot):
    phi = O.basis_eval(basis, sysm.mesh.xyz, 0)[0] @ cflo  # (npts, n)
    return 0.5 * np.sum(sysm.mesh.w[:, None] * vpot.T * phi ** 2, axis=0)


def test_dersic_is_the_centre_derivative():
    sysm = _system()
    cflo = np.hstack([O.flo_coefs(sysm, 1), O.flo_coefs(sysm, 2)])
    vpot = _potentials(sysm, cflo.shape[1])
    d = O.siclag_der(sysm, vpot, cflo)
    assert d.shape == (sysm.basis.natom, cflo.shape[1], 3)
    scale = np.abs(d).max()
    assert scale > 1e-3

    def central(a, x, h):
        xyz_p = sysm.basis.xyz.copy()
        xyz_m = sysm.basis.xyz.copy()
        xyz_p[a, x] += h
        xyz_m[a, x] -= h
        ep = _half_energy(sysm, dataclasses.replace(sysm.basis, xyz=xyz_p), cflo, vpot)
        em = _half_energy(sysm, dataclasses.replace(sysm.basis, xyz=xyz_m), cflo, vpot)
        return (ep - em) / (2 * h)

    # central differences converge as h^2 here (the residual falls 4x per halving of h); the Richardson
    # combination of h and h/2 removes that term
    h = 2e-4
    for a in range(sysm.basis.natom):
        for x in range(3):
            fd = (4.0 * central(a, x, h / 2) - central(a, x, h)) / 3.0
            np.testing.assert_allclose(d[a, :, x], fd, rtol=0, atol=1e-8 * scale, err_msg=f"atom {a} x {x}")


def test_spin_restricted_doubles():
    sysm = _system()
    cflo = np.hstack([O.flo_coefs(sysm, 1), O.flo_coefs(sysm, 2)])
    vpot = _potentials(sysm, cflo.shape[1])
    d2 = O.siclag_der(sysm, vpot, cflo)
    d1 = O.siclag_der(dataclasses.replace(sysm, nspn=1), vpot, cflo)
    np.testing.assert_array_equal(d1, 2.0 * d2)
