"""The oracle's DFT-side LSDA potential (oracle_dft_lsda, the APOTNL_DFT path of getvlxc.F/subvlxc_libxc.F90 for
NSPN = 2 and the LDA functional) pinned without the GPU:
  * PW91LC at general zeta (EC, VCUP, VCDN) against the reference routine: tests/test_oracle_pins.py;
  * exchange equals the spin-scaled Slater closed form -(3/4)(6/pi)^(1/3) sum_s rho_s^(4/3);
  * V_xc,s is the derivative of E_xc with respect to rho_s (central differences);
  * at rho_dn = 0 it reproduces the SIC path's XC (xc_points, a separate restatement of the same routines);
  * the Hartree energy and the spin charges are the plain weighted sums.
A mistake in the spin scaling, a swapped VCUP/VCDN or a dropped term fails one of these."""
import numpy as np

from oracle import oracle as O


def _rho(n, seed):
    rng = np.random.default_rng(seed)
    return 10.0 ** rng.uniform(-5, 2, n), 10.0 ** rng.uniform(-5, 2, n)


def test_exchange_closed_form_and_charges():
    ru, rd = _rho(200, 0)
    w = np.random.default_rng(1).uniform(0.1, 2.0, ru.size)
    vh = np.random.default_rng(2).uniform(0.0, 3.0, ru.size)
    pot, e = O.dft_lsda(ru, rd, w, vh)
    cx = -0.75 * (6.0 / np.pi) ** (1.0 / 3.0)
    ex = np.sum(w * cx * (ru ** (4.0 / 3.0) + rd ** (4.0 / 3.0)))
    assert abs(e["ex"] - ex) < 1e-8 * abs(ex)  # PW91EX's own 12-digit constants
    assert abs(e["eh"] - 0.5 * np.sum(w * (ru + rd) * vh)) < 1e-13 * abs(e["eh"])
    assert abs(e["n_up"] - np.sum(w * ru)) < 1e-13 * e["n_up"]
    assert abs(e["n_dn"] - np.sum(w * rd)) < 1e-13 * e["n_dn"]
    # Slater potential -(6 rho_s / pi)^(1/3) plus correlation and V_H
    _, e1 = O.dft_lsda(ru[:1], rd[:1], np.ones(1), np.zeros(1))
    assert e1["ec"] < 0


def _exc_point(ru, rd):
    _, e = O.dft_lsda(np.array([ru]), np.array([rd]), np.ones(1), np.zeros(1))
    return e["ex"] + e["ec"]


def test_potential_is_energy_derivative():
    # |zeta| <= 0.9: towards |zeta| = 1 the minority-spin derivative of the (1 -/+ zeta)^(4/3) terms makes central
    # differences inaccurate; the fully polarised limit is pinned against the SIC path below
    rng = np.random.default_rng(3)
    ru = 10.0 ** rng.uniform(-5, 2, 40)
    rd = ru * 10.0 ** rng.uniform(-1.2, 1.2, 40)
    pot, _ = O.dft_lsda(ru, rd, np.ones(ru.size), np.zeros(ru.size))
    for k in range(ru.size):
        for s, r in ((0, ru[k]), (1, rd[k])):
            h = 1e-5 * r
            if s == 0:
                d = (_exc_point(ru[k] + h, rd[k]) - _exc_point(ru[k] - h, rd[k])) / (2 * h)
            else:
                d = (_exc_point(ru[k], rd[k] + h) - _exc_point(ru[k], rd[k] - h)) / (2 * h)
            assert abs(pot[s, k] - d) < 1e-7 * abs(d) + 1e-9, (k, s, pot[s, k], d)


def test_fully_polarised_limit_matches_the_sic_path():
    rho = 10.0 ** np.random.default_rng(4).uniform(-5, 2, 60)
    rg = np.zeros((rho.size, 10))
    rg[:, 0] = rho
    v_sic, e4 = O.xc_points(0, rg)
    pot, _ = O.dft_lsda(rho, np.zeros_like(rho), np.ones(rho.size), np.zeros(rho.size))
    np.testing.assert_allclose(pot[0], v_sic, rtol=1e-13, atol=1e-15)
    for k in range(rho.size):
        _, e = O.dft_lsda(rho[k:k + 1], np.zeros(1), np.ones(1), np.zeros(1))
        np.testing.assert_allclose(e["ex"] + e["ec"], e4[k].sum(), rtol=1e-13, atol=1e-15)
