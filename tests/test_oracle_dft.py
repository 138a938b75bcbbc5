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


# ------------------------------------------------------------------------------------------------ PBE (IGGA = 1)
def _radial_fields(r, cs, es):
    """rho(r) = sum c e^{-e r^2} on the ray r*(1,1,1)/sqrt(3): value, gradient, Hessian (10 comps)."""
    u = np.ones(3) / np.sqrt(3.0)
    f = sum(c * np.exp(-e * r * r) for c, e in zip(cs, es))
    f1 = sum(-2 * e * r * c * np.exp(-e * r * r) for c, e in zip(cs, es))
    f2 = sum((4 * e * e * r * r - 2 * e) * c * np.exp(-e * r * r) for c, e in zip(cs, es))
    rg = np.zeros((r.size, 10))
    rg[:, 0] = f
    rg[:, 1:4] = f1[:, None] * u
    H = f2[:, None, None] * np.outer(u, u) + (f1 / r)[:, None, None] * (np.eye(3) - np.outer(u, u))
    rg[:, 4], rg[:, 5], rg[:, 6] = H[:, 0, 0], H[:, 1, 1], H[:, 2, 2]
    rg[:, 7], rg[:, 8], rg[:, 9] = H[:, 0, 1], H[:, 0, 2], H[:, 1, 2]
    return rg


def _random_fields(n, seed):
    rng = np.random.default_rng(seed)
    rg = rng.normal(size=(n, 10)) * 10.0 ** rng.uniform(-3, 1, (n, 1))
    rg[:, 0] = 10.0 ** rng.uniform(-4, 2, n)
    return rg


def test_gga_fully_polarised_limit_matches_the_sic_path():
    rg = _random_fields(80, 5)
    v_sic, e4 = O.xc_points(1, rg)
    pot, e = O.dft_gga(rg, np.zeros_like(rg), np.ones(rg.shape[0]), np.zeros(rg.shape[0]))
    # the two C paths see identical inputs; the compiler contracts a*b+c into fused multiply-adds differently in
    # each, which moves a potential formed with cancellation by ~1e-12 relative
    np.testing.assert_allclose(pot[0], v_sic, rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(e["ex"], e4[:, :2].sum(), rtol=1e-12)
    np.testing.assert_allclose(e["ec"], e4[:, 2:].sum(), rtol=1e-12)


def test_gga_spin_symmetry():
    ru, rd = _random_fields(80, 6), _random_fields(80, 7)
    w, vh = np.random.default_rng(8).uniform(0.1, 2.0, 80), np.random.default_rng(9).uniform(0.0, 3.0, 80)
    p1, e1 = O.dft_gga(ru, rd, w, vh)
    p2, e2 = O.dft_gga(rd, ru, w, vh)
    np.testing.assert_allclose(p2[::-1], p1, rtol=1e-12, atol=1e-14)  # swapping the spins swaps the potentials
    for k in ("ex", "ec", "eh"):
        assert abs(e1[k] - e2[k]) < 1e-12 * abs(e1[k])
    assert e1["n_up"] == e2["n_dn"] and e1["n_dn"] == e2["n_up"]
    p3, _ = O.dft_gga(ru, ru, w, vh)  # unpolarised: both spins see the same potential
    np.testing.assert_allclose(p3[0], p3[1], rtol=1e-12, atol=1e-14)


def test_gga_potential_is_the_functional_derivative_for_each_spin():
    """V_s = dE_xc/d rho_s: spherical spin densities of different shapes, each perturbed in turn by g,
    dE/d(delta) = int V_s g (radial quadrature, 40k points); this exercises the zeta-gradient term ww and the
    down-spin potential VCDN + DVCDN."""
    r = np.linspace(1e-4, 14.0, 40000)
    wr = 4 * np.pi * r * r * np.gradient(r)
    ru = _radial_fields(r, [1.0, 0.08], [1.0, 0.25])
    rd = _radial_fields(r, [0.5, 0.05], [0.7, 0.3])
    g = _radial_fields(r, [1.0], [0.6])
    zero = np.zeros(r.size)
    pot, _ = O.dft_gga(ru, rd, wr, zero)
    dl = 1e-4
    for s in (0, 1):
        a = [ru, rd]
        a[s] = a[s] + dl * g
        _, ep = O.dft_gga(a[0], a[1], wr, zero)
        a[s] = a[s] - 2 * dl * g
        _, em = O.dft_gga(a[0], a[1], wr, zero)
        dE = ((ep["ex"] + ep["ec"]) - (em["ex"] + em["ec"])) / (2 * dl)
        dv = (pot[s] * g[:, 0] * wr).sum()
        assert abs(dE - dv) < 2e-6 * abs(dv), (s, dE, dv)


def test_dft_matches_reference_subvlxc(golden):
    """The reference's own SUBVLXC (NSPN = 2, no centres; tests/golden/make_ref_vectors.py) on random spin densities
    with gradients and Hessians, ten of them fully polarised: LSDA (IGGA = 0, PW91EX + PW91LC/PW91NC) and PBE."""
    rgu, rgd, w = golden["dft_rgu"], golden["dft_rgd"], golden["dft_w"]
    zero = np.zeros(w.size)
    for name, pot_e in (("lsda", O.dft_lsda(rgu[:, 0], rgd[:, 0], w, zero)), ("pbe", O.dft_gga(rgu, rgd, w, zero))):
        pot, e = pot_e
        ref_pot, ref_e = golden[f"dft_pot_{name}"], golden[f"dft_e_{name}"]
        err = np.abs(pot - ref_pot).max() / np.abs(ref_pot).max()
        print(name, "max rel potential error", err, "energies", e["ex"], ref_e[0] + ref_e[1], e["ec"], ref_e[2] + ref_e[3])
        assert err < 1e-12, name
        assert abs(e["ex"] - (ref_e[0] + ref_e[1])) < 1e-12 * abs(ref_e[0]), name
        assert abs(e["ec"] - (ref_e[2] + ref_e[3])) < 1e-12 * abs(ref_e[2]), name
