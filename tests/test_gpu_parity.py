"""Parity of the HIP path (through the C ABI) with the CPU FP64 oracle, and properties at full size.

Tolerances (FP64 on both sides; differences come from summation order only):
  E_SIC                     |dE| < 1e-10 Ha
  RESULTS, SIC, V^i_ab      |d| < 1e-10 * max|x| + 1e-12
  FOD forces / gradient     |d| < 1e-10 Ha/bohr
  integer pair list         bit-exact
"""
import numpy as np
import pytest

from publicrelease_2020_amd import system as S

pytestmark = pytest.mark.gpu


def _flosic():
    from publicrelease_2020_amd import flosic as F
    return F


def _run(sysm, precision=0):
    F = _flosic()
    ctx = F.Flosic(0)
    ctx.load_system(sysm, precision)
    esic = ctx.evaluate()
    res = ctx.results()
    res["esic"] = esic
    return ctx, res


def _close(a, b, rel=1e-10, abs_=1e-12):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape
    tol = rel * max(np.abs(b).max(), 1e-300) + abs_
    err = np.abs(a - b).max()
    assert err <= tol, f"max abs diff {err:.3e} > {tol:.3e}"


SMALL = [
    ("K3-LSDA", dict(K=3, nocc=3, seed=7, xc=S.XC_LSDA)),
    ("K5-LSDA", dict(K=5, nocc=5, seed=11, xc=S.XC_LSDA)),
    ("K6-PBE", dict(K=6, nocc=6, seed=3, xc=S.XC_PBE)),
    ("K3-PBE", dict(K=3, nocc=3, seed=7, xc=S.XC_PBE)),
]


@pytest.fixture(scope="module", params=SMALL, ids=[s[0] for s in SMALL])
def small_case(request):
    from oracle import oracle as O
    a = request.param[1]
    sysm = S.make_synthetic(a["K"], a["nocc"], a["seed"], a["xc"], nrad=24, nang=50, box=5.0)
    ref = O.evaluate(sysm, want_potentials=True)
    assert ref["status"] == 0
    ctx, got = _run(sysm)
    return sysm, ref, ctx, got


def test_pair_list_bit_exact(small_case):
    from oracle import oracle as O
    sysm, ref, ctx, got = small_case
    assert np.array_equal(ctx.pairs(), O.pairs(sysm.basis))


def test_fermi_orbitals_parity(small_case):
    sysm, ref, ctx, got = small_case
    for s in (1, 2):
        fo = ctx.fermi_orbitals(s)
        r = ref["spins"][s - 1]
        _close(fo["fmat"], r["fmat"], 1e-11)
        _close(fo["over"], r["over"], 1e-11)
        _close(fo["q"], r["q"], 1e-11)
        _close(fo["tmat"], r["tmat"], 1e-10)


def test_energy_and_results_parity(small_case):
    sysm, ref, ctx, got = small_case
    assert abs(got["esic"] - ref["esic"]) < 1e-10
    for s in range(2):
        _close(got["spins"][s]["results"], ref["spins"][s]["results"])


def test_sic_matrix_and_vab_parity(small_case):
    sysm, ref, ctx, got = small_case
    for s in range(2):
        _close(got["spins"][s]["sic"], ref["spins"][s]["sic"])
        _close(got["spins"][s]["vab"], ref["spins"][s]["vab"])


def test_fod_forces_parity(small_case):
    sysm, ref, ctx, got = small_case
    for s in range(2):
        _close(got["spins"][s]["debdax"], ref["spins"][s]["debdax"], 1e-9, 1e-12)
        _close(got["spins"][s]["grad"], ref["spins"][s]["grad"], 1e-9, 1e-12)
        if sysm.xc == S.XC_LSDA:  # PBE: the gradient is the exact mesh derivative, DEBDAX the FRMORB2 formula
            _close(got["spins"][s]["grad"], got["spins"][s]["debdax"], 1e-10, 1e-13)


def test_potentials_parity(small_case):
    """ZPOT analogue: V_i = V_xc + V_H on the mesh and V_H alone."""
    sysm, ref, ctx, got = small_case
    for s in (1, 2):
        for i in range(sysm.nocc[s - 1]):
            pot, vh = ctx.potential(s, i + 1)
            _close(vh, ref["spins"][s - 1]["vh"][i], 1e-11)
            _close(pot, ref["spins"][s - 1]["vpot"][i], 1e-10)


def test_h2o_parity():
    """The water molecule with 5 FODs per spin (coarse mesh so the oracle runs in seconds)."""
    from oracle import oracle as O
    sysm = S.make_h2o(seed=1, nrad=24, nang=50)
    ref = O.evaluate(sysm)
    ctx, got = _run(sysm)
    assert abs(got["esic"] - ref["esic"]) < 1e-10
    for s in range(2):
        _close(got["spins"][s]["results"], ref["spins"][s]["results"])
        _close(got["spins"][s]["sic"], ref["spins"][s]["sic"])
        _close(got["spins"][s]["debdax"], ref["spins"][s]["debdax"], 1e-9, 1e-12)


def test_lowdin_failure_is_reported():
    F = _flosic()
    sysm = S.make_synthetic(3, 3, 7, nrad=16, nang=26, box=5.0)
    sysm.fod[0][1] = sysm.fod[0][0]
    ctx = F.Flosic(0)
    ctx.load_system(sysm)
    with pytest.raises(F.FlosicError) as e:
        ctx.evaluate()
    assert e.value.status == F.ELOWDIN


def test_bad_fod_count_is_rejected():
    F = _flosic()
    sysm = S.make_synthetic(3, 3, 7, nrad=16, nang=26, box=5.0)
    ctx = F.Flosic(0)
    ctx.load_system(sysm)
    with pytest.raises(F.FlosicError) as e:
        ctx.set_fods(1, sysm.fod[0][:2])
    assert e.value.status == F.EINVAL


@pytest.fixture(scope="module")
def k16():
    sysm = S.make_config(2)
    ctx, got = _run(sysm)
    return sysm, ctx, got


def test_k16_properties(k16):
    """Full-size (K=16, 186k points) size-independent checks: unit FLO charge up to mesh quadrature,
    lambda_ii = XDEL, V^i_ai = -SIC(a,i), adjoint gradient = FRMORB2 formula, E_SIC = sum RESULTS(9)."""
    sysm, ctx, got = k16
    esum = 0.0
    for s in range(2):
        sp = got["spins"][s]
        n = sysm.nocc[s]
        assert np.all(np.abs(sp["results"][:, 1] - 1.0) < 2e-3)
        _close(np.diag(sp["sic"][:n, :n]), sp["results"][:, 6], 1e-11)
        _close(np.array([sp["vab"][i, :, i] for i in range(n)]).T, -sp["sic"][:n, :n], 1e-10)
        _close(sp["grad"], sp["debdax"], 1e-9, 1e-13)
        esum += sp["results"][:, 8].sum()
    assert abs(esum - got["esic"]) < 1e-12


def test_k16_invariant_under_occupied_rotation(k16):
    sysm, ctx, got = k16
    rng = np.random.default_rng(17)
    F = _flosic()
    ctx2 = F.Flosic(0)
    ctx2.load_system(sysm)
    for s in (1, 2):
        n = sysm.nocc[s - 1]
        Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
        C2 = sysm.coef[s - 1].copy()
        C2[:, :n] = C2[:, :n] @ Q
        ctx2.set_orbitals(s, C2, n)
        ctx2.set_fods(s, sysm.fod[s - 1])
    e2 = ctx2.evaluate()
    assert abs(e2 - got["esic"]) < 1e-9
    for s in (1, 2):
        _close(ctx2.fod_forces(s), got["spins"][s - 1]["debdax"], 1e-8, 1e-12)


def test_k16_repeat_is_deterministic(k16):
    sysm, ctx, got = k16
    e2 = ctx.evaluate()
    assert e2 == got["esic"]


@pytest.mark.parametrize("prec", [1, 2])
def test_mixed_precision_within_tolerance(small_case, prec):
    """FP32-MFMA (1) and 3-piece split-BF16 (2) Coulomb with FP64 chunk sums: |dE_SIC| < 1e-6 Ha and FOD forces
    within 1e-5 Ha/bohr of the FP64 oracle."""
    sysm, ref, ctx, got = small_case
    F = _flosic()
    m = F.Flosic(0)
    m.load_system(sysm, prec)
    e = m.evaluate()
    assert abs(e - ref["esic"]) < 1e-6
    for s in (1, 2):
        assert np.abs(m.fod_forces(s) - ref["spins"][s - 1]["debdax"]).max() < 1e-5
        _close(m.sic_vab(s), ref["spins"][s - 1]["vab"], 1e-5, 1e-8)


def test_valu_coulomb_form_matches(small_case, monkeypatch):
    """The per-thread VALU Coulomb form gives the same potentials as the MFMA form (both FP64)."""
    sysm, ref, ctx, got = small_case
    F = _flosic()
    monkeypatch.setenv("FLOSIC_COULOMB", "valu")
    v = F.Flosic(0)
    v.load_system(sysm)
    e = v.evaluate()
    assert abs(e - ref["esic"]) < 1e-10
    for s in (1, 2):
        pot, vh = v.potential(s, 1)
        _close(vh, ref["spins"][s - 1]["vh"][0], 1e-11)


def test_bf16x2_error_is_bounded(small_case):
    """The 2-piece split (3 products, ~2^-16 per product) stays within 1e-4 Ha of the FP64 oracle."""
    sysm, ref, ctx, got = small_case
    F = _flosic()
    m = F.Flosic(0)
    m.load_system(sysm, F.PREC_BF16X2)
    e = m.evaluate()
    assert abs(e - ref["esic"]) < 1e-4


def test_lagmult_slv_parity(small_case):
    """LAGMULT_SLV: <phi_j| -1/2 lap + V_DFT |phi_k> added to the SIC block (Laplacian plane for LSDA, Hessian
    planes for PBE) against the oracle, with a synthetic DFT potential on the mesh."""
    from oracle import oracle as O
    sysm, ref, ctx, got = small_case
    r = np.linalg.norm(sysm.mesh.xyz - sysm.mesh.xyz.mean(0), axis=1)
    pot = -1.5 / np.sqrt(r * r + 0.3) + 0.05 * sysm.mesh.xyz[:, 0]
    for s in (1, 2):
        want = O.lagmult(sysm, s, pot)
        _close(ctx.lagmult_slv(s, pot), want, 1e-10)
        base = ctx.siclagm(s)
        n = sysm.nocc[s - 1]
        out = ctx.lagmult_slv(s, pot, base)
        _close(out[:n, :n], base[:n, :n] + want, 1e-10)
        assert np.array_equal(out[n:, :], base[n:, :])


def test_sic_hamiltonian_parity(small_case):
    """APATCH: the AO-basis SIC Hamiltonian against the oracle."""
    from oracle import oracle as O
    sysm, ref, ctx, got = small_case
    for s in (1, 2):
        want = O.sic_hamiltonian(sysm, s, ref["spins"][s - 1]["vpot"])
        h = ctx.sic_hamiltonian(s)
        _close(h, want, 1e-10)
        assert np.array_equal(h, h.T)
