"""Far chunks of the three-piece split-BF16 Coulomb form (flosic_set_far_threshold), through the C ABI.

A (64-point tile, chunk) pair whose every (point, exponent-pair group) pair has T = p|r - P|^2 above the threshold is
contracted with two bf16 pieces (3 products) instead of three (6 products).  Checked here:
  threshold 0 or 1e30        no far rows; results bit-identical to each other (three pieces everywhere)
  thresholds 10 < 30 < 100   far rows fall as the threshold rises, and are a proper fraction of the stream
  default (30, POISSON1's THIRTYSOMETHING)   E_SIC within 1e-6 Ha and FOD forces within 1e-5 Ha/bohr of FP64,
                             V_H within the three-piece form's error plus the two-piece error on the far part
                             (|dV_H| < 1e-4 * max|V_H| against FP64)
  batch vs single            the batched launch uses each context's far bits: identical results
"""
import numpy as np
import pytest

from publicrelease_2020_amd import system as S

pytestmark = pytest.mark.gpu


def _F():
    from publicrelease_2020_amd import flosic as F
    return F


def _forces(e):
    return np.concatenate([e.fod_forces(1), e.fod_forces(2)])


def _far_fraction(e):
    c = e.counts()
    tiles = (c[2] + 63) // 64
    return c[9] / (tiles * c[6])


@pytest.fixture(scope="module")
def system():
    # 80 FLOs on a mesh of 8 atoms in a 7-bohr box: tight and diffuse exponent pairs at every distance
    return S.make_synthetic(8, 40, 21, S.XC_LSDA, nrad=32, nang=110, box=7.0)


def test_far_threshold_fractions_and_off_switch(system):
    F = _F()
    e = F.Flosic(0)
    e.load_system(system, F.PREC_BF16X3)
    out = {}
    for t in (0.0, 1e30, 100.0, 30.0, 10.0):
        e.set_far_threshold(t)
        z = e.evaluate()
        out[t] = dict(z=z, frac=_far_fraction(e), vh=e.potential(1, 1)[1], f=_forces(e))
        print(f"T_far {t:g}: far rows {out[t]['frac']:.3f}, E_SIC {z!r}", flush=True)
    assert out[0.0]["frac"] == 0 and out[1e30]["frac"] == 0
    assert out[0.0]["z"] == out[1e30]["z"]
    assert np.array_equal(out[0.0]["vh"], out[1e30]["vh"])
    assert np.array_equal(out[0.0]["f"], out[1e30]["f"])
    assert 0 < out[100.0]["frac"] < out[30.0]["frac"] < out[10.0]["frac"] < 1
    # the default is 30
    d = F.Flosic(0)
    d.load_system(system, F.PREC_BF16X3)
    zd = d.evaluate()
    assert _far_fraction(d) == out[30.0]["frac"] and zd == out[30.0]["z"]
    d.close()
    e.close()


def test_far_chunks_within_gate(system):
    F = _F()
    ref = F.Flosic(0)
    ref.load_system(system, F.PREC_FP64)
    z64 = ref.evaluate()
    f64 = _forces(ref)
    v64 = [ref.potential(s, i)[1] for s, i in ((1, 1), (1, 40), (2, 7))]
    ref.close()
    e = F.Flosic(0)
    e.load_system(system, F.PREC_BF16X3)
    for t in (30.0, 10.0):
        e.set_far_threshold(t)
        z = e.evaluate()
        dF = float(np.abs(_forces(e) - f64).max())
        dV = max(float(np.abs(e.potential(s, i)[1] - v).max() / np.abs(v).max())
                 for (s, i), v in zip(((1, 1), (1, 40), (2, 7)), v64))
        print(f"T_far {t:g}: far rows {_far_fraction(e):.3f}, |dE| {abs(z - z64):.2e} Ha, max |dF| {dF:.2e}, "
              f"max |dV_H|/max|V_H| {dV:.2e}", flush=True)
        assert abs(z - z64) < 1e-6
        assert dF < 1e-5
        assert dV < 1e-4
    e.close()


def test_batch_uses_far_bits(system):
    F = _F()
    small = S.make_synthetic(4, 4, 5, S.XC_PBE, nrad=24, nang=50, box=5.0)

    def engines():
        out = []
        for sysm in (system, small):
            e = F.Flosic(0)
            e.load_system(sysm, F.PREC_BF16X3)
            out.append(e)
        return out

    single = engines()
    z1 = [e.evaluate() for e in single]
    v1 = [e.potential(1, 1)[1] for e in single]
    assert _far_fraction(single[0]) > 0
    batch = F.FlosicBatch(engines())
    z = batch.evaluate()
    np.testing.assert_allclose(z, z1, rtol=1e-13, atol=0)
    for e, v in zip(batch.engines, v1):
        np.testing.assert_allclose(e.potential(1, 1)[1], v, rtol=0, atol=1e-13 * np.abs(v).max())
    batch.close()
    for e in single:
        e.close()
