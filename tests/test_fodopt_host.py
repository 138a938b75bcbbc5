"""The FOD optimiser's host logic (publicrelease_2020_amd/fodopt.py) on an analytic stand-in for the library:
an energy of FOD positions with a known gradient (harmonic wells plus a pairwise Gaussian repulsion), so the
L-BFGS steps, the Armijo backtracking, the FOD_OPT convergence rule (max FOD force < 1e-4) and freezing are
checked without a GPU.  The GPU form is exercised on the library in tests/test_gpu_fodopt.py."""
import numpy as np

from publicrelease_2020_amd import fodopt


class FakeEngine:
    def __init__(self, centres_up, centres_dn):
        self.c = [np.asarray(centres_up, float), np.asarray(centres_dn, float)]
        self.f = [None, None]
        self.n_eval = 0

    def set_fods(self, ispn, f):
        self.f[ispn - 1] = np.asarray(f, float).reshape(-1, 3).copy()

    def _parts(self):
        e = 0.0
        g = [np.zeros_like(self.f[0]), np.zeros_like(self.f[1])]
        for s in range(2):
            d = self.f[s] - self.c[s]
            e += float((d ** 2).sum())
            g[s] += 2 * d
            for i in range(len(self.f[s])):
                for j in range(i + 1, len(self.f[s])):
                    r = self.f[s][i] - self.f[s][j]
                    v = 0.5 * np.exp(-float(r @ r))
                    e += v
                    g[s][i] += -2 * r * v
                    g[s][j] += 2 * r * v
        return e, g

    def evaluate(self):
        self.n_eval += 1
        return self._parts()[0]

    def fod_gradient(self, ispn):
        return self._parts()[1][ispn - 1]


class FakeBatch:
    def __init__(self, engines):
        self.engines = engines

    def evaluate(self):
        return np.array([e.evaluate() for e in self.engines])


def _system(seed):
    rng = np.random.default_rng(seed)
    cu, cd = rng.normal(size=(4, 3)), rng.normal(size=(3, 3))
    return cu, cd, cu + rng.normal(scale=0.4, size=cu.shape), cd + rng.normal(scale=0.4, size=cd.shape)


def test_lbfgs_converges_with_descent():
    cu, cd, fu, fd = _system(1)
    eng = FakeEngine(cu, cd)
    res = fodopt.optimize(eng, fu, fd)
    assert res.converged and res.fmax < 1e-4
    energies = [h[1] for h in res.history]
    assert all(b <= a + 1e-14 for a, b in zip(energies, energies[1:]))
    eng.set_fods(1, res.fod_up)
    eng.set_fods(2, res.fod_dn)
    g = np.vstack([eng.fod_gradient(1), eng.fod_gradient(2)])
    assert np.sqrt((g ** 2).sum(1)).max() < 1e-4
    assert res.steps < 60


def test_frozen_fods_stay():
    cu, cd, fu, fd = _system(2)
    frozen = np.array([True, False, False, True])
    res = fodopt.optimize(FakeEngine(cu, cd), fu, fd, frozen_up=frozen)
    np.testing.assert_array_equal(res.fod_up[frozen], fu[frozen])
    assert np.abs(res.fod_up[~frozen] - fu[~frozen]).max() > 1e-3
    assert res.converged  # over the free coordinates


def test_frozen_mask_rule():
    atoms = np.array([[0.0, 0, 0], [1.5, 0, 0]])
    fods = np.array([[0.05, 0, 0], [1.52, 0, 0], [0.5, 0, 0]])
    # within 0.1 bohr of a non-hydrogen atom only
    np.testing.assert_array_equal(fodopt.frozen_mask(fods, atoms, [8, 1]), [True, False, False])


def test_batch_matches_single_systems():
    systems = [_system(s) for s in (3, 4, 5)]
    single = [fodopt.optimize(FakeEngine(cu, cd), fu, fd) for cu, cd, fu, fd in systems]
    engines = [FakeEngine(cu, cd) for cu, cd, _, _ in systems]
    res = fodopt.optimize_batch(FakeBatch(engines), engines, [(fu, fd) for _, _, fu, fd in systems])
    for a, b in zip(res, single):
        assert a.converged
        assert abs(a.energy - b.energy) < 1e-8
        np.testing.assert_allclose(a.fod_up, b.fod_up, atol=1e-4)


class SingularRegion(FakeEngine):
    """Raises the library's Lowdin failure (status 2) when two FODs of a spin come within 0.3 bohr."""

    def evaluate(self):
        for f in self.f:
            for i in range(len(f)):
                for j in range(i + 1, len(f)):
                    if np.linalg.norm(f[i] - f[j]) < 0.3:
                        err = RuntimeError("flosic status 2: Lowdin eigenvalue below 1e-8")
                        err.status = 2
                        raise err
        return super().evaluate()


def test_singular_trial_points_are_backtracked():
    # wells at the same point pull the FODs together; the optimiser must stop short of the singular region
    cu = np.zeros((2, 3))
    cd = np.zeros((1, 3))
    fu = np.array([[0.6, 0.0, 0.0], [-0.6, 0.0, 0.0]])
    fd = np.array([[0.1, 0.2, 0.0]])
    eng = SingularRegion(cu, cd)
    res = fodopt.optimize(eng, fu, fd, max_steps=50, max_step=1.0)
    assert np.linalg.norm(res.fod_up[0] - res.fod_up[1]) >= 0.3
    energies = [h[1] for h in res.history]
    assert all(b <= a + 1e-14 for a, b in zip(energies, energies[1:]))

