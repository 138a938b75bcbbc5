"""FOD optimisation driven by the library's FOD gradient, in process (the role of FOD_OPT / FOD_LBFGS,
electronic_geometry_lbfgs.f:11-441, which the reference runs as one step per program invocation with its state
in files).

The energy is E_SIC(FODs) from flosic_evaluate and the gradient flosic_fod_gradient (the derivative of the
discrete energy).  The minimiser is limited-memory BFGS (the two-loop recursion of Nocedal, "Updating
quasi-Newton matrices with limited storage", Math. Comp. 35 (1980) 773) with the reference's settings: initial
inverse-Hessian scale DGUESS = 0.1 (electronic_geometry_lbfgs.f:200), convergence when the largest FOD force
|dE/da_i| falls below 1e-4 Ha/bohr (:171), optional freezing of FODs within 0.1 bohr of a non-hydrogen atom
(:112-122).  The line search is a backtracking search on the Armijo condition with a step cap, so every accepted
step lowers E_SIC.  `optimize_batch` moves the FODs of many systems at once, one batched evaluation per step.

Engines are duck-typed: set_fods(ispn, fods), evaluate() -> E_SIC, fod_gradient(ispn) -> (n, 3).
"""
from __future__ import annotations

import dataclasses
import math
import re
from typing import Callable, List, Optional, Sequence

import numpy as np

ELOWDIN = 2  # status of a singular Fermi-orbital overlap (include/flosic_hip.h FLOSIC_ELOWDIN)


@dataclasses.dataclass
class OptResult:
    fod_up: np.ndarray
    fod_dn: np.ndarray
    energy: float
    fmax: float
    steps: int
    evaluations: int
    converged: bool
    history: List[tuple]  # (step, energy, |g|, fmax) as FOD_OPT writes to fande.dat


def frozen_mask(fods: np.ndarray, atoms_xyz: np.ndarray, atoms_Z: Sequence[int], radius: float = 0.1) -> np.ndarray:
    """True for FODs that stay fixed: within `radius` bohr of an atom other than hydrogen (core FODs)."""
    fods = np.asarray(fods, dtype=np.float64).reshape(-1, 3)
    mask = np.zeros(fods.shape[0], dtype=bool)
    for a, z in zip(np.asarray(atoms_xyz).reshape(-1, 3), atoms_Z):
        if int(z) != 1:
            mask |= np.linalg.norm(fods - a, axis=1) < radius
    return mask


class _LBFGS:
    """Two-loop recursion over the last `memory` (s, y) pairs on a flat coordinate vector."""

    def __init__(self, memory: int, dguess: float):
        self.m = memory
        self.dguess = dguess
        self.s: List[np.ndarray] = []
        self.y: List[np.ndarray] = []

    def direction(self, g: np.ndarray) -> np.ndarray:
        q = g.copy()
        alphas = []
        for s, y in zip(reversed(self.s), reversed(self.y)):
            rho = 1.0 / float(y @ s)
            a = rho * float(s @ q)
            alphas.append((rho, a))
            q -= a * y
        if self.s:
            gamma = float(self.s[-1] @ self.y[-1]) / float(self.y[-1] @ self.y[-1])
        else:
            gamma = self.dguess
        r = gamma * q
        for (s, y), (rho, a) in zip(zip(self.s, self.y), reversed(alphas)):
            b = rho * float(y @ r)
            r += (a - b) * s
        return -r

    def update(self, s: np.ndarray, y: np.ndarray):
        if float(s @ y) > 1e-12 * float(np.linalg.norm(s) * np.linalg.norm(y)):  # keep the update positive
            self.s.append(s)
            self.y.append(y)
            if len(self.s) > self.m:
                self.s.pop(0)
                self.y.pop(0)


class _Problem:
    """Energy and gradient of one engine over the free FOD coordinates."""

    def __init__(self, engine, fod_up, fod_dn, frozen_up, frozen_dn):
        self.engine = engine
        self.base = [np.array(fod_up, dtype=np.float64).reshape(-1, 3), np.array(fod_dn, dtype=np.float64).reshape(-1, 3)]
        self.free = [~np.asarray(frozen_up, dtype=bool), ~np.asarray(frozen_dn, dtype=bool)]
        self.evals = 0

    def x0(self) -> np.ndarray:
        return np.concatenate([self.base[s][self.free[s]].ravel() for s in range(2)])

    def fods(self, x: np.ndarray):
        out, k = [], 0
        for s in range(2):
            f = self.base[s].copy()
            n = int(self.free[s].sum())
            f[self.free[s]] = x[k:k + 3 * n].reshape(n, 3)
            k += 3 * n
            out.append(f)
        return out

    def set(self, x: np.ndarray):
        up, dn = self.fods(x)
        self.engine.set_fods(1, up)
        self.engine.set_fods(2, dn)

    def grad(self) -> np.ndarray:
        return np.concatenate([np.asarray(self.engine.fod_gradient(s + 1)).reshape(-1, 3)[self.free[s]].ravel()
                               for s in range(2)])

    def value_grad(self, x: np.ndarray, qmin_floor: float = 0.0):
        """E_SIC and gradient at x.  A trial point whose Fermi-orbital overlap is singular (the library's
        FLOSIC_ELOWDIN, the reference's STOPIT in LOWSIC) or has an eigenvalue below qmin_floor counts as an
        infinite energy."""
        self.set(x)
        self.evals += 1
        try:
            e = float(self.engine.evaluate())
        except Exception as err:  # noqa: BLE001 - only the Lowdin failure is a property of the point
            if getattr(err, "status", None) == ELOWDIN:
                return math.inf, None
            raise
        if qmin_floor > 0.0 and self.qmin() < qmin_floor:
            return math.inf, None
        return e, self.grad()

    def qmin(self) -> float:
        """Smallest eigenvalue of the Fermi-orbital overlaps of both spins (engine.lowdin_q(ispn))."""
        return min(float(np.min(self.engine.lowdin_q(s))) for s in (1, 2) if len(self.base[s - 1]))


def _fmax(g: np.ndarray) -> float:
    return float(np.sqrt((g.reshape(-1, 3) ** 2).sum(axis=1).max())) if g.size else 0.0


def optimize(engine, fod_up, fod_dn, fmax_tol: float = 1e-4, max_steps: int = 200, memory: int = 8,
             dguess: float = 0.1, max_step: float = 0.3, frozen_up=None, frozen_dn=None,
             qmin_floor: float = 0.0, callback: Optional[Callable[[int, float, float], None]] = None) -> OptResult:
    """Minimise E_SIC over the FOD positions of one system; returns the final FODs and the FOD_OPT history.
    qmin_floor > 0 keeps the Fermi-orbital overlap eigenvalues above it (trial points below count as infinite
    energy; FRMORB2 flags eigenvalues below 0.01 as bothersome states, frmorb2.F:539-550), for energy surfaces
    that fall towards linearly dependent Fermi orbitals; the engine then provides lowdin_q(ispn)."""
    fu = np.asarray(fod_up, dtype=np.float64).reshape(-1, 3)
    fd = np.asarray(fod_dn, dtype=np.float64).reshape(-1, 3)
    P = _Problem(engine, fu, fd, np.zeros(len(fu), bool) if frozen_up is None else frozen_up,
                 np.zeros(len(fd), bool) if frozen_dn is None else frozen_dn)
    x = P.x0()
    e, g = P.value_grad(x)
    opt = _LBFGS(memory, dguess)
    hist = [(0, e, float(np.linalg.norm(g)), _fmax(g))]
    step = 0
    while _fmax(g) >= fmax_tol and step < max_steps:
        d = opt.direction(g)
        if float(d @ g) >= 0.0:  # not a descent direction: restart from steepest descent
            opt = _LBFGS(memory, dguess)
            d = -dguess * g
        dmax = np.sqrt((d.reshape(-1, 3) ** 2).sum(axis=1).max())
        t = min(1.0, max_step / dmax) if dmax > 0 else 1.0
        accepted = False
        for _ in range(30):  # backtracking on the Armijo condition
            xn = x + t * d
            en, gn = P.value_grad(xn, qmin_floor)
            if en <= e + 1e-4 * t * float(d @ g):
                accepted = True
                break
            t *= 0.5
        if not accepted:
            P.value_grad(x)  # leave the engine at the best point
            break
        if gn is None:
            break
        opt.update(xn - x, gn - g)
        x, e, g = xn, en, gn
        step += 1
        hist.append((step, e, float(np.linalg.norm(g)), _fmax(g)))
        if callback:
            callback(step, e, _fmax(g))
    up, dn = P.fods(x)
    return OptResult(fod_up=up, fod_dn=dn, energy=e, fmax=_fmax(g), steps=step, evaluations=P.evals,
                     converged=_fmax(g) < fmax_tol, history=hist)


def optimize_batch(batch, engines, fods: Sequence[tuple], fmax_tol: float = 1e-4, max_steps: int = 200,
                   memory: int = 8, dguess: float = 0.1, max_step: float = 0.3) -> List[OptResult]:
    """Minimise E_SIC of many systems at once: every trial step of every unconverged system is evaluated in one
    batch evaluation (`batch.evaluate()`, e.g. a replayed graph); each system keeps its own L-BFGS state and
    backtracking.  Converged systems keep their FODs (their evaluations repeat the same point)."""
    n = len(engines)
    P = [_Problem(engines[i], fods[i][0], fods[i][1], np.zeros(len(np.reshape(fods[i][0], (-1, 3))), bool),
                  np.zeros(len(np.reshape(fods[i][1], (-1, 3))), bool)) for i in range(n)]
    x = [p.x0() for p in P]
    for i in range(n):
        P[i].set(x[i])
    z = batch.evaluate()
    e = [float(v) for v in z]
    g = [P[i].grad() for i in range(n)]
    opt = [_LBFGS(memory, dguess) for _ in range(n)]
    hist = [[(0, e[i], float(np.linalg.norm(g[i])), _fmax(g[i]))] for i in range(n)]
    steps = [0] * n
    active = [_fmax(g[i]) >= fmax_tol for i in range(n)]
    d = [None] * n
    t = [1.0] * n
    tries = [0] * n
    evals = 1
    while any(active) and evals < 30 * max_steps:
        for i in range(n):
            if not active[i]:
                P[i].set(x[i])
                continue
            if tries[i] == 0:
                d[i] = opt[i].direction(g[i])
                if float(d[i] @ g[i]) >= 0.0:
                    opt[i] = _LBFGS(memory, dguess)
                    d[i] = -dguess * g[i]
                dmax = np.sqrt((d[i].reshape(-1, 3) ** 2).sum(axis=1).max())
                t[i] = min(1.0, max_step / dmax) if dmax > 0 else 1.0
            P[i].set(x[i] + t[i] * d[i])
        try:
            z = batch.evaluate()
        except Exception as err:  # noqa: BLE001
            m = re.search(r"system (\d+):", str(err))
            if getattr(err, "status", None) != ELOWDIN or not m:
                raise
            j = int(m.group(1))  # that system's trial point is singular: shorten its step and evaluate again
            t[j] *= 0.5
            tries[j] += 1
            evals += 1
            if tries[j] >= 30:
                active[j] = False
            continue
        evals += 1
        for i in range(n):
            if not active[i]:
                continue
            en, gn = float(z[i]), P[i].grad()
            if en <= e[i] + 1e-4 * t[i] * float(d[i] @ g[i]):
                xn = x[i] + t[i] * d[i]
                opt[i].update(xn - x[i], gn - g[i])
                x[i], e[i], g[i] = xn, en, gn
                steps[i] += 1
                tries[i] = 0
                hist[i].append((steps[i], e[i], float(np.linalg.norm(g[i])), _fmax(g[i])))
                if _fmax(g[i]) < fmax_tol or steps[i] >= max_steps:
                    active[i] = False
            else:
                t[i] *= 0.5
                tries[i] += 1
                if tries[i] >= 30:
                    active[i] = False
    out = []
    for i in range(n):
        up, dn = P[i].fods(x[i])
        out.append(OptResult(fod_up=up, fod_dn=dn, energy=e[i], fmax=_fmax(g[i]), steps=steps[i], evaluations=evals,
                             converged=_fmax(g[i]) < fmax_tol, history=hist[i]))
    return out
