"""Synthetic FLO-SIC systems: atoms, Cartesian Gaussian basis, Becke mesh, KS orbitals, FODs.

This module is host-side set-up (the part of the reference that lives in READINP/ISETUP/VMESH and the
SCF): it produces the data the SIC hot path consumes, in the layout the reference keeps in its module
data.

Basis layout (reference: ``gorbdrv.f:18-316``, ``frmorb2.F:262-266``, ``coupot_sic.F:380-400``):
  * every atom carries a list of "bare" primitive exponents (``BFALP(I_BARE,IFNCT)``);
  * for each angular momentum l = 0, 1, 2 it carries ``N_CON(l)`` contractions, each a coefficient vector
    over the bare exponents (``BFCON(I_BARE,ICON,l,IFNCT)``), normalisation folded in;
  * Cartesian components are s; x, y, z; xx, yy, zz, xy, xz, yz (``gorbdrv.f:78-98``);
  * the local order of basis functions on a site is l outer, Cartesian component, contraction inner.

Mesh: per-atom log radial grid times a randomly rotated Fibonacci sphere, Becke fuzzy-cell weights.
KS orbitals: seeded Gaussian coefficients, Löwdin-orthonormalised in the analytic AO overlap metric.
FODs: one per heavy atom, the rest uniform in 1.5-bohr balls around seeded atoms.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional, Sequence

import ctypes
import os
import subprocess

import numpy as np

CUTEXP = 40.0  # PARAMA2:31
_HERE = os.path.dirname(os.path.abspath(__file__))
_SETUP_SRC = os.path.join(_HERE, "csrc", "setup_host.cpp")
_SETUP_LIB = os.path.join(_HERE, "libflosic_setup.so")
_setup = None


def build_setup_lib() -> str:
    """Compile the OpenMP set-up helpers (Becke weights, AO overlap) in-tree."""
    tmp = _SETUP_LIB + ".tmp"
    subprocess.check_call(["g++", "-O3", "-fopenmp", "-fPIC", "-shared", "-std=c++17", _SETUP_SRC, "-o", tmp])
    os.replace(tmp, _SETUP_LIB)
    return _SETUP_LIB


def _setup_lib():
    global _setup
    if _setup is None:
        if not os.path.exists(_SETUP_LIB) or os.path.getmtime(_SETUP_LIB) < os.path.getmtime(_SETUP_SRC):
            build_setup_lib()
        L = ctypes.CDLL(_SETUP_LIB)
        dp, ip = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32)
        L.setup_becke.argtypes = [ctypes.c_int, dp, ctypes.c_int64, dp, ip, dp]
        L.setup_overlap.argtypes = [ctypes.c_int, dp, ip, dp, ip, ip, dp, ctypes.c_int, dp]
        _setup = L
    return _setup


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))

# Cartesian powers in reference order (gorbdrv.f:78-98)
CART = {
    0: [(0, 0, 0)],
    1: [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
    2: [(2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1), (0, 1, 1)],
}
NCART = {0: 1, 1: 3, 2: 6}

# exponent ladder: 0.1 * 3**k, k = 0..10  (0.1 ... 5904.9 bohr^-2)
EXPONENTS = 0.1 * 3.0 ** np.arange(11)
# which ladder entries each element carries, per l
SHELL_RULES = {
    1: {0: list(range(0, 6)), 1: [0, 1]},                      # H: s x6, p x2 (most diffuse p)
    6: {0: list(range(1, 11)), 1: list(range(1, 7)), 2: [2, 3]},  # C: s x10, p x6, d x2
    8: {0: list(range(1, 11)), 1: list(range(1, 7)), 2: [2, 3]},  # O: same ladder as C
}
XC_LSDA = 0
XC_PBE = 1


def prim_norm(alpha: float, l: int) -> float:
    """Normalisation of an xy-type Cartesian primitive x^a y^b z^c e^{-alpha r^2} with a,b,c <= 1."""
    return (2.0 * alpha / math.pi) ** 0.75 * (4.0 * alpha) ** (0.5 * l)


@dataclasses.dataclass
class Basis:
    """General-contraction Cartesian Gaussian basis, flat arrays in C-ABI layout."""

    xyz: np.ndarray          # (natom, 3)
    Z: np.ndarray            # (natom,) int
    bare_off: np.ndarray     # (natom+1,) int32 offsets into alpha
    alpha: np.ndarray        # (sum nbare,)
    ncon: np.ndarray         # (natom, 3) int32 contractions per l (0 if l absent)
    con_off: np.ndarray      # (natom, 3) int32 offsets into con
    con: np.ndarray          # flat: for (atom, l) a (ncon, nbare) row-major block

    @property
    def natom(self) -> int:
        return int(self.xyz.shape[0])

    def nbare(self, a: int) -> int:
        return int(self.bare_off[a + 1] - self.bare_off[a])

    def con_block(self, a: int, l: int) -> np.ndarray:
        nb = self.nbare(a)
        nc = int(self.ncon[a, l])
        o = int(self.con_off[a, l])
        return self.con[o:o + nc * nb].reshape(nc, nb)

    def functions(self):
        """List of (atom, l, cart_index, contraction) in reference local order."""
        out = []
        for a in range(self.natom):
            for l in range(3):
                for m in range(NCART[l]):
                    for c in range(int(self.ncon[a, l])):
                        out.append((a, l, m, c))
        return out

    @property
    def nbas(self) -> int:
        return int(sum(NCART[l] * int(self.ncon[a, l]) for a in range(self.natom) for l in range(3)))


@dataclasses.dataclass
class Mesh:
    xyz: np.ndarray   # (M, 3)
    w: np.ndarray     # (M,)

    @property
    def npts(self) -> int:
        return int(self.w.shape[0])


@dataclasses.dataclass
class System:
    """One FLO-SIC evaluation input (both spins)."""

    name: str
    basis: Basis
    mesh: Mesh
    coef: List[np.ndarray]   # per spin: (nbas, nwf) KS coefficients, columns orthonormal in S
    nocc: List[int]          # per spin: occupied orbitals == number of FODs
    fod: List[np.ndarray]    # per spin: (nocc, 3)
    xc: int = XC_LSDA
    nspn: int = 2
    meta: dict = dataclasses.field(default_factory=dict)

    @property
    def nbas(self) -> int:
        return self.basis.nbas


# ----------------------------------------------------------------------------------------------
# basis construction
# ----------------------------------------------------------------------------------------------

def build_basis(xyz: np.ndarray, Z: Sequence[int], rules=SHELL_RULES, ladder=EXPONENTS) -> Basis:
    """Uncontracted ladder basis: each (l, exponent) is its own contraction over the atom's bares."""
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    Z = np.asarray(Z, dtype=np.int32)
    bare_off = [0]
    alpha: list = []
    ncon = np.zeros((len(Z), 3), dtype=np.int32)
    con_off = np.zeros((len(Z), 3), dtype=np.int32)
    con: list = []
    for a, z in enumerate(Z):
        rule = rules[int(z)]
        used = sorted({k for l in rule for k in rule[l]})
        alpha.extend(float(ladder[k]) for k in used)
        nb = len(used)
        bare_off.append(bare_off[-1] + nb)
        for l in range(3):
            ks = rule.get(l, [])
            ncon[a, l] = len(ks)
            con_off[a, l] = len(con)
            blk = np.zeros((len(ks), nb))
            for c, k in enumerate(ks):
                i = used.index(k)
                blk[c, i] = prim_norm(float(ladder[k]), l)
            con.extend(blk.ravel().tolist())
    return Basis(xyz=xyz, Z=Z, bare_off=np.asarray(bare_off, dtype=np.int32),
                 alpha=np.asarray(alpha, dtype=np.float64), ncon=ncon, con_off=con_off,
                 con=np.asarray(con, dtype=np.float64))


def overlap_matrix(b: Basis) -> np.ndarray:
    """Analytic AO overlap (the OVMXSF integrals, ovmxsf.f:3-183, summed over contractions)."""
    N = b.nbas
    S = np.zeros((N, N))
    xyz = np.ascontiguousarray(b.xyz, np.float64)
    bo = np.ascontiguousarray(b.bare_off, np.int32)
    al = np.ascontiguousarray(b.alpha, np.float64)
    nc = np.ascontiguousarray(b.ncon, np.int32).ravel()
    co = np.ascontiguousarray(b.con_off, np.int32).ravel()
    cn = np.ascontiguousarray(b.con, np.float64)
    _setup_lib().setup_overlap(b.natom, _dp(xyz), _ip(bo), _dp(al), _ip(nc), _ip(co), _dp(cn), N, _dp(S))
    return 0.5 * (S + S.T)


def eval_basis(b: Basis, pts: np.ndarray, nderv: int = 1) -> np.ndarray:
    """Basis values (and gradients) at a few points: returns (1 or 4, npts, nbas). Set-up use only."""
    pts = np.asarray(pts, dtype=np.float64).reshape(-1, 3)
    ng = 1 if nderv == 0 else 4
    out = np.zeros((ng, pts.shape[0], b.nbas))
    k = 0
    for a in range(b.natom):
        r = pts - b.xyz[a]
        r2 = (r * r).sum(1)
        al = b.alpha[b.bare_off[a]:b.bare_off[a + 1]]
        arg = np.outer(r2, al)
        g = np.where(arg < CUTEXP, np.exp(-np.minimum(arg, CUTEXP)), 0.0)   # (npts, nbare)
        for l in range(3):
            nc = int(b.ncon[a, l])
            if nc == 0:
                continue
            cb = b.con_block(a, l)          # (nc, nbare)
            rad = g @ cb.T                  # (npts, nc)   sum_i c_i e^{-a_i r^2}
            drad = -2.0 * (g * al) @ cb.T   # (npts, nc)   sum_i c_i (-2 a_i) e^{-a_i r^2}
            for m, (px, py, pz) in enumerate(CART[l]):
                poly = r[:, 0] ** px * r[:, 1] ** py * r[:, 2] ** pz
                out[0, :, k:k + nc] = poly[:, None] * rad
                if ng == 4:
                    for dim, pw in enumerate((px, py, pz)):
                        dp = np.zeros_like(poly)
                        if pw > 0:
                            pows = [px, py, pz]
                            pows[dim] -= 1
                            dp = pw * r[:, 0] ** pows[0] * r[:, 1] ** pows[1] * r[:, 2] ** pows[2]
                        out[1 + dim, :, k:k + nc] = dp[:, None] * rad + (poly * r[:, dim])[:, None] * drad
                k += nc
    return out


# ----------------------------------------------------------------------------------------------
# mesh
# ----------------------------------------------------------------------------------------------

def fibonacci_sphere(n: int) -> np.ndarray:
    i = np.arange(n) + 0.5
    z = 1.0 - 2.0 * i / n
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = i * math.pi * (3.0 - math.sqrt(5.0))
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def random_rotation(rng: np.random.Generator) -> np.ndarray:
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def becke_mesh(xyz: np.ndarray, rng: np.random.Generator, nrad: int = 60, nang: int = 194,
               rmin: float = 5.0e-4, rmax: float = 25.0) -> Mesh:
    """Per-atom log radial grid x rotated Fibonacci sphere, Becke (1988) fuzzy-cell weights."""
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    K = xyz.shape[0]
    h = math.log(rmax / rmin) / (nrad - 1)
    rad = rmin * np.exp(h * np.arange(nrad))
    wrad = h * rad ** 3                     # int f r^2 dr = int f r^3 d(ln r)
    sph = fibonacci_sphere(nang)
    wang = 4.0 * math.pi / nang
    pts = []
    wts = []
    owner = []
    for a in range(K):
        R = random_rotation(rng)
        s = sph @ R.T
        p = xyz[a] + (rad[:, None, None] * s[None, :, :]).reshape(-1, 3)
        w = (wrad[:, None] * np.full((1, nang), wang)).reshape(-1)
        pts.append(p)
        wts.append(w)
        owner.append(np.full(p.shape[0], a, dtype=np.int64))
    pts = np.concatenate(pts)
    wts = np.concatenate(wts)
    owner = np.concatenate(owner)
    if K > 1:
        pts = np.ascontiguousarray(pts)
        owner32 = np.ascontiguousarray(owner, np.int32)
        _setup_lib().setup_becke(K, _dp(np.ascontiguousarray(xyz)), pts.shape[0], _dp(pts), _ip(owner32), _dp(wts))
    keep = wts > 0.0
    return Mesh(xyz=np.ascontiguousarray(pts[keep]), w=np.ascontiguousarray(wts[keep]))


# ----------------------------------------------------------------------------------------------
# orbitals and FODs
# ----------------------------------------------------------------------------------------------

def _sym_inv_sqrt(A: np.ndarray) -> np.ndarray:
    q, U = np.linalg.eigh(0.5 * (A + A.T))
    return (U / np.sqrt(q)) @ U.T


def random_orbitals(S: np.ndarray, nocc: int, nwf: int, rng: np.random.Generator) -> np.ndarray:
    """(nbas, nwf) coefficients, C^T S C = 1; first nocc columns are the occupied orbitals."""
    N = S.shape[0]
    X = rng.standard_normal((N, nocc))
    Cocc = X @ _sym_inv_sqrt(X.T @ S @ X)
    if nwf <= nocc:
        return np.ascontiguousarray(Cocc[:, :nwf])
    Y = rng.standard_normal((N, nwf - nocc))
    Y = Y - Cocc @ (Cocc.T @ (S @ Y))
    Y = Y - Cocc @ (Cocc.T @ (S @ Y))
    Cv = Y @ _sym_inv_sqrt(Y.T @ S @ Y)
    return np.ascontiguousarray(np.concatenate([Cocc, Cv], axis=1))


def lowdin_qmin(b: Basis, C: np.ndarray, fod: np.ndarray) -> float:
    """Smallest eigenvalue of the Fermi-orbital overlap for these FODs (frmorb2.F:296-336,460-503)."""
    n = fod.shape[0]
    chi = eval_basis(b, fod, nderv=0)[0]
    psi = chi @ C[:, :n]
    rho = (psi ** 2).sum(1)
    T = psi / np.sqrt(rho)[:, None]
    return float(np.linalg.eigvalsh(T @ T.T).min())


def place_fods(xyz: np.ndarray, Z: np.ndarray, nocc: int, rng: np.random.Generator) -> np.ndarray:
    heavy = [i for i, z in enumerate(Z) if z > 1]
    f = [xyz[i].copy() for i in heavy[:nocc]]
    while len(f) < nocc:
        a = int(rng.integers(len(Z)))
        while True:
            v = rng.uniform(-1.5, 1.5, 3)
            if v @ v <= 1.5 * 1.5:
                break
        f.append(xyz[a] + v)
    return np.asarray(f)


def random_atoms(K: int, rng: np.random.Generator, box: float = 10.0, dmin: float = 2.0):
    Z = rng.choice(np.array([1, 6, 8]), size=K, p=[0.5, 0.35, 0.15])
    xyz = []
    tries = 0
    while len(xyz) < K:
        tries += 1
        if tries > 200000:
            raise RuntimeError("atom placement failed")
        p = rng.uniform(0.0, box, 3)
        if all(np.linalg.norm(p - q) >= dmin for q in xyz):
            xyz.append(p)
    return np.asarray(xyz), Z.astype(np.int32)


def make_synthetic(K: int, nocc: int, seed: int, xc: int = XC_LSDA, nwf: Optional[int] = None,
                   nrad: int = 60, nang: int = 194, name: str = "synthetic", qmin_floor: float = 1e-3,
                   box: float = 10.0) -> System:
    """The generator behind the synthetic workloads (K atoms, Z in {1,6,8}, ladder basis, Becke mesh)."""
    ss = np.random.SeedSequence(seed)
    g_geo, g_mesh, g_c, g_fod = [np.random.default_rng(s) for s in ss.spawn(4)]
    xyz, Z = random_atoms(K, g_geo, box=box)
    b = build_basis(xyz, Z)
    mesh = becke_mesh(xyz, g_mesh, nrad=nrad, nang=nang)
    S = overlap_matrix(b)
    N = b.nbas
    nwf = N if nwf is None else min(nwf, N)
    coef, fods = [], []
    for spin in range(2):
        C = random_orbitals(S, nocc, nwf, g_c)
        for attempt in range(64):
            f = place_fods(xyz, Z, nocc, g_fod)
            if lowdin_qmin(b, C, f) > qmin_floor:
                break
        coef.append(C)
        fods.append(f)
    meta = dict(K=K, seed=seed, nbas=N, npts=mesh.npts, nocc=nocc, nwf=nwf, nrad=nrad, nang=nang,
                xc="PBE" if xc == XC_PBE else "LSDA")
    return System(name=name, basis=b, mesh=mesh, coef=coef, nocc=[nocc, nocc], fod=fods, xc=xc,
                  meta=meta)


def make_h2o(seed: int = 1, nrad: int = 60, nang: int = 194, sigma: float = 0.02) -> System:
    """H2O: O at origin, O-H 1.809 bohr, H-O-H 104.5 deg; 5 FODs per spin with seeded jitter."""
    th = math.radians(104.5 / 2.0)
    xyz = np.array([[0.0, 0.0, 0.0],
                    [1.809 * math.sin(th), 1.809 * math.cos(th), 0.0],
                    [-1.809 * math.sin(th), 1.809 * math.cos(th), 0.0]])
    Z = np.array([8, 1, 1], dtype=np.int32)
    rng = np.random.default_rng(seed)
    b = build_basis(xyz, Z)
    mesh = becke_mesh(xyz, rng, nrad=nrad, nang=nang)
    S = overlap_matrix(b)
    base = np.array([[0.0, 0.0, 0.0], 0.6 * xyz[1], 0.6 * xyz[2],
                     [0.0, -0.9 * math.cos(math.radians(54.74)), 0.9 * math.sin(math.radians(54.74))],
                     [0.0, -0.9 * math.cos(math.radians(54.74)), -0.9 * math.sin(math.radians(54.74))]])
    coef, fods = [], []
    for spin in range(2):
        coef.append(random_orbitals(S, 5, b.nbas, rng))
        fods.append(base + rng.normal(0.0, sigma, base.shape))
    meta = dict(K=3, seed=seed, nbas=b.nbas, npts=mesh.npts, nocc=5, nwf=b.nbas, nrad=nrad, nang=nang,
                xc="LSDA")
    return System(name="h2o", basis=b, mesh=mesh, coef=coef, nocc=[5, 5], fod=fods, xc=XC_LSDA, meta=meta)


def make_config(cfg: int, index: int = 0, **kw) -> System:
    """Workload configurations: 1 H2O, 2 K=16, 3 K=64, 4 K=32 PBE, 5 batch member ``index``."""
    if cfg == 1:
        return make_h2o(seed=1000 + index, **kw)
    if cfg == 2:
        return make_synthetic(16, 24, 1000 * 2 + index, XC_LSDA, name="K16", **kw)
    if cfg == 3:
        return make_synthetic(64, 96, 1000 * 3 + index, XC_LSDA, name="K64", **kw)
    if cfg == 4:
        return make_synthetic(32, 48, 1000 * 4 + index, XC_PBE, name="K32-PBE", **kw)
    if cfg == 5:
        rng = np.random.default_rng(5000 + index)
        K = int(rng.integers(4, 97))
        xc = XC_PBE if rng.random() < 0.5 else XC_LSDA
        nocc = max(2, min(128, (3 * K) // 2))
        return make_synthetic(K, nocc, 1000 * 5 + index, xc, name=f"batch{index}-K{K}", **kw)
    raise ValueError(f"unknown config {cfg}")


def subsample_mesh(sysm: System, npts: int) -> System:
    """The same system on every m-th mesh point (m = M // npts), weights times m: a bounded sample of a large
    workload on which the CPU oracle finishes in seconds."""
    M = sysm.mesh.npts
    step = max(1, M // max(1, npts))
    idx = np.arange(0, M, step)
    return dataclasses.replace(sysm, name=f"{sysm.name}-every{step}",
                               mesh=Mesh(xyz=np.ascontiguousarray(sysm.mesh.xyz[idx]),
                                         w=np.ascontiguousarray(sysm.mesh.w[idx] * step)),
                               meta=dict(sysm.meta, mesh_step=step))


def make_tiny(K: int = 3, nocc: int = 3, seed: int = 7, xc: int = XC_LSDA, nrad: int = 24, nang: int = 50,
              nwf: Optional[int] = None, box: float = 4.5) -> System:
    """Small system for parity tests (oracle finishes in seconds)."""
    return make_synthetic(K, nocc, seed, xc, nwf=nwf, nrad=nrad, nang=nang, name=f"tiny{K}", box=box)
