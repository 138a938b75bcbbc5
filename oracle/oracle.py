"""ctypes front-end of the CPU FP64 oracle (flosic_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the CPU-baseline leg of bench.py,
always as the checker.  The product path never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class ObBasis(C.Structure):
    _fields_ = [("natom", C.c_int32), ("xyz", _dp), ("bare_off", _ip), ("alpha", _dp),
                ("ncon", _ip), ("con_off", _ip), ("con", _dp)]


class ObSpin(C.Structure):
    _fields_ = [("nwf", C.c_int32), ("nocc", C.c_int32), ("coef", _dp), ("fod", _dp),
                ("fmat", _dp), ("over", _dp), ("q", _dp), ("U", _dp), ("tmat", _dp),
                ("results", _dp), ("sic", _dp), ("vab", _dp), ("debdax", _dp), ("grad", _dp),
                ("vpot", _dp), ("vh", _dp)]


def build(force: bool = False) -> str:
    src = os.path.join(HERE, "flosic_oracle.c")
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.oracle_eval.restype = C.c_int
        L.oracle_eval.argtypes = [C.POINTER(ObBasis), C.c_int64, _dp, _dp, C.c_int, C.c_int, C.POINTER(ObSpin),
                                  _dp, C.c_int]
        L.oracle_boys.argtypes = [C.c_int, C.c_int, _dp, _dp]
        L.oracle_basis_eval.argtypes = [C.POINTER(ObBasis), C.c_int, C.c_int, _dp, _dp]
        L.oracle_pairs.restype = C.c_int
        L.oracle_pairs.argtypes = [C.POINTER(ObBasis), C.c_int, _ip]
        L.oracle_poisson_batch.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_int, _dp, _dp]
        L.oracle_xc_points.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp]
        L.oracle_eigh.argtypes = [C.c_int, _dp, _dp, _dp]
        L.oracle_fod_forces.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int, _dp, _dp]
        L.oracle_xc_kernel.argtypes = [C.c_int, _dp, _dp]
        L.oracle_stage_times.argtypes = [_dp]
        L.oracle_flo_coef.restype = C.c_int
        L.oracle_flo_coef.argtypes = [C.POINTER(ObBasis), C.POINTER(ObSpin), _dp]
        L.oracle_lagmult.argtypes = [C.POINTER(ObBasis), C.c_int64, _dp, _dp, C.c_int, _dp, _dp, _dp]
        L.oracle_sic_hamiltonian.argtypes = [C.POINTER(ObBasis), C.c_int64, _dp, _dp, C.c_int, _dp, _dp, _dp]
        L.oracle_siclag_der.argtypes = [C.POINTER(ObBasis), C.c_int, C.c_int64, _dp, _dp, C.c_int, _dp, _dp, _dp]
        L.oracle_pbe_edens.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.oracle_dft_lsda.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        L.oracle_dft_gga.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        L.oracle_nbas.restype = C.c_int
        L.oracle_nbas.argtypes = [C.POINTER(ObBasis)]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


class _BasisHold:
    def __init__(self, b):
        self.xyz = np.ascontiguousarray(b.xyz, dtype=np.float64)
        self.bare_off = np.ascontiguousarray(b.bare_off, dtype=np.int32)
        self.alpha = np.ascontiguousarray(b.alpha, dtype=np.float64)
        self.ncon = np.ascontiguousarray(b.ncon, dtype=np.int32).ravel()
        self.con_off = np.ascontiguousarray(b.con_off, dtype=np.int32).ravel()
        self.con = np.ascontiguousarray(b.con, dtype=np.float64)
        self.s = ObBasis(b.natom, _d(self.xyz), _i(self.bare_off), _d(self.alpha), _i(self.ncon),
                         _i(self.con_off), _d(self.con))


def boys(top: int, T) -> np.ndarray:
    T = np.ascontiguousarray(T, dtype=np.float64).ravel()
    out = np.zeros((T.size, top + 1))
    lib().oracle_boys(top, T.size, _d(T), _d(out))
    return out


def basis_eval(basis, pts, nderv: int = 0) -> np.ndarray:
    h = _BasisHold(basis)
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    ng = {0: 1, 1: 4, 2: 10}[nderv]
    out = np.zeros((ng, pts.shape[0], basis.nbas))
    lib().oracle_basis_eval(C.byref(h.s), nderv, pts.shape[0], _d(pts), _d(out))
    return out


def pairs(basis) -> np.ndarray:
    h = _BasisHold(basis)
    n = lib().oracle_pairs(C.byref(h.s), 0, _i(np.zeros(6, np.int32)))
    out = np.zeros((max(n, 1), 6), np.int32)
    lib().oracle_pairs(C.byref(h.s), n, _i(out))
    return out[:n]


def poisson_batch(alpha, A, beta, B, rho, pts) -> np.ndarray:
    """Potential of pairs with Fortran-ordered RHO(10,10,npair) (element (mu,nu) = rho[ip, nu, mu])."""
    alpha = np.ascontiguousarray(alpha, np.float64)
    beta = np.ascontiguousarray(beta, np.float64)
    A = np.ascontiguousarray(A, np.float64)
    B = np.ascontiguousarray(B, np.float64)
    rho = np.ascontiguousarray(rho, np.float64)
    pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
    V = np.zeros(pts.shape[0])
    lib().oracle_poisson_batch(alpha.size, _d(alpha), _d(A), _d(beta), _d(B), _d(rho), pts.shape[0], _d(pts), _d(V))
    return V


def xc_points(xc: int, rhog) -> tuple:
    rhog = np.ascontiguousarray(rhog, np.float64).reshape(-1, 10)
    n = rhog.shape[0]
    v = np.zeros(n)
    e = np.zeros((n, 4))
    lib().oracle_xc_points(xc, n, _d(rhog), _d(v), _d(e))
    return v, e


def pbe_edens(rho, g) -> tuple:
    """PBE energy density e(rho, |grad rho|) of a fully polarised density and its exact partial derivatives."""
    rho = np.ascontiguousarray(rho, np.float64).ravel()
    g = np.ascontiguousarray(g, np.float64).ravel()
    e, er, eg = np.zeros(rho.size), np.zeros(rho.size), np.zeros(rho.size)
    lib().oracle_pbe_edens(rho.size, _d(rho), _d(g), _d(e), _d(er), _d(eg))
    return e, er, eg


def xc_kernel(which: int, args) -> np.ndarray:
    """0 PW91EX(dkf,s,u,v) 1 PW91LC(rs,zet) 2 PBEEX(rho,s,u,v) 3 PBECOR(rs,zet,t,uu,vv,ww)."""
    a = np.ascontiguousarray(args, np.float64)
    out = np.zeros(6)
    lib().oracle_xc_kernel(which, _d(a), _d(out))
    return out


def dft_lsda(rho_up, rho_dn, w, vh):
    """DFT-side LSDA potential of the total spin densities plus the Hartree potential vh of the total density
    (APOTNL_DFT with NSPN = 2): returns pot (2, n) and {E_x, E_c, E_H, N_up, N_dn}."""
    ru, rd, ww, v = (np.ascontiguousarray(a, np.float64) for a in (rho_up, rho_dn, w, vh))
    n = ru.size
    pot = np.zeros(2 * n)
    e = np.zeros(5)
    lib().oracle_dft_lsda(n, _d(ru), _d(rd), _d(ww), _d(v), _d(pot), _d(e))
    return pot.reshape(2, n), dict(zip(("ex", "ec", "eh", "n_up", "n_dn"), e))


def dft_gga(rg_up, rg_dn, w, vh):
    """DFT-side PBE potential (SUBVLXC with NSPN = 2, IGGA = 1) of spin densities given as (n, 10) arrays (value,
    gradient, Hessian xx,yy,zz,xy,xz,yz) plus the Hartree potential vh of the total density: pot (2, n) and
    {E_x, E_c, E_H, N_up, N_dn}."""
    ru, rd = (np.ascontiguousarray(a, np.float64).reshape(-1, 10) for a in (rg_up, rg_dn))
    ww, v = (np.ascontiguousarray(a, np.float64) for a in (w, vh))
    n = ru.shape[0]
    pot = np.zeros(2 * n)
    e = np.zeros(5)
    lib().oracle_dft_gga(n, _d(ru), _d(rd), _d(ww), _d(v), _d(pot), _d(e))
    return pot.reshape(2, n), dict(zip(("ex", "ec", "eh", "n_up", "n_dn"), e))


def eigh(A):
    A = np.ascontiguousarray(A, np.float64)
    n = A.shape[0]
    q = np.zeros(n)
    U = np.zeros((n, n))
    lib().oracle_eigh(n, _d(A), _d(q), _d(U))
    return q, U.reshape(n, n).T  # column-major -> U[:, k] eigenvector k


def fod_forces(fmat, U, q, sinvh, tmat, sic, ldsic):
    n = q.size
    deb = np.zeros(3 * n)
    grad = np.zeros(3 * n)
    f = np.ascontiguousarray(fmat)  # [K][i][j] memory == FMAT(j,i,K)
    lib().oracle_fod_forces(n, _d(f), _d(np.asfortranarray(U)), _d(np.ascontiguousarray(q)),
                            _d(np.asfortranarray(sinvh)), _d(np.asfortranarray(tmat)),
                            _d(np.asfortranarray(sic)), ldsic, _d(deb), _d(grad))
    return deb.reshape(n, 3), grad.reshape(n, 3)


def evaluate(sysm, nthreads: int = 0, want_potentials: bool = False, want_vab: bool = True,
             fods=None, coefs=None) -> dict:
    """Full FP64 evaluation of a System: returns per-spin dicts and E_SIC."""
    h = _BasisHold(sysm.basis)
    pts = np.ascontiguousarray(sysm.mesh.xyz, np.float64)
    w = np.ascontiguousarray(sysm.mesh.w, np.float64)
    M = w.size
    spins = (ObSpin * 2)()
    keep = []
    outs = []
    for s in range(2):
        C_ = np.asfortranarray((coefs or sysm.coef)[s], np.float64)
        fod = np.ascontiguousarray((fods or sysm.fod)[s], np.float64)
        n = fod.shape[0]
        nwf = C_.shape[1]
        o = dict(fmat=np.zeros((4, n, n)), over=np.zeros((n, n)), q=np.zeros(n), U=np.zeros((n, n)),
                 tmat=np.zeros((n, n)), results=np.zeros((n, 13)), sic=np.zeros((n, nwf)),
                 vab=np.zeros((n, n, n)) if want_vab else None, debdax=np.zeros((n, 3)), grad=np.zeros((n, 3)),
                 vpot=np.zeros((n, M)) if want_potentials else None, vh=np.zeros((n, M)) if want_potentials else None)
        keep += [C_, fod]
        sp = spins[s]
        sp.nwf = nwf
        sp.nocc = n
        sp.coef = _d(C_)
        sp.fod = _d(fod)
        for k in ("fmat", "over", "q", "U", "tmat", "results", "sic", "vab", "debdax", "grad", "vpot", "vh"):
            setattr(sp, k, _d(o[k]) if o[k] is not None else _dp())
        outs.append(o)
    esic = np.zeros(1)
    st = lib().oracle_eval(C.byref(h.s), M, _d(pts), _d(w), int(sysm.nspn), int(sysm.xc), spins, _d(esic),
                           int(nthreads))
    st_t = np.zeros(2)
    lib().oracle_stage_times(_d(st_t))
    res = dict(status=int(st), esic=float(esic[0]), spins=[], stage_s=dict(fixed=float(st_t[0]), mesh=float(st_t[1])))
    for s in range(2):
        o = outs[s]
        n = o["q"].size
        # raw buffers are column-major (Fortran); present them as natural numpy arrays
        d = dict(
            fmat=o["fmat"],                               # fmat[K, i, j] = FMAT(j,i,K)
            over=o["over"].T.copy(),
            q=o["q"],
            U=o["U"].T.copy(),                            # U[:, a] eigenvector a
            tmat=o["tmat"].T.copy(),                      # tmat[i, j] = TMAT(i,j)
            results=o["results"],                         # results[i, 0:13]
            sic=o["sic"].T.copy(),                        # sic[j, i] = SIC(j,i)
            vab=None if o["vab"] is None else np.transpose(o["vab"], (0, 2, 1)).copy(),  # vab[i, a, b]
            debdax=o["debdax"], grad=o["grad"], vpot=o["vpot"], vh=o["vh"])
        res["spins"].append(d)
    return res


def flo_coefs(sysm, ispn: int, fods=None) -> np.ndarray:
    """FLO AO coefficients (nbas x nocc) of spin ispn (1 or 2)."""
    h = _BasisHold(sysm.basis)
    C_ = np.asfortranarray(sysm.coef[ispn - 1], np.float64)
    fod = np.ascontiguousarray((fods or sysm.fod)[ispn - 1], np.float64)
    sp = ObSpin()
    sp.nwf = C_.shape[1]
    sp.nocc = fod.shape[0]
    sp.coef = _d(C_)
    sp.fod = _d(fod)
    out = np.zeros((fod.shape[0], sysm.basis.nbas))  # row f = FLO f (column-major nbas x n)
    st = lib().oracle_flo_coef(C.byref(h.s), C.byref(sp), _d(out))
    if st:
        raise RuntimeError(f"oracle_flo_coef status {st}")
    return out.T.copy()


def lagmult(sysm, ispn: int, pot) -> np.ndarray:
    """LAGMULT_SLV: [j, k] = <phi_j| -1/2 lap + pot |phi_k> over the FLOs of spin ispn, on the mesh."""
    cf = np.asfortranarray(flo_coefs(sysm, ispn))
    h = _BasisHold(sysm.basis)
    pts = np.ascontiguousarray(sysm.mesh.xyz, np.float64)
    w = np.ascontiguousarray(sysm.mesh.w, np.float64)
    n = cf.shape[1]
    lam = np.zeros(n * n)
    lib().oracle_lagmult(C.byref(h.s), w.size, _d(pts), _d(w), n, _d(cf), _d(np.ascontiguousarray(pot, np.float64)),
                         _d(lam))
    return lam.reshape(n, n).T.copy()


def sic_hamiltonian(sysm, ispn: int, vpot) -> np.ndarray:
    """APATCH (MM=2): H_ab = sum_i 1/2 (<chi_a|V_i|phi_i><phi_i|chi_b> + (a <-> b)), both brackets on the mesh;
    vpot[i] = V_i on the mesh (n x npts)."""
    cf = np.asfortranarray(flo_coefs(sysm, ispn))
    h = _BasisHold(sysm.basis)
    pts = np.ascontiguousarray(sysm.mesh.xyz, np.float64)
    w = np.ascontiguousarray(sysm.mesh.w, np.float64)
    n = cf.shape[1]
    N = sysm.basis.nbas
    H = np.zeros(N * N)
    lib().oracle_sic_hamiltonian(C.byref(h.s), w.size, _d(pts), _d(w), n, _d(cf),
                                 _d(np.ascontiguousarray(vpot, np.float64)), _d(H))
    return H.reshape(N, N).T.copy()


def siclag_der(sysm, vpot, cflo=None, basis=None) -> np.ndarray:
    """SICLAG_DER: dersic[A, i, x] = sum_k w dphi_i^A/dR_A,x V_i phi_i over the FLOs of both spins (spin 1 first);
    vpot[i] = V_i on the mesh (nF x npts); cflo (nbas x nF) defaults to the FLOs of both spins, basis to the
    system's (a basis with moved centres gives the finite-difference checks their fields)."""
    if cflo is None:
        cflo = np.hstack([flo_coefs(sysm, 1), flo_coefs(sysm, 2)])
    cflo = np.asfortranarray(cflo, np.float64)
    b = sysm.basis if basis is None else basis
    h = _BasisHold(b)
    pts = np.ascontiguousarray(sysm.mesh.xyz, np.float64)
    w = np.ascontiguousarray(sysm.mesh.w, np.float64)
    n = cflo.shape[1]
    out = np.zeros(3 * n * b.natom)
    lib().oracle_siclag_der(C.byref(h.s), int(sysm.nspn), w.size, _d(pts), _d(w), n, _d(cflo),
                            _d(np.ascontiguousarray(vpot, np.float64)), _d(out))
    return out.reshape(b.natom, n, 3)


def flo_fields(sysm, ispn: int):
    """FLO values on the mesh, [f, k]."""
    cf = flo_coefs(sysm, ispn)
    chi = basis_eval(sysm.basis, sysm.mesh.xyz, 0)[0]
    return (chi @ cf).T.copy()


def diagv_fo(ham, n: int, maxiter: int = 600, convg: float = 1.0e-6):
    """DIAGV_FO (diagv_fo.F:5-156) restated with numpy, loop for loop: the symmetric part of the m x m orbital
    Hamiltonian ham[J, I] = HAM(J,I) is diagonalised (DIAGGE with a unit overlap, diagv_fo.F:48-71; here this
    oracle's eigensolver, so eigenvector signs may differ from LAPACK's), its eigenvectors by ascending eigenvalue
    start the orbitals; each sweep rotates every (occupied i < n, virtual j >= n) pair by
    t = atan(2 h_ij / (h_ii - h_jj)) / 2 with h_ii, h_ij, h_jj from two matrix-vector products per pair
    (diagv_fo.F:74-112), and stops once max |phi_occ^T HAM phi_virt| < convg (diagv_fo.F:117-140).
    Returns (phires, sweep): phires[inew, iold] = PHIRES(INEW,IOLD) after the final transpose, sweep the converged
    sweep or 0 after maxiter sweeps."""
    H = np.asarray(ham, np.float64)
    m = H.shape[0]
    _, P = eigh(0.5 * (H + H.T))
    P = P.copy()
    with np.errstate(divide="ignore"):
        for sweep in range(1, maxiter + 1):
            for i in range(n):
                for j in range(n, m):
                    f = H @ P[:, i]
                    hii = P[:, i] @ f
                    f = H @ P[:, j]
                    hij = P[:, i] @ f
                    hjj = P[:, j] @ f
                    t = 0.5 * np.arctan(2.0 * hij / (hii - hjj))
                    pi, pj = P[:, i].copy(), P[:, j].copy()
                    P[:, i] = np.cos(t) * pi + np.sin(t) * pj
                    P[:, j] = -np.sin(t) * pi + np.cos(t) * pj
            if n == 0 or n == m or np.abs(P[:, :n].T @ H @ P[:, n:]).max() < convg:
                return P.T.copy(), sweep
    return P.T.copy(), 0


def scf_sic_orbitals(ham_ao, coef_flo_virt, sic, n: int):
    """SCFSICW's FLO-SIC orbital update (scfsicw.f:353-614, NITER > 1, no unified Hamiltonian) restated:
    coef_flo_virt = C (nbas x m: the FLOs C_occ TMAT^T, then the virtual orbitals), sic[I, J] = SIC(I,J) for the
    FLOs J < n.  HAM = C^T ham_ao C; HAM(J,I) += SIC(I,J); DIAGV_FO; EVAL(i) = phi_i^T HAM phi_i; C_new = C PHIRES^T.
    Returns (C_new, EVAL, sweep)."""
    Cm = np.asarray(coef_flo_virt, np.float64)
    ham = Cm.T @ np.asarray(ham_ao, np.float64) @ Cm
    ham[:n, :] += np.asarray(sic, np.float64)[:, :n].T
    phi, sweep = diagv_fo(ham, n)
    ev = np.einsum("ik,lk,il->i", phi, ham, phi)
    return Cm @ phi.T, ev, sweep
