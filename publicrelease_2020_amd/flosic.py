"""Python host mirror of the FLO-SIC hot path over the C ABI of libflosic_hip.so.

The methods carry the names of the reference subroutines they replace (FLOSIC/PublicRelease_2020):
  frmorb()         FRMORB2(0,0) + FRMIORB      Fermi orbitals at the FODs, Lowdin, TMAT   frmorb2.F:140-525
  apotnl_sic()     APOTNL_SIC over all FLOs    RESULTS(13,:) per spin, E_SIC               apotnl_sic.F:31-668
  siclagm()        SICLAGM                     SIC(j,i), j over all orbitals               frmorb_mpi.F:2-178
  sic_vab()        AHAMMAT in FLO space        V^i_ab                                       sicmat.F:380-411
  fod_forces()     FRMORB2(-ISPN,-1)           DEBDAX = +dE_SIC/da                          frmorb2.F:589-729
  fod_gradient()   adjoint dE_SIC/da           (no reference entry point; equals DEBDAX for LSDA, the exact
                                               derivative of the discrete energy for PBE)
  potential()      ZPOTnnnn                    V_i = V_xc + V_H on the mesh                 apotnl_sic.F:497-522
Errors are raised as FlosicError carrying the library status, the way the reference stops (STOPIT).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libflosic_hip.so")

OK, EINVAL, ELOWDIN, ECOUPOT, EDEVICE, ESTATE, ENOMEM = range(7)
XC_LSDA, XC_PBE = 0, 1
PREC_FP64, PREC_MIXED, PREC_BF16X3, PREC_BF16X2 = 0, 1, 2, 3

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_lp = C.POINTER(C.c_int64)


class FlosicError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"flosic status {status}: {msg}")
        self.status = status


_lib = None


def library():
    """Load the HIP library; raises if it has not been built (there is no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FlosicError(ESTATE, f"{LIB_PATH} is missing: run publicrelease_2020_amd/build.py")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        sig = {
            "flosic_create": [C.c_int32, C.POINTER(vp)],
            "flosic_destroy": [vp],
            "flosic_set_basis": [vp, C.c_int32, _dp, _ip, _dp, _ip, _ip, _dp],
            "flosic_set_mesh": [vp, C.c_int64, _dp, _dp],
            "flosic_set_options": [vp, C.c_int32, C.c_int32, C.c_int32],
            "flosic_set_orbitals": [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _dp],
            "flosic_set_fods": [vp, C.c_int32, C.c_int32, _dp],
            "flosic_frmorb": [vp],
            "flosic_get_fo": [vp, C.c_int32, _dp, _dp, _dp, _dp, _dp],
            "flosic_evaluate": [vp, _dp],
            "flosic_apotnl_sic": [vp, C.c_int32, _dp, _dp],
            "flosic_apotnl_dft": [vp, _dp, _dp],
            "flosic_apotnl_dft_noalloc": [vp, _dp, _dp],
            "flosic_siclag_der": [vp, _dp],
            "flosic_siclagm": [vp, C.c_int32, _dp],
            "flosic_sic_vab": [vp, C.c_int32, _dp],
            "flosic_fod_forces": [vp, C.c_int32, _dp],
            "flosic_fod_gradient": [vp, C.c_int32, _dp],
            "flosic_get_potential": [vp, C.c_int32, C.c_int32, _dp, _dp],
            "flosic_lagmult_slv": [vp, C.c_int32, _dp, C.c_int32, _dp],
            "flosic_sic_hamiltonian": [vp, C.c_int32, _dp],
            "flosic_get_pairs": [vp, _ip, _ip],
            "flosic_get_timings": [vp, C.c_int32, _dp],
            "flosic_get_counts": [vp, C.c_int32, _lp],
            "flosic_set_mesh_block": [vp, C.c_int64],
            "flosic_set_coulomb_cols": [vp, C.c_int32],
            "flosic_set_far_threshold": [vp, C.c_double],
            "flosic_diagv_fo": [C.c_int32, C.c_int32, _dp, C.c_int32, _dp, _ip],
            "flosic_scf_sic_orbitals": [vp, C.c_int32, _dp, C.c_int32, _dp, _dp, _ip],
            "flosic_batch_create": [C.c_int32, C.POINTER(vp), C.c_int32, C.POINTER(vp)],
            "flosic_batch_evaluate": [vp, _dp],
            "flosic_batch_get_timings": [vp, C.c_int32, _dp],
            "flosic_batch_get_counts": [vp, C.c_int32, _lp],
            "flosic_batch_destroy": [vp],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int32
        L.flosic_last_error.argtypes = [vp]
        L.flosic_last_error.restype = C.c_char_p
        L.flosic_batch_last_error.argtypes = [vp]
        L.flosic_batch_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def exported_symbols() -> list:
    return ["flosic_create", "flosic_destroy", "flosic_last_error", "flosic_set_basis", "flosic_set_mesh",
            "flosic_set_options", "flosic_set_orbitals", "flosic_set_fods", "flosic_frmorb", "flosic_get_fo",
            "flosic_evaluate", "flosic_apotnl_sic", "flosic_apotnl_dft", "flosic_apotnl_dft_noalloc", "flosic_siclag_der", "flosic_siclagm", "flosic_sic_vab", "flosic_fod_forces",
            "flosic_fod_gradient", "flosic_get_potential", "flosic_lagmult_slv", "flosic_sic_hamiltonian",
            "flosic_get_pairs", "flosic_get_timings", "flosic_get_counts", "flosic_set_mesh_block", "flosic_set_coulomb_cols",
            "flosic_set_far_threshold", "flosic_diagv_fo", "flosic_scf_sic_orbitals",
            "flosic_batch_create", "flosic_batch_evaluate", "flosic_batch_last_error", "flosic_batch_get_timings",
            "flosic_batch_get_counts", "flosic_batch_destroy"]


def _d(a):
    return a.ctypes.data_as(_dp)


def diagv_fo(ham, n: int):
    """DIAGV_FO (flosic_diagv_fo) of the m x m orbital Hamiltonian ham[J, I] = HAM(J,I) with n occupied orbitals:
    (PHIRES with phires[inew, iold], converged sweep or 0)."""
    L = library()
    h = np.asfortranarray(np.asarray(ham, np.float64))
    m = h.shape[0]
    phi = np.zeros((m, m), order="F")
    sw = C.c_int32()
    st = L.flosic_diagv_fo(m, int(n), _d(h), m, _d(phi), C.byref(sw))
    if st:
        raise FlosicError(st, "flosic_diagv_fo")
    return np.ascontiguousarray(phi), sw.value


class Flosic:
    """One evaluation context on a HIP device, holding basis, mesh, orbitals and FODs."""

    def __init__(self, device: int = 0):
        self._L = library()
        h = C.c_void_p()
        st = self._L.flosic_create(int(device), C.byref(h))
        if st != OK:
            raise FlosicError(st, "flosic_create failed (no HIP device?)")
        self._h = h
        self._n = [0, 0]
        self._nwf = [0, 0]
        self.nbas = 0
        self.npts = 0

    def close(self):
        if getattr(self, "_h", None):
            self._L.flosic_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, st: int):
        if st != OK:
            raise FlosicError(st, self._L.flosic_last_error(self._h).decode())

    # ---- set-up (module data of the reference) ----
    def set_basis(self, basis):
        self._keep_basis = (np.ascontiguousarray(basis.xyz, np.float64), np.ascontiguousarray(basis.bare_off, np.int32),
                            np.ascontiguousarray(basis.alpha, np.float64),
                            np.ascontiguousarray(basis.ncon, np.int32).ravel(),
                            np.ascontiguousarray(basis.con_off, np.int32).ravel(),
                            np.ascontiguousarray(basis.con, np.float64))
        x, bo, al, nc, co, cn = self._keep_basis
        self._chk(self._L.flosic_set_basis(self._h, basis.natom, _d(x), bo.ctypes.data_as(_ip), _d(al),
                                           nc.ctypes.data_as(_ip), co.ctypes.data_as(_ip), _d(cn)))
        self.nbas = basis.nbas
        self.natom = basis.natom

    def set_mesh(self, xyz, w):
        xyz = np.ascontiguousarray(xyz, np.float64)
        w = np.ascontiguousarray(w, np.float64)
        self._chk(self._L.flosic_set_mesh(self._h, w.size, _d(xyz), _d(w)))
        self.npts = w.size

    def set_options(self, nspn: int = 2, xc: int = XC_LSDA, precision: int = PREC_FP64):
        self._chk(self._L.flosic_set_options(self._h, nspn, xc, precision))
        if nspn == 1:  # spin-restricted: spin 2 is empty
            self._n[1] = self._nwf[1] = 0
        self._nspn = nspn

    def set_orbitals(self, ispn: int, coef, nocc: int):
        c = np.asfortranarray(coef, np.float64)
        self._chk(self._L.flosic_set_orbitals(self._h, ispn, c.shape[0], c.shape[1], nocc, _d(c)))
        self._n[ispn - 1] = nocc
        self._nwf[ispn - 1] = c.shape[1]

    def set_fods(self, ispn: int, fod):
        f = np.ascontiguousarray(fod, np.float64).reshape(-1, 3)
        if ispn == 2 and getattr(self, "_nspn", 2) == 1 and f.shape[0] == 0:
            return  # NSPN = 1 has no second spin channel
        self._chk(self._L.flosic_set_fods(self._h, ispn, f.shape[0], _d(f)))

    def load_system(self, sysm, precision: int = PREC_FP64):
        self.set_basis(sysm.basis)
        self.set_mesh(sysm.mesh.xyz, sysm.mesh.w)
        self.set_options(sysm.nspn, sysm.xc, precision)
        for s in range(1, sysm.nspn + 1):
            self.set_orbitals(s, sysm.coef[s - 1], sysm.nocc[s - 1])
            self.set_fods(s, sysm.fod[s - 1])

    # ---- hot path ----
    def frmorb(self):
        self._chk(self._L.flosic_frmorb(self._h))

    def fermi_orbitals(self, ispn: int) -> dict:
        n = self._n[ispn - 1]
        fm, ov, q, U, tm = (np.zeros(4 * n * n), np.zeros(n * n), np.zeros(n), np.zeros(n * n), np.zeros(n * n))
        self._chk(self._L.flosic_get_fo(self._h, ispn, _d(fm), _d(ov), _d(q), _d(U), _d(tm)))
        return dict(fmat=fm.reshape(4, n, n), over=ov.reshape(n, n).T, q=q, U=U.reshape(n, n).T,
                    tmat=tm.reshape(n, n).T)

    def lowdin_q(self, ispn: int) -> np.ndarray:
        """Eigenvalues of the Fermi-orbital overlap of spin ispn (EVAL before inversion, lowden_files.f:71-99)."""
        return self.fermi_orbitals(ispn)["q"]

    def evaluate(self) -> float:
        z = np.zeros(1)
        self._chk(self._L.flosic_evaluate(self._h, _d(z)))
        return float(z[0])

    def apotnl_sic(self, ispn: int) -> np.ndarray:
        n = self._n[ispn - 1]
        r = np.zeros(13 * n)
        self._chk(self._L.flosic_apotnl_sic(self._h, ispn, _d(r), None))
        return r.reshape(n, 13)

    def apotnl_dft(self):
        """DFT-side LSDA potential of the total spin densities (APOTNL_DFT): pot (2, nmsh) = V_H + V_xc per spin and
        {ex, ec, eh, n_up, n_dn}."""
        M = int(self.counts()[2])
        pot = np.zeros(2 * M)
        e = np.zeros(5)
        self._chk(self._L.flosic_apotnl_dft(self._h, _d(pot), _d(e)))
        return pot.reshape(2, M), dict(zip(("ex", "ec", "eh", "n_up", "n_dn"), (float(x) for x in e)))

    def apotnl_dft_noalloc(self):
        """flosic_apotnl_dft_noalloc: the same outputs as apotnl_dft from the library's fixed scratch (no device
        allocation; needs a current evaluation with resident fields)."""
        M = int(self.counts()[2])
        pot = np.zeros(2 * M)
        e = np.zeros(5)
        self._chk(self._L.flosic_apotnl_dft_noalloc(self._h, _d(pot), _d(e)))
        return pot.reshape(2, M), dict(zip(("ex", "ec", "eh", "n_up", "n_dn"), (float(x) for x in e)))

    def siclag_der(self) -> np.ndarray:
        """SICLAG_DER: [A, i, x] = sum_k w dphi_i^A/dR_A,x V_i phi_i, FLOs of both spins (spin 1 first)."""
        nF = self._n[0] + self._n[1]
        d = np.zeros(3 * nF * self.natom)
        self._chk(self._L.flosic_siclag_der(self._h, _d(d)))
        return d.reshape(self.natom, nF, 3)

    def siclagm(self, ispn: int) -> np.ndarray:
        n, nwf = self._n[ispn - 1], self._nwf[ispn - 1]
        s = np.zeros(nwf * n)
        self._chk(self._L.flosic_siclagm(self._h, ispn, _d(s)))
        return s.reshape(n, nwf).T  # [j, i] = SIC(j,i)

    def sic_vab(self, ispn: int) -> np.ndarray:
        n = self._n[ispn - 1]
        v = np.zeros(n ** 3)
        self._chk(self._L.flosic_sic_vab(self._h, ispn, _d(v)))
        return np.transpose(v.reshape(n, n, n), (0, 2, 1)).copy()  # [i, a, b]

    def fod_forces(self, ispn: int) -> np.ndarray:
        n = self._n[ispn - 1]
        d = np.zeros(3 * n)
        self._chk(self._L.flosic_fod_forces(self._h, ispn, _d(d)))
        return d.reshape(n, 3)

    def fod_gradient(self, ispn: int) -> np.ndarray:
        n = self._n[ispn - 1]
        d = np.zeros(3 * n)
        self._chk(self._L.flosic_fod_gradient(self._h, ispn, _d(d)))
        return d.reshape(n, 3)

    def potential(self, ispn: int, iorb: int):
        p = np.zeros(self.npts)
        vh = np.zeros(self.npts)
        self._chk(self._L.flosic_get_potential(self._h, ispn, iorb, _d(p), _d(vh)))
        return p, vh

    def lagmult_slv(self, ispn: int, pot, sic=None) -> np.ndarray:
        """LAGMULT_SLV: sic[j, k] += <phi_j| -1/2 lap + pot |phi_k> over the FLOs of spin ispn (sic defaults to
        zeros, so the DFT part alone is returned)."""
        n = self._n[ispn - 1]
        out = np.zeros((n, n)) if sic is None else np.array(sic, dtype=np.float64)
        f = np.asfortranarray(out)
        p = np.ascontiguousarray(pot, np.float64)
        self._chk(self._L.flosic_lagmult_slv(self._h, ispn, _d(p), f.shape[0], _d(f)))
        return np.ascontiguousarray(f)

    def sic_hamiltonian(self, ispn: int) -> np.ndarray:
        """APATCH: H_ab = sum_i 1/2 (<chi_a|V_i|phi_i><phi_i|chi_b> + (a <-> b)) over the FLOs of spin ispn."""
        h = np.zeros(self.nbas * self.nbas)
        self._chk(self._L.flosic_sic_hamiltonian(self._h, ispn, _d(h)))
        return h.reshape(self.nbas, self.nbas).T.copy()

    def scf_sic_orbitals(self, ispn: int, ham_ao):
        """SCFSICW's FLO-SIC orbital update (flosic_scf_sic_orbitals): new orbitals (nbas x nwf), EVAL (nwf) and the
        DIAGV_FO sweep count from the spin's Kohn-Sham Hamiltonian in the basis (nbas x nbas)."""
        nwf = self._nwf[ispn - 1]
        h = np.asfortranarray(np.asarray(ham_ao, np.float64))
        assert h.shape == (self.nbas, self.nbas)
        coef = np.zeros((self.nbas, nwf), order="F")
        ev = np.zeros(nwf)
        sw = C.c_int32()
        self._chk(self._L.flosic_scf_sic_orbitals(self._h, ispn, _d(h), self.nbas, _d(coef), _d(ev), C.byref(sw)))
        return np.ascontiguousarray(coef), ev, sw.value

    def pairs(self) -> np.ndarray:
        n = C.c_int32()
        self._chk(self._L.flosic_get_pairs(self._h, C.byref(n), None))
        out = np.zeros((max(n.value, 1), 6), np.int32)
        self._chk(self._L.flosic_get_pairs(self._h, C.byref(n), out.ctypes.data_as(_ip)))
        return out[:n.value]

    def timings(self) -> np.ndarray:
        t = np.zeros(10)
        self._chk(self._L.flosic_get_timings(self._h, 10, _d(t)))
        return t

    def counts(self) -> np.ndarray:
        c = np.zeros(10, np.int64)
        self._chk(self._L.flosic_get_counts(self._h, 10, c.ctypes.data_as(_lp)))
        return c

    def set_far_threshold(self, tfar: float):
        """Three-piece split-BF16 form: (tile, chunk) pairs with every T = p|r-P|^2 > tfar run with two pieces;
        tfar <= 0 keeps three pieces everywhere."""
        self._chk(self._L.flosic_set_far_threshold(self._h, float(tfar)))

    def set_coulomb_cols(self, ncols: int):
        """FLO columns per Coulomb block: 0 widest, -1 autotune, or 64/128/192/256."""
        self._chk(self._L.flosic_set_coulomb_cols(self._h, int(ncols)))

    def coulomb_cols(self) -> int:
        return int(self.counts()[8])

    def peak_device_bytes(self) -> int:
        return int(self.counts()[7])

    def set_mesh_block(self, block_points: int):
        """0: resident basis planes and whole-mesh fields; > 0: mesh-block recomputation."""
        self._chk(self._L.flosic_set_mesh_block(self._h, int(block_points)))

    def results(self) -> dict:
        """Everything the evaluation produced, per spin, in natural numpy layouts."""
        out = dict(esic=None, spins=[])
        for s in (1, 2):
            out["spins"].append(dict(results=self.apotnl_sic(s), sic=self.siclagm(s), vab=self.sic_vab(s),
                                     debdax=self.fod_forces(s), grad=self.fod_gradient(s)))
        return out


class FlosicBatch:
    """Independent systems (Flosic contexts on one device) evaluated together: the Coulomb contractions of all
    systems of one precision in one launch, optionally the whole device part as a replayed hipGraph."""

    def __init__(self, engines, use_graph: bool = False):
        self._L = library()
        self.engines = list(engines)
        arr = (C.c_void_p * len(self.engines))(*[e._h for e in self.engines])
        h = C.c_void_p()
        st = self._L.flosic_batch_create(len(self.engines), arr, 1 if use_graph else 0, C.byref(h))
        if st != OK:
            raise FlosicError(st, "flosic_batch_create failed")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._L.flosic_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def evaluate(self) -> np.ndarray:
        z = np.zeros(len(self.engines))
        st = self._L.flosic_batch_evaluate(self._h, _d(z))
        if st != OK:
            raise FlosicError(st, self._L.flosic_batch_last_error(self._h).decode())
        return z

    def timings(self) -> np.ndarray:
        t = np.zeros(4)
        self._L.flosic_batch_get_timings(self._h, 4, _d(t))
        return t

    def counts(self) -> np.ndarray:
        c = np.zeros(5, np.int64)
        self._L.flosic_batch_get_counts(self._h, 5, c.ctypes.data_as(_lp))
        return c
