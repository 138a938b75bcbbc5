"""Write golden vectors from the reference's own routines (oracle/_ref/libflosic_ref.so).

The library is compiled from the reference sources by ``make -C oracle ref`` (only where the reference is
mounted).  This script calls, through oracle/ref_driver.f90:
  FFFMT3/FFFMT4/FFFMTC  Boys functions on a T grid
  POISSON2 (DEBUG)      exact potential of random s/p/d Cartesian pair densities at random points
  GORBDRV               values, gradients and Hessians of an O-atom ladder basis at 25 points
  PW91EX/PW91LC/PBEEX/PBECOR  at random arguments
  FRMORB2(-ISPN,-1)     FOD force (DEBDAX) and SLAT for FMAT/SIC of a small seeded system
  SUBVLXC (NSPN = 2)    DFT-side LSDA and PBE potentials and energies of random spin densities
and stores inputs and outputs in ref_vectors.npz.  Run from the repository root:
    python tests/golden/make_ref_vectors.py
"""
import ctypes as C
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from publicrelease_2020_amd import system as S  # noqa: E402
from oracle import oracle as O  # noqa: E402

LIB = os.path.join(ROOT, "oracle", "_ref", "libflosic_ref.so")
dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int32)


def d(a):
    return a.ctypes.data_as(dp)


def main():
    R = C.CDLL(LIB)
    out = {}
    here = os.getcwd()
    os.chdir(tempfile.mkdtemp())
    # Boys
    T = np.concatenate([np.linspace(0.0, 20.0, 801), np.random.default_rng(11).uniform(0.0, 60.0, 200)])
    for top in (2, 3, 4):
        F = np.zeros((5, T.size))
        R.ref_boys(C.c_int(top), C.c_int(T.size), d(T), d(F))
        out[f"boys_T"] = T
        out[f"boys_F{top}"] = F[:top + 1].T.copy()
    # POISSON
    rng = np.random.default_rng(12)
    npair = 60
    al = rng.uniform(0.15, 8.0, npair)
    be = rng.uniform(0.15, 8.0, npair)
    A = rng.uniform(-1.5, 1.5, (npair, 3))
    B = rng.uniform(-1.5, 1.5, (npair, 3))
    rho = rng.standard_normal((npair, 10, 10))
    pts = rng.uniform(-5.0, 5.0, (400, 3))
    V = np.zeros(pts.shape[0])
    R.ref_poisson(C.c_int(npair), d(np.ascontiguousarray(al)), d(np.ascontiguousarray(A)), d(np.ascontiguousarray(be)),
                  d(np.ascontiguousarray(B)), d(np.ascontiguousarray(rho)), C.c_int(pts.shape[0]), d(pts), d(V))
    out.update(poisson_alpha=al, poisson_A=A, poisson_beta=be, poisson_B=B, poisson_rho=rho, poisson_pts=pts,
               poisson_V=V)
    # GORBDRV
    b = S.build_basis(np.zeros((1, 3)), [8])
    nb = b.nbare(0)
    ncon = b.ncon[0].astype(np.int32).copy()
    maxcon = int(ncon.max())
    con = np.zeros((3, maxcon, nb))
    for l in range(3):
        if ncon[l]:
            con[l, :ncon[l], :] = b.con_block(0, l)
    gpts = np.random.default_rng(13).uniform(-0.9, 0.9, (25, 3))
    grad = np.zeros((3, 15, 6, 10, 25))
    R.ref_gorbdrv(C.c_int(2), C.c_int(nb), d(np.ascontiguousarray(b.alpha)), ncon.ctypes.data_as(ip),
                  C.c_int(maxcon), d(np.ascontiguousarray(con)), C.c_int(2), C.c_int(25), d(gpts), d(grad))
    ref = []
    for l in range(3):
        for m in range(S.NCART[l]):
            for c in range(ncon[l]):
                ref.append(grad[l, c, m, :, :])
    out.update(gorbdrv_pts=gpts, gorbdrv_out=np.stack(ref, axis=-1))  # (10, 25, nbas)
    # XC kernels
    rng = np.random.default_rng(14)
    n = 300
    xin = np.stack([rng.uniform(0.05, 4.0, n), rng.uniform(0.0, 4.0, n), rng.uniform(-3, 3, n), rng.uniform(-3, 3, n)], 1)
    pw91ex = np.zeros((n, 3))
    pbeex = np.zeros((n, 3))
    for i in range(n):
        o = (C.c_double * 3)()
        R.ref_pw91ex(*[C.c_double(x) for x in xin[i]], C.byref(o, 0), C.byref(o, 8), C.byref(o, 16))
        pw91ex[i] = list(o)
        o = (C.c_double * 3)()
        R.ref_pbeex(*[C.c_double(x) for x in xin[i]], C.byref(o, 0), C.byref(o, 8), C.byref(o, 16))
        pbeex[i] = list(o)
    cin = np.stack([rng.uniform(0.01, 20.0, n), np.where(rng.random(n) < 0.5, 1.0, rng.uniform(-0.9, 0.9, n)),
                    rng.uniform(0.0, 3.0, n), rng.uniform(-2, 2, n), rng.uniform(-2, 2, n), np.zeros(n)], 1)
    pw91lc = np.zeros((n, 3))
    pbecor = np.zeros((n, 6))
    for i in range(n):
        o = (C.c_double * 3)()
        R.ref_pw91lc(C.c_double(cin[i, 0]), C.c_double(cin[i, 1]), C.byref(o, 0), C.byref(o, 8), C.byref(o, 16))
        pw91lc[i] = list(o)
        o = (C.c_double * 6)()
        R.ref_pbecor(*[C.c_double(x) for x in cin[i]], *[C.byref(o, 8 * k) for k in range(6)])
        pbecor[i] = list(o)
    out.update(xc_xin=xin, xc_pw91ex=pw91ex, xc_pbeex=pbeex, xc_cin=cin, xc_pw91lc=pw91lc, xc_pbecor=pbecor)
    # FRMORB2 force from a seeded small system
    sy = S.make_synthetic(3, 3, 7, nrad=24, nang=50, box=4.5)
    res = O.evaluate(sy, want_vab=False)
    sp = res["spins"]
    n1, n2 = sp[0]["q"].size, sp[1]["q"].size
    f1 = np.ascontiguousarray(sp[0]["fmat"])
    f2 = np.ascontiguousarray(sp[1]["fmat"])
    for s in (1, 2):
        nn = [n1, n2][s - 1]
        sic = np.ascontiguousarray(sp[s - 1]["sic"][:nn, :nn].T)
        deb = np.zeros((nn, 3))
        slat = np.zeros((nn, nn))
        R.ref_frmorb2_force(C.c_int(s), C.c_int(n1), C.c_int(n2), d(f1), d(f2), d(sic), d(deb), d(slat))
        out[f"frm_sic{s}"] = sp[s - 1]["sic"][:nn, :nn].copy()
        out[f"frm_debdax{s}"] = deb
        out[f"frm_slat{s}"] = slat.T.copy()
    out.update(frm_fmat1=f1, frm_fmat2=f2)
    # SUBVLXC (NSPN = 2): spin densities with value, gradient and Hessian, LSDA (IGGA = 0) and PBE (IGGA = 1)
    rng = np.random.default_rng(23)
    nd = 250
    rgu = rng.normal(size=(nd, 10)) * 10.0 ** rng.uniform(-3, 1, (nd, 1))
    rgd = rng.normal(size=(nd, 10)) * 10.0 ** rng.uniform(-3, 1, (nd, 1))
    rgu[:, 0] = 10.0 ** rng.uniform(-4, 2, nd)
    rgd[:, 0] = rgu[:, 0] * 10.0 ** rng.uniform(-3, 0.5, nd)
    rgd[:10, :] = 0.0  # fully polarised points
    wd = rng.uniform(0.1, 2.0, nd)
    for name, gga in (("lsda", 0), ("pbe", 1)):
        pot = np.zeros((2, nd))
        e = np.zeros(4)
        R.ref_subvlxc(C.c_int(gga), C.c_int(nd), d(rgu), d(rgd), d(wd), d(pot), d(e))
        out[f"dft_pot_{name}"] = pot
        out[f"dft_e_{name}"] = e
    out.update(dft_rgu=rgu, dft_rgd=rgd, dft_w=wd)
    os.chdir(here)
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "ref_vectors.npz"), **out)
    print("wrote", sorted(out.keys()))


if __name__ == "__main__":
    main()
