"""The reference's on-disk formats around the SIC path, read and written so that a reference run directory can be
evaluated with the library and its outputs handed back in the files the reference's tools read.

  SYMBOL + ISYMGEN -> INPUT   symbolic input expansion                         symbol.F:1-270
  INPUT                       atom types, positions, basis, electrons          readinp.F:48-236
  FRMORB / FRMIDT             "n_up n_dn" then one FOD per line                main.F:142-164
  fforce.dat                  DEBDAX per FOD, ' ',3F20.12                      main.F:708-715, 794
  WFOUT                       Fortran unformatted KS wavefunction records      wfout.F:21-41
  OLDMESH / VMOLD             Fortran unformatted mesh (NMSH, ICALC; R; W)     dvpmesh.F:897-902, vmesh.F:184-193

Unformatted files use the usual sequential layout: every record is framed by its byte length as a 4-byte integer
before and after (the record-marker convention of gfortran and flang).  Symmetry is not expanded: every position
in INPUT is one atom (point group NONE, the setting FLO-SIC runs with, where UNRAVEL is the identity).
"""
from __future__ import annotations

import dataclasses
import io
import os
import re
import struct
from typing import List, Optional, Tuple

import numpy as np

from .system import Basis, Mesh, System, XC_LSDA, XC_PBE

NCART = (1, 3, 6)


# ------------------------------------------------------------------------------------------------ SYMBOL
@dataclasses.dataclass
class SymbolFile:
    mode: str                  # relaxation mode line (SCF-ONLY, LBFGS, ...)
    functional: str            # e.g. LDA-PW91*LDA-PW91
    mesh: str                  # OLDMESH or DVPMESH
    files: List[Tuple[str, str]]   # (symbolic file, expanded file)
    symbols: List[Tuple[str, str]]  # (name, value) of the selected calculation block, in file order


def _split_symbol_line(line: str) -> Tuple[str, str]:
    """symbol.F:206-229: the name is the non-blank characters before '=', the value keeps single blanks where
    the line had runs of blanks; '!' starts a comment; names are cut to 16 characters."""
    line = line[:80]
    name, value, side, blank = [], [], 0, False
    for ch in line:
        if ch == "!":
            break
        if ch == "=":
            side = 1
        elif ch != " " and ch != "\t":
            blank = False
            (name if side == 0 else value).append(ch)
        elif side == 1 and not blank:
            blank = True
            value.append(" ")
    return "".join(name)[:16], "".join(value)


def read_symbol(text: str, ncalc: int = 1) -> SymbolFile:
    """The SYMBOL file: header, symbolic file pairs, identity list and the symbol block of calculation ncalc
    (the first block carrying that index, symbol.F:197-205)."""
    lines = text.splitlines()
    mode, functional, mesh = lines[0].strip(), lines[1].split()[0], lines[2][:7].strip()
    k = 3
    nfiles = int(lines[k].split()[0])
    k += 1
    files = []
    for _ in range(nfiles):
        a, _, b = lines[k].split("!")[0].partition("=")
        files.append((a.replace(" ", ""), b.replace(" ", "")))
        k += 1
    nsymbs = int(lines[k].split()[0])
    nidsym = int(lines[k + 1].split()[0])
    k += 2 + nidsym
    while k < len(lines):
        if not lines[k].strip():
            k += 1
            continue
        icalc = int(lines[k].split()[0])
        block = lines[k + 1:k + 1 + nsymbs]
        if icalc == ncalc:
            return SymbolFile(mode, functional, mesh, files, [_split_symbol_line(l) for l in block])
        k += 1 + nsymbs
    raise ValueError(f"SYMBOL has no block for calculation {ncalc}")


def expand_symbolic(template: str, symbols: List[Tuple[str, str]]) -> str:
    """symbol.F:240-262: a line that starts with a symbol's name is replaced by that symbol's value (the last
    matching symbol wins); other lines are copied."""
    out = []
    for line in template.splitlines():
        line80 = line[:80]
        rep = None
        for name, value in symbols:
            if name and line80[:len(name)] == name:
                rep = value
        out.append(rep if rep is not None else line80)
    return "\n".join(out) + "\n"


# ------------------------------------------------------------------------------------------------ INPUT
@dataclasses.dataclass
class InputFile:
    basis: Basis
    charges: np.ndarray        # nuclear charge per atom
    electrons: Tuple[float, float]   # E_UP, E_DN as written (a negative E_DN flags the down count)
    wffile: str
    spin_tags: List[int]       # +1 SUP, -1 SDN, 0 otherwise, per atom


def _fortran_floats(s: str) -> List[float]:
    return [float(t.replace("D", "E").replace("d", "e")) for t in s.replace(",", " ").split()]


class _ListReader:
    """List-directed READ over the lines of a formatted file: each READ starts at a new record and takes as
    many values as it needs, continuing over following records."""

    def __init__(self, text: str):
        self.lines = text.splitlines()
        self.k = 0

    def line(self) -> str:
        s = self.lines[self.k]
        self.k += 1
        return s

    def values(self, n: int, conv=float) -> list:
        out: list = []
        while len(out) < n:
            toks = self.line().replace(",", " ").split()
            for t in toks:
                if len(out) == n:
                    break
                out.append(conv(float(t.replace("D", "E").replace("d", "e"))) if conv is not str else t)
        return out


def read_input(text: str) -> InputFile:
    """INPUT as READINP reads it (readinp.F:55-220): per atom type the charges, potential type, positions,
    EXTRABASIS, bare exponents (sorted from large to small, readinp.F:147-166, with the coefficients following
    their exponents) and the contraction rows of s, p, d; supplementary functions are added when EXTRABASIS > 0
    and skipped otherwise; then the electron counts and the wavefunction file name."""
    r = _ListReader(text)
    nfnct = int(r.values(1)[0])
    xyz, Z, tags = [], [], []
    bare_off, alpha, ncon_l, con_off_l, con = [0], [], [], [], []
    for _ in range(nfnct):
        zelc, znuc = r.values(2)
        psp = r.line()[:7]
        if psp[:3] != "ALL":
            raise ValueError("pseudopotential atom types are not supported by the SIC path")
        npos = int(r.values(1)[0])
        pos = []
        for _ in range(npos):
            line = r.line()
            tag = 1 if "SUP" in line[:80] else (-1 if "SDN" in line[:80] else 0)
            pos.append((_fortran_floats(" ".join(line.split()[:3]))[:3], tag))
        iextra = int(r.values(1)[0])
        nbare = int(r.values(1)[0])
        ncon = [int(v) for v in r.values(3)]
        nfadd = [int(v) for v in r.values(3)]
        alp = np.array(r.values(nbare))
        order = np.argsort(-alp, kind="stable")  # large to small
        blocks = []
        for l in range(3):
            n = ncon[l] + (nfadd[l] if iextra > 0 else 0)
            rows = [np.array(r.values(nbare)) for _ in range(n)]
            if iextra <= 0:
                for _ in range(nfadd[l]):
                    r.values(nbare)
            blocks.append(np.array(rows).reshape(n, nbare)[:, order] if n else np.zeros((0, nbare)))
        for (p, tag) in pos:
            xyz.append(p)
            Z.append(int(round(znuc)))
            tags.append(tag)
            alpha.extend(alp[order].tolist())
            bare_off.append(bare_off[-1] + nbare)
            offs, ns = [], []
            for l in range(3):
                offs.append(len(con))
                ns.append(blocks[l].shape[0])
                con.extend(blocks[l].ravel().tolist())
            ncon_l.append(ns)
            con_off_l.append(offs)
    e_up, e_dn = r.values(2)
    wffile = r.line().strip() if r.k < len(r.lines) else "WFOUT"
    basis = Basis(xyz=np.array(xyz, dtype=np.float64).reshape(-1, 3), Z=np.array(Z, dtype=np.int32),
                  bare_off=np.array(bare_off, dtype=np.int32), alpha=np.array(alpha, dtype=np.float64),
                  ncon=np.array(ncon_l, dtype=np.int32).reshape(-1, 3),
                  con_off=np.array(con_off_l, dtype=np.int32).reshape(-1, 3), con=np.array(con, dtype=np.float64))
    return InputFile(basis=basis, charges=basis.Z.astype(np.float64), electrons=(e_up, e_dn), wffile=wffile,
                     spin_tags=tags)


# ------------------------------------------------------------------------------------------------ FRMORB, fforce
def read_frmorb(text: str) -> Tuple[np.ndarray, np.ndarray]:
    """FRMORB (main.F:150-163): LFRM(1) LFRM(2) on the first record, then one FOD per record, up then down.  Each
    READ takes the first values of its record, so symfrm.f's layout (four counts in the header; each FOD followed
    by its weight and equivalence indices; blank records between the spins) reads the same way."""
    r = _ListReader(text)
    n_up, n_dn = (int(v) for v in r.values(2))
    up = np.array([r.values(3) for _ in range(n_up)], dtype=np.float64).reshape(n_up, 3)
    dn = np.array([r.values(3) for _ in range(n_dn)], dtype=np.float64).reshape(n_dn, 3)
    return up, dn


def write_frmorb(fod_up, fod_dn) -> str:
    """FRMORB in symfrm.f's layout (symfrm.f:136-165) without symmetry: header LFRM(1) LFRM(2) LFRM(1) LFRM(2)
    (every FOD its own identity member), each FOD at full precision followed by its weight 1 and its own index,
    two blank records after each spin.  main.F reads the first two counts and three coordinates per record; the FOD
    optimiser (electronic_geometry_lbfgs.f:70-83) also reads the weight and index."""
    up = np.asarray(fod_up, dtype=np.float64).reshape(-1, 3)
    dn = np.asarray(fod_dn, dtype=np.float64).reshape(-1, 3)
    out = io.StringIO()
    out.write(f"{up.shape[0]:5d}{dn.shape[0]:5d}{up.shape[0]:9d}{dn.shape[0]:5d}\n")
    for fods in (up, dn):
        for i, p in enumerate(fods):
            out.write(" ".join(f"{x:24.16E}" for x in p) + f" {1:4d} {i + 1:4d}\n")
        out.write("  \n  \n")
    return out.getvalue()


def write_fforce(debdax_up, debdax_dn=None) -> str:
    """fforce.dat as main.F writes it (' ',3F20.12 per FOD, spin 1 then spin 2, main.F:708-735)."""
    rows = [np.asarray(debdax_up, dtype=np.float64).reshape(-1, 3)]
    if debdax_dn is not None:
        rows.append(np.asarray(debdax_dn, dtype=np.float64).reshape(-1, 3))
    return "".join(" " + "".join(f"{x:20.12f}" for x in p) + "\n" for p in np.vstack(rows))


def read_fforce(text: str) -> np.ndarray:
    return np.array([_fortran_floats(l) for l in text.splitlines() if l.strip()], dtype=np.float64).reshape(-1, 3)


# ------------------------------------------------------------------------------------------------ unformatted
def _records(data: bytes) -> List[bytes]:
    out, k = [], 0
    while k < len(data):
        (n,) = struct.unpack_from("<i", data, k)
        rec = data[k + 4:k + 4 + n]
        (m,) = struct.unpack_from("<i", data, k + 4 + n)
        if m != n:
            raise ValueError("corrupt unformatted record markers")
        out.append(rec)
        k += 8 + n
    return out


def _record(payload: bytes) -> bytes:
    return struct.pack("<i", len(payload)) + payload + struct.pack("<i", len(payload))


@dataclasses.dataclass
class Wavefunctions:
    """WFOUT contents (wfout.F:21-41) for one representation per spin (point group NONE)."""
    nspn: int
    nwf: int
    nwfs: List[int]
    evlocc: np.ndarray
    occupancy: List[np.ndarray]   # per spin
    coef: List[np.ndarray]        # per spin (ns_tot, n_occ): PSI_COEF columns, rows in the file's SALC order


def read_wfout(data: bytes) -> Wavefunctions:
    recs = _records(data)
    k = 0
    (nspn,) = struct.unpack("<i", recs[k]); k += 1
    ints = struct.unpack(f"<{len(recs[k]) // 4}i", recs[k]); k += 1
    nwf, nwfs = ints[0], list(ints[1:1 + nspn])
    evlocc = np.frombuffer(recs[k], dtype="<f8").copy(); k += 1
    (n_rep,) = struct.unpack("<i", recs[k]); k += 1
    if n_rep != 1:
        raise ValueError("WFOUT with symmetry representations: the SIC path runs with point group NONE")
    occ, coef = [], []
    for _ in range(nspn):
        n_occ, ns_tot = struct.unpack("<2i", recs[k]); k += 1
        occ.append(np.frombuffer(recs[k], dtype="<f8").copy()); k += 1
        cols = []
        for _ in range(n_occ):
            cols.append(np.frombuffer(recs[k], dtype="<f8").copy()); k += 1
        coef.append(np.array(cols, dtype=np.float64).reshape(n_occ, ns_tot).T.copy())
    return Wavefunctions(nspn=nspn, nwf=nwf, nwfs=nwfs, evlocc=evlocc, occupancy=occ, coef=coef)


def salc_order(basis: Basis) -> np.ndarray:
    """Index map between WFOUT's coefficient order and the local basis order, point group NONE.

    WFOUT holds PSI_COEF in the symmetry-adapted (SALC) basis.  With point group NONE every site is its own shell and
    U_MAT is the identity, so UNRAVEL's loops (unravel.f:51-89) read, per site, l outer, then the contraction IBASE,
    then the SALC index IQ = the Cartesian component: coefficient j of the file is local function ``order[j]`` of the
    site-major (l, component, contraction) order of gorbdrv.f / frmorb2.F:262-266."""
    local = {}
    k = 0
    for a in range(basis.natom):
        for l in range(3):
            for m in range(NCART[l]):
                for c in range(int(basis.ncon[a, l])):
                    local[(a, l, m, c)] = k
                    k += 1
    return np.array([local[(a, l, m, c)] for a in range(basis.natom) for l in range(3)
                     for c in range(int(basis.ncon[a, l])) for m in range(NCART[l])], dtype=np.int64)


def salc_to_local(basis: Basis, coef: np.ndarray) -> np.ndarray:
    """WFOUT coefficient rows (SALC order) -> local basis order."""
    out = np.empty_like(coef)
    out[salc_order(basis)] = coef
    return out


def local_to_salc(basis: Basis, coef: np.ndarray) -> np.ndarray:
    """Local basis order -> WFOUT coefficient rows (SALC order)."""
    return np.ascontiguousarray(np.asarray(coef)[salc_order(basis)])


def write_wfout(wf: Wavefunctions) -> bytes:
    out = [_record(struct.pack("<i", wf.nspn)),
           _record(struct.pack(f"<{1 + wf.nspn}i", wf.nwf, *wf.nwfs)),
           _record(np.asarray(wf.evlocc, dtype="<f8").tobytes()),
           _record(struct.pack("<i", 1))]
    for s in range(wf.nspn):
        c = np.asarray(wf.coef[s], dtype=np.float64)
        out.append(_record(struct.pack("<2i", c.shape[1], c.shape[0])))
        out.append(_record(np.asarray(wf.occupancy[s], dtype="<f8").tobytes()))
        for j in range(c.shape[1]):
            out.append(_record(np.ascontiguousarray(c[:, j], dtype="<f8").tobytes()))
    return b"".join(out)


def read_mesh(data: bytes) -> Tuple[np.ndarray, np.ndarray, int]:
    """OLDMESH / VMOLD: NMSH, ICALC; RMSH(3,NMSH); WMSH(NMSH).  Returns (xyz (NMSH,3), w, icalc)."""
    recs = _records(data)
    nmsh, icalc = struct.unpack("<2i", recs[0])
    xyz = np.frombuffer(recs[1], dtype="<f8").reshape(nmsh, 3).copy()
    w = np.frombuffer(recs[2], dtype="<f8").copy()
    return xyz, w, icalc


def write_mesh(xyz, w, icalc: int = 1) -> bytes:
    xyz = np.asarray(xyz, dtype="<f8").reshape(-1, 3)
    return b"".join([_record(struct.pack("<2i", xyz.shape[0], icalc)), _record(xyz.tobytes()),
                     _record(np.asarray(w, dtype="<f8").tobytes())])


# ------------------------------------------------------------------------------------------------ run directory
def functional_xc(functional: str) -> int:
    """SETDFTYP names of the two functionals the SIC path supports (setdftyp.F:5-110)."""
    f = functional.upper()
    if f.startswith("GGA-PBE"):
        return XC_PBE
    if f.startswith("LDA-PW91") or f.startswith("LDA-PZ") or f.startswith("LDA"):
        return XC_LSDA
    raise ValueError(f"functional {functional} is not supported by the SIC path")


def load_run(directory: str, frmorb: Optional[str] = None) -> System:
    """A reference run directory as one SIC evaluation input: SYMBOL (+ the symbolic ISYMGEN, or INPUT), the
    FODs (FRMORB, else FRMIDT), the KS wavefunctions (the WFOUT named in INPUT) and the mesh (VMOLD, else
    OLDMESH).  The occupied orbitals are the columns with occupancy > 0.5."""
    def rd(name, mode="r"):
        with open(os.path.join(directory, name), mode) as f:
            return f.read()

    sym = read_symbol(rd("SYMBOL"))
    inp_text = None
    for src, dst in sym.files:
        if dst.upper() == "INPUT" and os.path.exists(os.path.join(directory, src)):
            inp_text = expand_symbolic(rd(src), sym.symbols)
    if inp_text is None:
        inp_text = rd("INPUT")
    inp = read_input(inp_text)
    fname = frmorb or ("FRMORB" if os.path.exists(os.path.join(directory, "FRMORB")) else "FRMIDT")
    up, dn = read_frmorb(rd(fname))
    wf = read_wfout(rd(inp.wffile, "rb"))
    mesh_name = "VMOLD" if os.path.exists(os.path.join(directory, "VMOLD")) else "OLDMESH"
    xyz, w, _ = read_mesh(rd(mesh_name, "rb"))
    coef, nocc = [], []
    for s, fods in enumerate((up, dn)):
        c = wf.coef[min(s, wf.nspn - 1)]  # NSPN = 1: both spins share the orbitals
        if c.shape[1] < fods.shape[0]:
            raise ValueError(f"spin {s + 1}: {fods.shape[0]} FODs but {c.shape[1]} orbitals in {inp.wffile}")
        if c.shape[0] != inp.basis.nbas:
            raise ValueError(f"{inp.wffile}: {c.shape[0]} coefficients per orbital, the basis has {inp.basis.nbas}")
        coef.append(salc_to_local(inp.basis, c))
        nocc.append(fods.shape[0])  # one FOD per occupied orbital (frmiorb.f:443-449)
    return System(name=os.path.basename(os.path.normpath(directory)), basis=inp.basis, mesh=Mesh(xyz=xyz, w=w),
                  coef=coef, nocc=nocc, fod=[up, dn], xc=functional_xc(sym.functional), nspn=wf.nspn)
