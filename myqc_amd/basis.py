"""STO-3G basis construction with the reference's index layout.

Restates src/myQC/basis.f90:23-226 (buildBasis): reads the ``mybasis`` text
file and produces, in the reference's flat 0-based layouts,

  bas(set*OpS + k)            contraction coefficient of the k-th orbital of a set
  set(s)                      the set's primitive exponent
  basinfo = [OpS, norb, {pri, ang, ori, nprim, center} per orbital]
  setinfo = [nset, setl, {norb_in_set, maxL, center, orb0..orb(OpS-1)} per set]

A *set* is one primitive exponent on one atom shared by every function that
uses it (1s sets; 2s/2p "sp" sets).  AO order: atoms in input order, then the
file's sections in order, p as x,y,z (basis.f90:180-191).

The arrays are exactly those the C-ABI ``myqc_set_system`` takes.
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
ATOM_NAMES = ["H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne"]
BASIS_NAMES = ["STO-3G", "tester1", "tester2", "tester3"]


@dataclass
class Basis:
    bas: np.ndarray        # float64 [Anum*almax*OpS]
    set_alpha: np.ndarray  # float64 [Anum*almax]
    basinfo: np.ndarray    # int32 [2 + 5*Omax*Anum]
    setinfo: np.ndarray    # int32 [2 + Anum*almax*(3+OpS)]
    maxN: int
    maxL: int

    @property
    def OpS(self) -> int:
        return int(self.basinfo[0])

    @property
    def norb(self) -> int:
        return int(self.basinfo[1])

    @property
    def nset(self) -> int:
        return int(self.setinfo[0])

    @property
    def setl(self) -> int:
        return int(self.setinfo[1])


def _records(path):
    with open(path) as f:
        for line in f:
            toks = line.split()
            if toks:
                yield toks


_PARSED: dict = {}


def _parsed(path: str, basis_name: str):
    """The basis section of ``mybasis`` parsed once per (file, name, mtime):
    the header numbers and, per element, (sec, orb, nset, sections) with
    sections = [(func, coef, pri, ang, ori, [(values, setn), ...])], read as
    basis.f90:60-150 reads them."""
    key = (path, basis_name, os.stat(path).st_mtime_ns)
    hit = _PARSED.get(key)
    if hit is not None:
        return hit
    recs = list(_records(path))
    try:
        start = next(i for i, r in enumerate(recs) if r[0] == basis_name)
    except StopIteration:
        raise ValueError(f"basis {basis_name!r} not in {path}")
    head = tuple(int(x) for x in recs[start + 1][:5])
    maxN, maxL = (int(x) for x in recs[start + 2][:2])
    elements = {}
    for name in ATOM_NAMES:
        pos = next((k for k in range(start + 3, len(recs)) if recs[k][0] == name), None)
        if pos is None:
            continue
        sec, orb, nset = (int(x) for x in recs[pos + 1][:3])
        p = pos + 2
        sections = []
        for _ in range(sec):
            func, coef, pri, ang, ori = (int(x) for x in recs[p][:5])
            p += 1
            rows = []
            for _k in range(func):
                r = recs[p]
                p += 1
                rows.append(([float(x) for x in r[:coef]], int(round(float(r[coef])))))
            sections.append((func, coef, pri, ang, ori, rows))
        elements[name] = (sec, orb, nset, sections)
    out = (head, maxN, maxL, elements)
    _PARSED[key] = out
    return out


def build_basis(Z, basis_name: str = "STO-3G", path: str | None = None) -> Basis:
    path = path or os.path.join(DATA_DIR, "mybasis")
    (Smax, Cmax, Omax, almax, OpS), maxN, maxL, elements = _parsed(path, basis_name)
    anum = len(Z)
    bas = np.zeros(anum * almax * OpS)
    sets = np.zeros(anum * almax)
    basinfo = np.zeros(2 + 5 * Omax * anum, dtype=np.int32)
    setinfo = np.zeros(2 + anum * almax * (3 + OpS), dtype=np.int32)
    setnum = 0
    orbnum = 0
    setl = 3 + OpS
    for i, z in enumerate(Z):
        name = ATOM_NAMES[z - 1]
        if name not in elements:
            raise ValueError(f"atom {name} not in basis {basis_name}")
        sec, orb, nset, sections = elements[name]
        basinfo[0] = OpS
        basinfo[1] += orb
        setinfo[0] += nset
        setinfo[1] = setl
        for func, coef, pri, ang, ori, rows in sections:
            for val, setn in rows:
                s = setnum + setn
                sets[s] = val[coef - 1]
                setorbs = int(setinfo[2 + s * setl])
                setinfo[2 + s * setl + 2] = i
                if ori == -1:
                    setinfo[2 + s * setl + 3 + setorbs] = orbnum
                    bas[setorbs + s * OpS] = val[0]
                    setorbs += 1
                elif ori == 2:
                    for m in range(3):
                        setinfo[2 + s * setl + 3 + setorbs] = orbnum + m
                        bas[setorbs + s * OpS] = val[0]
                        setorbs += 1
                    if setinfo[2 + s * setl + 1] < 1:
                        setinfo[2 + s * setl + 1] = 1
                else:
                    raise ValueError("only s and p sections are supported (basis.f90:169-174)")
                setinfo[2 + s * setl] = setorbs
            if ori == -1:
                basinfo[2 + 5 * orbnum: 2 + 5 * (orbnum + 1)] = [pri, ang, ori, func, i]
                orbnum += 1
            else:
                for m in range(3):
                    basinfo[2 + 5 * (orbnum + m): 2 + 5 * (orbnum + m + 1)] = [pri, ang, m, func, i]
                orbnum += 3
        setnum = int(setinfo[0])
    return Basis(bas=bas, set_alpha=sets, basinfo=basinfo, setinfo=setinfo, maxN=maxN, maxL=maxL)


def read_ftab(path: str | None = None) -> np.ndarray:
    """Read the Boys table ``Ft(0:120,0:22)`` (one unformatted record, Tk
    fastest; read at src/integrals/int2e.f90:161-163).  Returns a float64 array
    of 121*23 values in that column-major order."""
    path = path or os.path.join(DATA_DIR, "Ftab")
    with open(path, "rb") as f:
        b = f.read()
    n = int.from_bytes(b[:4], "little", signed=True)
    if n != 121 * 23 * 8 or len(b) < n + 8:
        raise ValueError(f"{path}: not a 121x23 Boys table record")
    return np.frombuffer(b[4:4 + n], dtype="<f8").astype(np.float64).copy()
