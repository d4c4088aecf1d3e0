"""Structure readers: the file formats of the reference's atomicStructure
module (src/atoms/atomicStructure.f90), for ingesting requests.

Each reader returns what the reference's reader makes of the file, down to
its quirks, so that a structure read here and one read by a Fortran program
linked against the reference are the same:

- ``read_xyz``      as_readXYZ      :143-166  bohr, nonperiodic
- ``read_fhiaims``  as_readFhiaims  :168-219  Angstrom; lines starting with
  ``atom`` (column 1, so ``atom_frac`` too, read as Cartesian) and
  ``lattice_vector``; periodic when any lattice vector is given
- ``read_ascii``    as_readAscii    :245-285  v_sim ascii, always periodic;
  lower-triangular lattice on two lines; Angstrom when the first line says
  ``angstroem`` or ``is_angstrom`` is set
- ``read_poscar``   as_readPOSCAR   :427-475  VASP 5, Direct coordinates
  only; the scale factor on line 2 is read but not applied (as there)
- ``read_data``     as_readData     :354-425  RuNNer input.data blocks, bohr
  and Hartree; periodic when exactly three lattice lines are given

Values are read the way Fortran list-directed input reads them: each read
starts on a new line, takes further lines until it has its values and drops
the rest of its last line; ``d``/``D`` exponents are accepted; an element
symbol keeps its first two characters and maps to Z by exact match (-1 when
unknown, as elemSymToNum, constants.f90:95-111).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List

import numpy as np

from .workloads import Structure

ANG2BOHR = 1.8897259886                      # constants.f90:7

_SYMBOLS = (
    "H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge As Se Br Kr "
    "Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb "
    "Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn Fr Ra Ac Th Pa U Np Pu Am Cm Bk Cf Es Fm Md No Lr Rf "
    "Db Sg Bh Hs Mt Ds Rg Cn Nh Fl Mc Lv Ts Og").split()
_Z = {s: k + 1 for k, s in enumerate(_SYMBOLS)}


class ReadError(ValueError):
    """A file the reference's reader would stop on."""


def element_number(sym: str) -> int:
    """elemSymToNum of a two-character item: Z, or -1."""
    return _Z.get(sym[:2].strip(), -1)


def _real(tok: str) -> float:
    return float(tok.replace("d", "e").replace("D", "E"))


class _Records:
    """The lines of a file, read as Fortran formatted records."""

    def __init__(self, path: str):
        with open(path) as f:
            self.lines = f.read().splitlines()
        self.i = 0

    def line(self) -> str:
        """A '(A)' read: the next line verbatim."""
        if self.i >= len(self.lines):
            raise ReadError("unexpected end of file")
        self.i += 1
        return self.lines[self.i - 1]

    def items(self, n: int) -> List[str]:
        """A list-directed read of n items."""
        out: List[str] = []
        while len(out) < n:
            out.extend(_split(self.line()))
        return out[:n]


def _split(line: str) -> List[str]:
    return line.replace(",", " ").split()


def _items_of(line: str, n: int, what: str) -> List[str]:
    """A list-directed read of n items from one line (an internal read)."""
    t = _split(line)
    if len(t) < n:
        raise ReadError(f"{what}: expected {n} items in {line!r}")
    return t[:n]


def read_xyz(path: str) -> Structure:
    r = _Records(path)
    nat = int(r.items(1)[0])
    r.line()
    Z = np.zeros(nat, dtype=np.int32)
    pos = np.zeros((nat, 3))
    for i in range(nat):
        t = r.items(4)
        Z[i] = element_number(t[0])
        pos[i] = [_real(x) for x in t[1:4]]
    return Structure(pos, Z, np.zeros((3, 3)), False)


def read_fhiaims(path: str) -> Structure:
    r = _Records(path)
    pos, Z, lat, nlat = [], [], np.zeros((3, 3)), 0
    for ln in r.lines:
        ln = ln[:200]
        if ln[:4] == "atom":
            t = _items_of(ln, 5, "atom")
            pos.append([_real(x) for x in t[1:4]])
            Z.append(element_number(t[4]))
        elif ln[:14] == "lattice_vector":
            if nlat >= 3:
                raise ReadError("more than 3 lattice vectors")
            t = _items_of(ln, 4, "lattice_vector")
            lat[nlat] = [_real(x) for x in t[1:4]]
            nlat += 1
    pos = np.array(pos, dtype=np.float64).reshape(-1, 3) * ANG2BOHR
    if nlat:
        return Structure(pos, np.array(Z, dtype=np.int32), lat * ANG2BOHR, True)
    return Structure(pos, np.array(Z, dtype=np.int32), np.zeros((3, 3)), False)


def read_ascii(path: str, is_angstrom: bool = False) -> Structure:
    r = _Records(path)
    head = _split(r.line())
    nat = int(head[0])
    angs = is_angstrom or (len(head) > 1 and head[1][:9] == "angstroem")
    a = [_real(x) for x in r.items(3)]
    b = [_real(x) for x in r.items(3)]
    lat = np.array([[a[0], 0.0, 0.0], [a[1], a[2], 0.0], [b[0], b[1], b[2]]])
    Z = np.zeros(nat, dtype=np.int32)
    pos = np.zeros((nat, 3))
    for i in range(nat):
        t = r.items(4)
        pos[i] = [_real(x) for x in t[:3]]
        Z[i] = element_number(t[3])
    if angs:
        pos, lat = pos * ANG2BOHR, lat * ANG2BOHR
    return Structure(pos, Z, lat, True)


def read_poscar(path: str) -> Structure:
    r = _Records(path)
    r.line()                                   # comment
    r.items(1)                                 # scale: read, not applied
    lat = np.array([[_real(x) for x in r.items(3)] for _ in range(3)]) * ANG2BOHR
    syms = _split(r.line())
    counts = [int(x) for x in r.items(len(syms))]
    mode = r.line().lower().strip()
    if mode != "direct":
        raise ReadError(f"POSCAR without DIRECT not implemented ({mode!r})")
    Z = np.concatenate([np.full(c, element_number(s), dtype=np.int32) for s, c in zip(syms, counts)]
                       ) if syms else np.zeros(0, np.int32)
    pos = np.zeros((len(Z), 3))
    for i in range(len(Z)):
        s = [_real(x) for x in r.items(3)]
        # matMulVec3(lat, s) (linalg.f90:192-201): component a = sum_k lat(a,k) s(k)
        for c in range(3):
            pos[i, c] = (lat[0, c] * s[0] + lat[1, c] * s[1]) + lat[2, c] * s[2]
    return Structure(pos, Z, lat, True)


@dataclass
class DataStructure(Structure):
    """A structure of an input.data file with its reference data."""
    energy: float = 0.0
    qtot: float = 0.0
    charges: np.ndarray = field(default_factory=lambda: np.zeros(0))
    forces: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))


def read_data(path: str) -> List[DataStructure]:
    r = _Records(path)
    out: List[DataStructure] = []
    cur = None
    for raw in r.lines:
        ln = raw[:512].strip()
        if ln[:5] == "begin":
            cur = {"lat": np.zeros((3, 3)), "nlat": 0, "pos": [], "Z": [], "q": [], "f": [],
                   "energy": 0.0, "qtot": 0.0}
        if cur is None:
            continue
        if ln[:7] == "lattice":
            if cur["nlat"] >= 3:
                raise ReadError("more than 3 lattice lines in a structure")
            cur["lat"][cur["nlat"]] = [_real(x) for x in _items_of(ln, 4, "lattice")[1:4]]
            cur["nlat"] += 1
        if ln[:4] == "atom":
            t = _items_of(ln, 10, "atom")
            cur["pos"].append([_real(x) for x in t[1:4]])
            cur["Z"].append(element_number(t[4]))
            cur["q"].append(_real(t[5]))
            cur["f"].append([_real(x) for x in t[7:10]])
        if ln[:6] == "energy":
            cur["energy"] = _real(_items_of(ln, 2, "energy")[1])
        if ln[:6] == "charge":
            cur["qtot"] = _real(_items_of(ln, 2, "charge")[1])
        if ln[:3] == "end":
            periodic = cur["nlat"] == 3
            out.append(DataStructure(np.array(cur["pos"], dtype=np.float64).reshape(-1, 3),
                                     np.array(cur["Z"], dtype=np.int32),
                                     cur["lat"] if periodic else np.zeros((3, 3)), periodic,
                                     energy=cur["energy"], qtot=cur["qtot"],
                                     charges=np.array(cur["q"], dtype=np.float64),
                                     forces=np.array(cur["f"], dtype=np.float64).reshape(-1, 3)))
    return out


READERS = {"xyz": read_xyz, "fhiaims": read_fhiaims, "ascii": read_ascii, "poscar": read_poscar,
           "data": read_data}


def read_structures(path: str, fmt: str) -> List[Structure]:
    """Every structure of a file in one of the READERS formats."""
    if fmt not in READERS:
        raise ReadError(f"unknown format {fmt!r} (one of {sorted(READERS)})")
    got = READERS[fmt](path)
    return got if isinstance(got, list) else [got]
