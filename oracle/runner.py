"""Oracle side: RuNNer parameter parser and ctypes binding of liboracle.so.

TEST INFRASTRUCTURE ONLY (the checker). tests/, the smoke check and the
cpu_baseline leg of bench.py may import this module; product code must not.

``read_potential`` restates parseRuNNerInput (src/potentials/hdnnps.f90:34-233)
with its helpers sortSymFunctions / symfunctionstring (:235-267), readScaling
(:269-300), readWeights (:302-330), splitString (src/util/util.f90:128-161),
elemSymToNum (src/constants/constants.f90:95-111) and the stable merge sorts
(src/util/sorting.f90:26-177).
"""
from __future__ import annotations

import ctypes
import dataclasses
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

ELEMENT_SYMBOLS = (
    "H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge "
    "As Se Br Kr Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce Pr Nd Pm Sm "
    "Eu Gd Tb Dy Ho Er Tm Yb Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn Fr Ra Ac Th Pa U "
    "Np Pu Am Cm Bk Cf Es Fm Md No Lr Rf Db Sg Bh Hs Mt Ds Rg Cn Nh Fl Mc Lv Ts Og").split()
NMAX_SFS = 200   # hdnnps.f90:48


class RunnerError(RuntimeError):
    pass


def elem_sym_to_num(sym: str) -> int:
    """constants.f90:95-111 on a character(len=2) value: -1 if unknown."""
    s = sym[:2].strip()
    for i, e in enumerate(ELEMENT_SYMBOLS, start=1):
        if s == e:
            return i
    return -1


def split_string(line: str):
    """util.f90:128-161: blanks separate words; a tab is not a separator."""
    return [w for w in line.split(" ") if w != ""]


def _freal(tok: str) -> float:
    return float(tok.replace("d", "e").replace("D", "e"))


def _fmt_f(x: float, w: int, d: int) -> str:
    s = f"{x:{w}.{d}f}"
    return "*" * w if len(s) > w else s


def _fmt_i(n: int, w: int) -> str:
    s = f"{n:{w}d}"
    return "*" * w if len(s) > w else s


@dataclass
class SymFunction:
    name: int                 # 2, 3, or 9 (RuNNer's wide angular G5, beyond the reference)
    els: tuple                # (Z1,) or (Z1, Z2) as written in the file
    rcut: float
    eta: float
    rs: float = 0.0
    lam: float = 0.0
    zeta: float = 0.0
    sfmin: float = 0.0
    sfmax: float = 0.0
    sfmean: float = 0.0

    def key(self) -> str:
        """symfunctionstring, hdnnps.f90:254-267, padded to 256 characters."""
        if self.name == 2:
            k = (_fmt_i(2, 2) + _fmt_f(self.rcut, 9, 5) + _fmt_f(self.eta, 9, 5)
                 + _fmt_f(self.rs, 9, 5) + _fmt_i(self.els[0], 3))
        else:                 # type 3, and 9 in type 3's form
            k = (_fmt_i(self.name, 2) + _fmt_f(self.rcut, 9, 5) + _fmt_f(self.eta, 9, 5)
                 + _fmt_f(self.zeta, 9, 5) + _fmt_f(self.lam + 10, 9, 5)
                 + _fmt_i(self.els[0], 3) + _fmt_i(self.els[1], 3))
        return k.ljust(256)


@dataclass
class Potential:
    elements: list                       # Z, ascending
    sfs: dict                            # Z -> list[SymFunction] (sorted)
    layers: dict                         # Z -> list of (W (nOut, nIn), b (nOut,))
    acts: list                           # activation char per layer
    nodes: list                          # nOut per layer
    atom_energy: dict = field(default_factory=dict)
    scalesf: bool = False
    centersf: bool = False
    maxrc: float = 0.0
    emin: float = 0.0
    emax: float = 0.0
    cutoff_type: int = 2                 # RuNNer keywords the reference does not read
    scale_range: bool = False
    scale_min: float = 0.0
    scale_max: float = 1.0


def _stable_sort_index(keys):
    return sorted(range(len(keys)), key=lambda i: keys[i])


def read_potential(path: str) -> Potential:
    """parseRuNNerInput, hdnnps.f90:34-233. ``path`` is the directory."""
    path = os.path.join(path, "")
    try:
        with open(path + "input.nn") as f:
            lines = f.read().split("\n")
    except OSError:
        raise RunnerError(path + "input.nn not found")
    nels = 0
    elements = []
    sflists = {}
    atom_energy = {}
    scalesf = centersf = False
    extra = {}
    nlayers = 0
    nnodes, acts = [], []
    maxrc = 0.0
    for raw in lines:
        line = raw[:256]
        w = split_string(line)
        if not w:
            continue
        key = w[0]
        if key == "number_of_elements":
            nels = int(w[1])
        elif key == "elements":
            elements = [w[1 + i][:2] for i in range(nels)]
            for e in elements:
                sflists[elem_sym_to_num(e)] = []
            z = [elem_sym_to_num(e) for e in elements]
            order = _stable_sort_index(z)
            elements = [elements[i] for i in order]
        elif key == "atom_energy":
            atom_energy[elem_sym_to_num(w[1][:2])] = _freal(w[2])
        elif key == "scale_symmetry_functions":
            scalesf = True
        elif key == "center_symmetry_functions":
            centersf = True
        elif key == "cutoff_type":
            extra["cutoff_type"] = int(w[1])
            if not 0 <= extra["cutoff_type"] <= 8:
                raise RunnerError(f"cutoff_type {w[1]} is not one of RuNNer's 0..8")
            if len(w) > 2 and _freal(w[2]) != 0.0:
                raise RunnerError("cutoff_type with an inner cutoff is not supported")
        elif key == "scale_min_short":
            extra["scale_min"], extra["scale_range"] = _freal(w[1]), True
        elif key == "scale_max_short":
            extra["scale_max"], extra["scale_range"] = _freal(w[1]), True
        elif key == "global_hidden_layers_short":
            nlayers = int(w[1])
        elif key == "global_nodes_short":
            nnodes = [int(w[1 + i]) for i in range(nlayers)] + [1]
        elif key == "global_activation_short":
            acts = [w[1 + i][0] for i in range(nlayers + 1)]
        elif key == "symfunction_short":
            name = int(w[2])
            centre = elem_sym_to_num(w[1][:2])
            if name == 2:
                sf = SymFunction(2, (elem_sym_to_num(w[3][:2]),), rcut=_freal(w[6]),
                                 eta=_freal(w[4]), rs=_freal(w[5]))
            elif name in (3, 9):
                sf = SymFunction(name, (elem_sym_to_num(w[3][:2]), elem_sym_to_num(w[4][:2])),
                                 rcut=_freal(w[8]), eta=_freal(w[5]), lam=_freal(w[6]),
                                 zeta=_freal(w[7]))
            else:
                continue
            maxrc = max(maxrc, sf.rcut)
            if centre not in sflists:
                raise RunnerError(f"symmetry function for element {w[1]} not in 'elements'")
            if len(sflists[centre]) + 1 > NMAX_SFS:
                raise RunnerError("more symmetry functions in input file than nmaxSFs allows")
            sflists[centre].append(sf)
    zs = [elem_sym_to_num(e) for e in elements]
    for z in zs:                                             # :204-212
        lst = sflists[z]
        order = _stable_sort_index([s.key() for s in lst])
        sflists[z] = [lst[i] for i in order]
    try:                                                      # readScaling :269-300
        with open(path + "scaling.data") as f:
            toks_lines = [ln.split() for ln in f.read().split("\n")]
    except OSError:
        raise RunnerError(path + "scaling.data not found")
    li = 0

    def record(n):
        """one list-directed READ of n values: it takes values from as many
        lines as needed; the next READ starts on the following line"""
        nonlocal li
        out = []
        while len(out) < n:
            if li >= len(toks_lines):
                raise RunnerError("unexpected end of scaling.data")
            out.extend(toks_lines[li][:n - len(out)])
            li += 1
        return out

    for iel, z in enumerate(zs, start=1):
        for isf, sf in enumerate(sflists[z], start=1):
            t = record(5)
            riel, risf = int(t[0]), int(t[1])
            sf.sfmin, sf.sfmax, sf.sfmean = _freal(t[2]), _freal(t[3]), _freal(t[4])
            if riel != iel or risf != isf:
                raise RunnerError("wrong index reading scaling.data")
    t = record(2)
    emin, emax = _freal(t[0]), _freal(t[1])
    layers = {}
    for z in zs:                                              # readWeights :302-330
        fn = path + f"weights.{z:03d}.data"
        try:
            with open(fn) as f:
                vals = [_freal(ln.split()[0]) for ln in f if ln.split()]
        except OSError:
            raise RunnerError(fn + " not found")
        pos = 0
        nin = len(sflists[z])
        ls = []
        for nout in nnodes:
            W = np.array(vals[pos:pos + nin * nout]).reshape(nin, nout).T.copy()
            pos += nin * nout
            b = np.array(vals[pos:pos + nout])
            pos += nout
            ls.append((W, b))
            nin = nout
        layers[z] = ls
    if extra.get("scale_range") and not extra.get("scale_max", 1.0) > extra.get("scale_min", 0.0):
        raise RunnerError("scale_max_short must be greater than scale_min_short")
    return Potential(zs, {z: sflists[z] for z in zs}, layers, acts, nnodes, atom_energy,
                     scalesf, centersf, maxrc, emin, emax, **extra)


# ---------------------------------------------------------------------------
# ctypes binding of liboracle.so
# ---------------------------------------------------------------------------

MAXEL, MAXLAYERS = 16, 8
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class _Model(ctypes.Structure):
    _fields_ = [
        ("nel", ctypes.c_int), ("Z", ctypes.c_int * MAXEL), ("nsf", ctypes.c_int * MAXEL),
        ("sf_off", ctypes.c_int * MAXEL),
        ("sf_type", _ip), ("sf_e1", _ip), ("sf_e2", _ip),
        ("sf_rcut", _dp), ("sf_eta", _dp), ("sf_rs", _dp), ("sf_lambda", _dp), ("sf_zeta", _dp),
        ("sf_min", _dp), ("sf_max", _dp), ("sf_mean", _dp),
        ("scalesf", ctypes.c_int), ("centersf", ctypes.c_int),
        ("atom_energy", ctypes.c_double * MAXEL), ("maxrc", ctypes.c_double),
        ("nlayers", ctypes.c_int), ("nodes", ctypes.c_int * MAXLAYERS),
        ("act", ctypes.c_char * MAXLAYERS), ("weights", _dp * MAXEL),
        ("cutoff_type", ctypes.c_int), ("scale_range", ctypes.c_int),
        ("scale_min", ctypes.c_double), ("scale_max", ctypes.c_double),
    ]


_LIB = None


def load_library():
    global _LIB
    if _LIB is None:
        p = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(p):
            raise RunnerError(f"{p} is not built (make -C oracle oracle)")
        lib = ctypes.CDLL(p)
        lib.orc_evaluate.restype = ctypes.c_int
        lib.orc_atoms.restype = ctypes.c_int
        lib.orc_neighbor_entries.restype = ctypes.c_long
        lib.orc_neighbor_entries.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_double,
                                             ctypes.c_long, _ip]
        lib.orc_cutoff.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp]
        _LIB = lib
    return _LIB


def _ptr(a, t):
    return a.ctypes.data_as(t)


class Oracle:
    """FP64 CPU restatement of hdnnpEnergyAndForces for one potential."""

    def __init__(self, pot: Potential):
        self.pot = pot
        self.lib = load_library()
        m = _Model()
        zs = pot.elements
        if len(zs) > MAXEL or len(pot.nodes) > MAXLAYERS:
            raise RunnerError("potential too large for the oracle")
        allsf = [sf for z in zs for sf in pot.sfs[z]]
        self._keep = []

        def arr(vals, dt):
            a = np.ascontiguousarray(np.array(vals if vals else [0], dtype=dt))
            self._keep.append(a)
            return a

        m.nel = len(zs)
        off = 0
        for e, z in enumerate(zs):
            m.Z[e] = z
            m.nsf[e] = len(pot.sfs[z])
            m.sf_off[e] = off
            off += len(pot.sfs[z])
            m.atom_energy[e] = pot.atom_energy.get(z, 0.0)
        m.sf_type = _ptr(arr([s.name for s in allsf], np.int32), _ip)
        m.sf_e1 = _ptr(arr([s.els[0] for s in allsf], np.int32), _ip)
        m.sf_e2 = _ptr(arr([s.els[1] if s.name != 2 else 0 for s in allsf], np.int32), _ip)
        for name, attr in (("sf_rcut", "rcut"), ("sf_eta", "eta"), ("sf_rs", "rs"),
                           ("sf_lambda", "lam"), ("sf_zeta", "zeta"), ("sf_min", "sfmin"),
                           ("sf_max", "sfmax"), ("sf_mean", "sfmean")):
            setattr(m, name, _ptr(arr([getattr(s, attr) for s in allsf], np.float64), _dp))
        m.scalesf = int(pot.scalesf)
        m.centersf = int(pot.centersf)
        m.cutoff_type = int(pot.cutoff_type)
        m.scale_range = int(pot.scale_range)
        m.scale_min = float(pot.scale_min)
        m.scale_max = float(pot.scale_max)
        m.maxrc = pot.maxrc
        m.nlayers = len(pot.nodes)
        for l, n in enumerate(pot.nodes):
            m.nodes[l] = n
        m.act = "".join(pot.acts).encode()
        for e, z in enumerate(zs):
            flat = []
            for W, b in pot.layers[z]:
                flat.extend(W.reshape(-1, order="F").tolist())
                flat.extend(b.tolist())
            m.weights[e] = _ptr(arr(flat, np.float64), _dp)
        self.model = m
        self.maxsf = max(len(pot.sfs[z]) for z in zs)

    def evaluate(self, s, want_sf=False, nthreads=0):
        pos = np.ascontiguousarray(s.pos, dtype=np.float64)
        Z = np.ascontiguousarray(s.Z, dtype=np.int32)
        lat = np.ascontiguousarray(s.lat, dtype=np.float64)
        nat = Z.shape[0]
        E = ctypes.c_double(0.0)
        F = np.zeros((nat, 3))
        L = np.zeros((3, 3))
        G = np.zeros((nat, self.maxsf)) if want_sf else None
        D = np.zeros((nat, self.maxsf)) if want_sf else None
        rc = self.lib.orc_evaluate(
            ctypes.byref(self.model), ctypes.c_int(nat), _ptr(Z, _ip), _ptr(pos, _dp),
            ctypes.c_int(int(s.periodic)), _ptr(lat, _dp), ctypes.byref(E), _ptr(F, _dp),
            _ptr(L, _dp), _ptr(G, _dp) if want_sf else None, _ptr(D, _dp) if want_sf else None,
            ctypes.c_int(self.maxsf), ctypes.c_int(nthreads))
        if rc:
            raise RunnerError(f"orc_evaluate failed with status {rc}")
        out = {"energy": E.value, "force": F, "dEdLat": L}
        if want_sf:
            out["G"], out["dEdG"] = G, D
        return out

    def atoms(self, s, sel):
        pos = np.ascontiguousarray(s.pos, dtype=np.float64)
        Z = np.ascontiguousarray(s.Z, dtype=np.int32)
        lat = np.ascontiguousarray(s.lat, dtype=np.float64)
        sel = np.ascontiguousarray(sel, dtype=np.int32)
        n = sel.shape[0]
        Ei = np.zeros(n)
        G = np.zeros((n, self.maxsf))
        D = np.zeros((n, self.maxsf))
        rc = self.lib.orc_atoms(ctypes.byref(self.model), ctypes.c_int(Z.shape[0]), _ptr(Z, _ip),
                                _ptr(pos, _dp), ctypes.c_int(int(s.periodic)), _ptr(lat, _dp),
                                ctypes.c_int(n), _ptr(sel, _ip), _ptr(Ei, _dp), _ptr(G, _dp),
                                _ptr(D, _dp), ctypes.c_int(self.maxsf))
        if rc:
            raise RunnerError(f"orc_atoms failed with status {rc}")
        return Ei, G, D


def committee(members, s):
    """globalHdnnpEnergyAndForces (src/potentials/globalHdnnps.f90:31-70): the
    neighbour lists and scaled symmetry functions of the first potential
    (:44-46), then hdnnpCalcNNs of every member on them (:47-49; only the
    member's networks and atom energies enter, hdnnps.f90:523-526); returns the
    mean energy, forces and dEdLat, the population standard deviation of the
    member energies (:57-67) and the member energies."""
    first = members[0]
    res = []
    for p in members:
        hybrid = dataclasses.replace(first, layers=p.layers, atom_energy=p.atom_energy, acts=p.acts,
                                     nodes=p.nodes)
        res.append(Oracle(hybrid).evaluate(s))
    E = np.array([r["energy"] for r in res])
    mean = E.sum() / len(E)
    return {"energy": mean, "force": sum(r["force"] for r in res) / len(res),
            "dEdLat": sum(r["dEdLat"] for r in res) / len(res),
            "std": float(np.sqrt(((E - mean) ** 2).sum() / len(E))), "members": E}


def neighbor_entries(s, rc: float) -> np.ndarray:
    """(i, j, Lx, Ly, Lz) rows in the reference's list order (self excluded)."""
    lib = load_library()
    pos = np.ascontiguousarray(s.pos, dtype=np.float64)
    lat = np.ascontiguousarray(s.lat, dtype=np.float64)
    nat = pos.shape[0]
    cap = max(64, nat * 64)
    while True:
        ent = np.zeros((cap, 5), dtype=np.int32)
        n = lib.orc_neighbor_entries(nat, _ptr(pos, _dp), int(s.periodic), _ptr(lat, _dp), rc,
                                     cap, _ptr(ent, _ip))
        if n >= 0:
            return ent[:n]
        cap = -n


def canonical(ent: np.ndarray) -> np.ndarray:
    """Neighbour entries sorted lexicographically by (i, j, Lx, Ly, Lz)."""
    ent = np.asarray(ent, dtype=np.int32).reshape(-1, 5)
    order = np.lexsort(ent.T[::-1])
    return np.ascontiguousarray(ent[order])


def cutoff(r: float, rc: float, cutoff_type: int = 2):
    lib = load_library()
    fc, dfc = ctypes.c_double(), ctypes.c_double()
    lib.orc_cutoff(int(cutoff_type), r, rc, ctypes.byref(fc), ctypes.byref(dfc))
    return fc.value, dfc.value
