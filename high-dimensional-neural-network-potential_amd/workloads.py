"""Synthetic structures and the synthetic 5-element RuNNer potential.

Everything here is seeded and deterministic (numpy PCG64), so the same call
gives the same bytes on any machine.

Structures follow the reference's ``atStruct`` conventions
(src/atoms/atomicStructure.f90:17-28): positions ``pos[nat, 3]`` in bohr
(memory order x1 y1 z1 x2 ...), ``lat[3, 3]`` with row k = lattice vector k
(the reference's column ``lat(:, k)``), ``Z[nat]`` int32 atomic numbers.

Perovskite supercells replicate the unit cell in the order of
``as_repeatUnitCell`` (atomicStructure.f90:636-671): x outermost, then y,
then z, atoms of the unit innermost.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

A_PEROVSKITE = 11.9          # bohr
RC = 12.0                    # bohr, cutoff of every symmetry function
ELEMENTS = (1, 6, 7, 53, 82)  # H C N I Pb, sorted by Z
SYMBOL = {1: "H", 6: "C", 7: "N", 53: "I", 82: "Pb"}
# atomic masses used by the MD harness (constants.f90:38-53), amu
MASS = {1: 1.008, 6: 12.011, 7: 14.007, 53: 126.904, 82: 207.2}
MASS_UNIT_TO_ELECTRON = 1822.888513874368        # constants.f90:11
FEMTOSECONDS_TO_ATOMIC = 41.34137333519408      # constants.f90:12

# methylammonium geometry of the reference example (example/input.xyz), bohr;
# the C-N bond lies on x and its midpoint is the origin
_MA_Z = np.array([7, 6, 1, 1, 1, 1, 1, 1], dtype=np.int32)
_MA_POS = np.array([
    [0.139424E+01, 0.000000E+00, 0.000000E+00],
    [-0.139424E+01, 0.000000E+00, 0.000000E+00],
    [0.210232E+01, -0.595831E+00, -0.170907E+01],
    [0.210251E+01, -0.118221E+01, 0.137043E+01],
    [0.210232E+01, 0.177804E+01, 0.338450E+00],
    [-0.202541E+01, -0.192941E+01, -0.367174E+00],
    [-0.202522E+01, 0.646664E+00, 0.185458E+01],
    [-0.202522E+01, 0.128275E+01, -0.148721E+01],
])


@dataclass
class Structure:
    """Host-side mirror of the reference ``atStruct`` (atomicStructure.f90:17-28)."""
    pos: np.ndarray                  # (nat, 3) float64, bohr
    Z: np.ndarray                    # (nat,) int32
    lat: np.ndarray = field(default_factory=lambda: np.zeros((3, 3)))  # row k = vector k
    periodic: bool = False

    @property
    def nat(self) -> int:
        return int(self.Z.shape[0])

    def copy(self) -> "Structure":
        return Structure(self.pos.copy(), self.Z.copy(), self.lat.copy(), self.periodic)


def _random_rotation(rng: np.random.Generator) -> np.ndarray:
    """Uniform random rotation matrix (Shoemake's quaternion method)."""
    u1, u2, u3 = rng.random(3)
    q = np.array([
        np.sqrt(1 - u1) * np.sin(2 * np.pi * u2),
        np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
        np.sqrt(u1) * np.sin(2 * np.pi * u3),
        np.sqrt(u1) * np.cos(2 * np.pi * u3),
    ])
    x, y, z, w = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def perovskite(n1: int, n2: int, n3: int, seed: int = 1234, sigma: float = 0.2,
               a: float = A_PEROVSKITE) -> Structure:
    """Pseudocubic ABX3 supercell, 12 atoms per formula unit.

    Unit cell: Pb at (0,0,0); I at (a/2,0,0), (0,a/2,0), (0,0,a/2);
    CH3NH3+ at the body centre with a seeded uniform random orientation per
    unit cell. Every coordinate then gets N(0, sigma^2) jitter.
    """
    rng = np.random.default_rng(seed)
    h = a / 2
    frame_z = np.array([82, 53, 53, 53], dtype=np.int32)
    frame_pos = np.array([[0, 0, 0], [h, 0, 0], [0, h, 0], [0, 0, h]], dtype=np.float64)
    unit_z = np.concatenate([frame_z, _MA_Z])
    pos = []
    for x in range(n1):
        for y in range(n2):
            for z in range(n3):
                origin = np.array([x * a, y * a, z * a])
                rot = _random_rotation(rng)
                mol = _MA_POS @ rot.T + np.array([h, h, h])
                pos.append(np.concatenate([frame_pos, mol]) + origin)
    pos = np.concatenate(pos)
    pos = pos + rng.normal(0.0, sigma, size=pos.shape)
    Z = np.tile(unit_z, n1 * n2 * n3).astype(np.int32)
    lat = np.diag([n1 * a, n2 * a, n3 * a]).astype(np.float64)
    return Structure(pos, Z, lat, True)


def cluster(nat: int, seed: int, box: float = 20.0, dmin: float = 1.8,
            probs=(0.5, 0.25, 0.25)) -> Structure:
    """Nonperiodic C/N/H cluster: uniform positions in a cube of side ``box``
    with every pair at least ``dmin`` apart (rejection sampling); elements
    H, C, N drawn with probabilities ``probs``."""
    rng = np.random.default_rng(seed)
    pos = np.zeros((0, 3))
    while pos.shape[0] < nat:
        p = rng.random(3) * box
        if pos.shape[0] == 0 or np.min(np.sum((pos - p) ** 2, axis=1)) >= dmin * dmin:
            pos = np.vstack([pos, p])
    Z = rng.choice(np.array([1, 6, 7], dtype=np.int32), size=nat, p=probs).astype(np.int32)
    return Structure(pos, Z, np.zeros((3, 3)), False)


def cluster_scaled(nat: int, seed: int, dmin: float = 1.8, probs=(0.5, 0.25, 0.25),
                   ref_nat: int = 36, ref_box: float = 20.0) -> Structure:
    """Nonperiodic C/N/H cluster of any size: the rejection recipe of
    ``cluster`` in a cube of side ref_box * max(1, nat / ref_nat)^(1/3), so a
    large cluster keeps the density of a ref_nat-atom one in the 20-bohr box.
    Candidates are checked against a hash grid of cell size dmin."""
    rng = np.random.default_rng(seed)
    box = ref_box * max(1.0, nat / ref_nat) ** (1.0 / 3.0)
    ncell = max(1, int(box / dmin))
    cell = box / ncell
    grid: dict = {}
    pos = np.zeros((nat, 3))
    n = 0
    d2 = dmin * dmin
    while n < nat:
        cand = rng.random((256, 3)) * box
        for p in cand:
            c = np.minimum((p / cell).astype(int), ncell - 1)
            ok = True
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dz in (-1, 0, 1):
                        for q in grid.get((c[0] + dx, c[1] + dy, c[2] + dz), ()):
                            if np.sum((pos[q] - p) ** 2) < d2:
                                ok = False
                                break
                        if not ok:
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if ok:
                pos[n] = p
                grid.setdefault((c[0], c[1], c[2]), []).append(n)
                n += 1
                if n == nat:
                    break
    Z = rng.choice(np.array([1, 6, 7], dtype=np.int32), size=nat, p=probs).astype(np.int32)
    return Structure(pos, Z, np.zeros((3, 3)), False)


def perovskite_near(nat: int, seed: int) -> Structure:
    """Perovskite supercell n1 x n2 x n3 of about nat atoms (12 per formula
    unit, at least one), as near a cube as the count allows."""
    nfu = max(1, int(round(nat / 12)))
    n1 = max(1, int(round(nfu ** (1.0 / 3.0))))
    n2 = max(1, int(round((nfu / n1) ** 0.5)))
    n3 = max(1, int(round(nfu / (n1 * n2))))
    return perovskite(n1, n2, n3, seed=seed)


def serving_requests(nreq: int, seed: int = 77, nmin: int = 8, nmax: int = 20000):
    """Request structures of a serving stream: atom counts log-uniform in
    nmin..nmax, alternately periodic perovskite supercells (perovskite_near)
    and nonperiodic C/N/H clusters (cluster_scaled)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(nreq):
        n = int(round(np.exp(rng.uniform(np.log(nmin), np.log(nmax)))))
        sub = int(rng.integers(1 << 30))
        out.append(perovskite_near(n, sub) if k % 2 == 0 else cluster_scaled(n, sub))
    return out


def batch_mix(nstruct: int, seed: int = 2024, nmax_cell: int = 4, cl_min: int = 8, cl_max: int = 64):
    """Batch of independent structures: half perovskite supercells
    n1 x n2 x n3 with each n uniform in 1..nmax_cell (12 to 12*nmax_cell^3
    atoms, jitter 0.2 bohr), half nonperiodic C/N/H clusters of cl_min..cl_max
    atoms in a 20-bohr box (minimum distance 1.8 bohr), interleaved."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(nstruct):
        sub = int(rng.integers(1 << 30))
        if k % 2 == 0:
            n = rng.integers(1, nmax_cell + 1, size=3)
            out.append(perovskite(int(n[0]), int(n[1]), int(n[2]), seed=sub))
        else:
            out.append(cluster(int(rng.integers(cl_min, cl_max + 1)), seed=sub))
    return out


def methylammonium() -> Structure:
    """The reference example molecule (example/input.xyz), nonperiodic."""
    return Structure(_MA_POS.copy(), _MA_Z.copy(), np.zeros((3, 3)), False)


def write_structure(path: str, s: Structure) -> None:
    """Plain-text structure: 'nat periodic', three lattice-vector lines,
    then 'Z x y z' per atom (full double precision)."""
    with open(path, "w") as f:
        f.write(f"{s.nat} {1 if s.periodic else 0}\n")
        for k in range(3):
            f.write(" ".join(repr(float(v)) for v in s.lat[k]) + "\n")
        for z, p in zip(s.Z, s.pos):
            f.write(f"{int(z)} {float(p[0])!r} {float(p[1])!r} {float(p[2])!r}\n")


def read_structure(path: str) -> Structure:
    with open(path) as f:
        nat, per = (int(t) for t in f.readline().split())
        lat = np.array([[float(t) for t in f.readline().split()] for _ in range(3)])
        rows = [f.readline().split() for _ in range(nat)]
    Z = np.array([int(r[0]) for r in rows], dtype=np.int32)
    pos = np.array([[float(t) for t in r[1:4]] for r in rows])
    return Structure(pos, Z, lat, bool(per))


# ---------------------------------------------------------------------------
# synthetic potential (RuNNer files)
# ---------------------------------------------------------------------------

G2_ETAS = [float(f"{v:.6g}") for v in np.geomspace(1e-3, 0.5, 8)]
G4_ETAS = [float(f"{v:.6g}") for v in np.geomspace(1e-3, 1e-1, 5)]
G4_ZETAS = [1.0, 2.0, 4.0, 16.0]
G4_LAMBDAS = [-1.0, 1.0]
HIDDEN = (25, 25)
ACTIVATIONS = ("t", "t", "l")


def element_pairs(elements=ELEMENTS):
    """The 15 unordered element pairs of 5 elements, sorted by Z."""
    out = []
    for i, a in enumerate(elements):
        for b in elements[i:]:
            out.append((a, b))
    return out


def symfunction_lines(elements=ELEMENTS, rc: float = RC):
    """RuNNer ``symfunction_short`` lines: per centre element 40 G2
    (8 eta x 5 neighbour elements, Rs = 0) and 40 G4 (5 eta x 4 zeta x
    2 lambda, in that nesting); G4 number s uses element pair s mod 15."""
    pairs = element_pairs(elements)
    lines = []
    for c in elements:
        cs = SYMBOL[c]
        for eta in G2_ETAS:
            for e in elements:
                lines.append(f"symfunction_short {cs:>2} 2 {SYMBOL[e]:>2} {eta:.6g} 0.0 {rc:.4f}")
        s = 0
        for eta in G4_ETAS:
            for zeta in G4_ZETAS:
                for lam in G4_LAMBDAS:
                    a, b = pairs[s % len(pairs)]
                    lines.append(f"symfunction_short {cs:>2} 3 {SYMBOL[a]:>2} {SYMBOL[b]:>2} "
                                 f"{eta:.6g} {lam:+.1f} {zeta:.1f} {rc:.4f}")
                    s += 1
    return lines


def write_input_nn(path: str, elements=ELEMENTS, rc: float = RC, scale: bool = True,
                   cutoff_type: int = 2) -> None:
    with open(path, "w") as f:
        f.write("### synthetic perovskite-like HDNNP (H C N I Pb), generated\n")
        f.write(f"number_of_elements {len(elements)}\n")
        f.write("elements " + " ".join(SYMBOL[e] for e in elements) + "\n")
        f.write(f"global_hidden_layers_short {len(HIDDEN)}\n")
        f.write("global_nodes_short " + " ".join(str(n) for n in HIDDEN) + "\n")
        f.write("global_activation_short " + " ".join(ACTIVATIONS) + "\n")
        f.write(f"cutoff_type {cutoff_type}\n")
        for line in symfunction_lines(elements, rc):
            f.write(line + "\n")
        if scale:
            f.write("scale_symmetry_functions\ncenter_symmetry_functions\n")


def write_weights(dirpath: str, nsf: int, seed: int = 4321, elements=ELEMENTS) -> None:
    """weights.ZZZ.data per element: per layer, for each input j, for each
    output i, W(i, j); then the nOut biases (hdnnps.f90:318-327).
    W ~ N(0, 1/fan_in), biases 0."""
    rng = np.random.default_rng(seed)
    sizes = [nsf, *HIDDEN, 1]
    for z in elements:
        vals = []
        for l in range(len(sizes) - 1):
            nin, nout = sizes[l], sizes[l + 1]
            w = rng.normal(0.0, 1.0 / np.sqrt(nin), size=(nout, nin))
            vals.extend(w.T.reshape(-1).tolist())     # input-major
            vals.extend([0.0] * nout)
        with open(os.path.join(dirpath, f"weights.{z:03d}.data"), "w") as f:
            for v in vals:
                f.write(f"{float(v)!r}\n")


def write_weights_repulsive(dirpath: str, nsf: int, ng2: int, out_scale: float = 1.0, seed: int = 4322,
                            elements=ELEMENTS) -> None:
    """weights.ZZZ.data of a purely repulsive network of the same shape as
    write_weights: the first layer reads only the ng2 radial inputs (the G2
    come first in the sorted order), every weight is positive (uniform in
    [0.5, 1.5] times 0.06, 1/fan_in and out_scale/fan_in per layer) and the
    biases are 0, so each atomic energy increases with every G2 -- that is,
    as any neighbour comes closer. A soft-sphere model on which velocity
    Verlet conserves energy; its cost per atom is that of the random one."""
    rng = np.random.default_rng(seed)
    sizes = [nsf, *HIDDEN, 1]
    gains = [0.06, 1.0 / sizes[1], out_scale / sizes[2]]
    for z in elements:
        vals = []
        for l in range(len(sizes) - 1):
            nin, nout = sizes[l], sizes[l + 1]
            w = gains[l] * rng.uniform(0.5, 1.5, size=(nout, nin))
            if l == 0:
                w[:, ng2:] = 0.0
            vals.extend(w.T.reshape(-1).tolist())     # input-major
            vals.extend([0.0] * nout)
        with open(os.path.join(dirpath, f"weights.{z:03d}.data"), "w") as f:
            for v in vals:
                f.write(f"{float(v)!r}\n")


def write_scaling(path: str, stats, energy_range=(-1.0, 1.0)) -> None:
    """scaling.data: 'iel isf min max mean' per (Z-rank, sorted-SF rank),
    then 'emin emax' (hdnnps.f90:269-300). ``stats`` is a list over elements
    of arrays (nsf, 3) = (min, max, mean)."""
    with open(path, "w") as f:
        for iel, st in enumerate(stats, start=1):
            for isf, (mn, mx, me) in enumerate(st, start=1):
                f.write(f"{iel:4d} {isf:4d} {float(mn)!r} {float(mx)!r} {float(me)!r}\n")
        f.write(f"{float(energy_range[0])!r} {float(energy_range[1])!r}\n")
