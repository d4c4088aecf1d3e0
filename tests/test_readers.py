"""Structure readers (atomicStructure.f90:143-475) against what the
reference's own readers made of the same files (tests/golden/readers.npz,
written by refdump read), in Python and in the Fortran atomicStructure
module. No GPU."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from conftest import GOLD, ROOT

CASES = [("fhiaims", "geometry.in"), ("fhiaims", "molecule.in"), ("ascii", "cell_ang.ascii"),
         ("ascii", "cell_bohr.ascii"), ("asciiang", "cell_bohr.ascii"), ("poscar", "POSCAR"),
         ("data", "input.data"), ("data", "big.data")]


def golden_structures(fmt, fn):
    d = np.load(os.path.join(GOLD, "readers.npz"))
    out, k = [], 0
    while f"{fmt}:{fn}:{k}:Z" in d.files:
        key = f"{fmt}:{fn}:{k}:"
        out.append({f: d[key + f] for f in ("periodic", "lat", "Z", "pos", "energy", "qtot", "charges", "forces")
                    if key + f in d.files})
        k += 1
    return out


def py_read(pkg, fmt, path):
    rd = pkg.readers
    if fmt == "asciiang":
        return [rd.read_ascii(path, is_angstrom=True)]
    return rd.read_structures(path, fmt)


@pytest.mark.parametrize("fmt,fn", CASES)
def test_python_reader_matches_reference(pkg, fmt, fn):
    got = py_read(pkg, fmt, os.path.join(GOLD, "readers", fn))
    want = golden_structures(fmt, fn)
    assert len(got) == len(want) and want
    for s, w in zip(got, want):
        assert bool(s.periodic) == bool(w["periodic"])
        assert np.array_equal(s.Z, w["Z"])
        np.testing.assert_array_equal(s.pos, w["pos"])
        np.testing.assert_array_equal(s.lat if s.periodic else np.zeros((3, 3)), w["lat"])
        if "energy" in w:
            assert s.energy == w["energy"] and s.qtot == w["qtot"]
            np.testing.assert_array_equal(s.charges, w["charges"])
            np.testing.assert_array_equal(s.forces, w["forces"])


def test_reader_errors(pkg, tmp_path):
    rd = pkg.readers
    p = tmp_path / "POSCAR"
    p.write_text("x\n1.0\n1 0 0\n0 1 0\n0 0 1\nH\n1\nCartesian\n0 0 0\n")
    with pytest.raises(rd.ReadError):
        rd.read_poscar(str(p))
    q = tmp_path / "a.in"
    q.write_text("".join("lattice_vector 1 0 0\n" for _ in range(4)))
    with pytest.raises(rd.ReadError):
        rd.read_fhiaims(str(q))
    with pytest.raises(rd.ReadError):
        rd.read_structures(str(q), "cif")
    assert rd.element_number("Pb") == 82 and rd.element_number("Xx") == -1 and rd.element_number("PB") == -1


def parse_dump(path):
    out, cur = [], None
    for ln in open(path).read().splitlines():
        p = ln.split()
        if not p:
            continue
        if p[0] == "structure":
            cur = {"periodic": int(p[2]), "lat": [], "Z": [], "pos": [], "aux": []}
            out.append(cur)
        elif p[0] == "atom":
            cur["Z"].append(int(p[1]))
            cur["pos"].append([float(t) for t in p[2:5]])
        elif p[0] == "energy":
            cur["energy"], cur["qtot"] = float(p[1]), float(p[2])
        elif p[0] == "aux":
            cur["aux"].append([float(t) for t in p[1:5]])
        else:
            cur["lat"].append([float(t) for t in p[:3]])
    return out


@pytest.mark.parametrize("fmt,fn", CASES)
def test_fortran_reader_matches_reference(fmt, fn, tmp_path):
    """The readers of the Fortran atomicStructure module (driver mode
    `hdnnp_md read`), written in the same text form as refdump read."""
    exe = os.path.join(ROOT, "high-dimensional-neural-network-potential_amd", "lib", "hdnnp_md")
    out = tmp_path / "r.txt"
    subprocess.run([exe, "read", fmt, os.path.join(GOLD, "readers", fn), str(out)], check=True,
                   capture_output=True)
    got = parse_dump(str(out))
    want = golden_structures(fmt, fn)
    assert len(got) == len(want)
    for s, w in zip(got, want):
        assert s["periodic"] == int(w["periodic"])
        assert np.array_equal(np.array(s["Z"]), w["Z"])
        np.testing.assert_array_equal(np.array(s["pos"]).reshape(-1, 3), w["pos"])
        np.testing.assert_array_equal(np.array(s["lat"]), w["lat"])
        if "energy" in w:
            assert s["energy"] == w["energy"] and s["qtot"] == w["qtot"]
            np.testing.assert_array_equal(np.array(s["aux"])[:, 0], w["charges"])
            np.testing.assert_array_equal(np.array(s["aux"])[:, 1:], w["forces"])


@pytest.mark.parametrize("case", ["perov1", "ma_tric2"])
def test_fortran_repeat_unit_cell_matches_reference(wl, case, tmp_path):
    """as_repeatUnitCell (atomicStructure.f90:636-671) of the Fortran
    atomicStructure module (driver mode `hdnnp_md repeat`) against the
    reference's own routine (tests/golden/repeat.npz, written by refdump
    repeat): atom order (x, y, z repetitions, atoms innermost), positions and
    lattice bit for bit."""
    d = np.load(os.path.join(GOLD, "repeat.npz"))
    n = [int(k) for k in d[f"{case}_n"]]
    s = wl.Structure(d[f"{case}_in_pos"], d[f"{case}_in_Z"].astype(np.int32), d[f"{case}_in_lat"], True)
    sf, out = tmp_path / "s.txt", tmp_path / "o.txt"
    wl.write_structure(str(sf), s)
    exe = os.path.join(ROOT, "high-dimensional-neural-network-potential_amd", "lib", "hdnnp_md")
    subprocess.run([exe, "repeat", str(sf), *map(str, n), str(out)], check=True, capture_output=True)
    got = parse_dump(str(out))[0]
    assert np.array_equal(np.array(got["Z"]), d[f"{case}_Z"])
    np.testing.assert_array_equal(np.array(got["pos"]).reshape(-1, 3), d[f"{case}_pos"])
    np.testing.assert_array_equal(np.array(got["lat"]), d[f"{case}_lat"])
    assert len(got["Z"]) == s.nat * n[0] * n[1] * n[2]
