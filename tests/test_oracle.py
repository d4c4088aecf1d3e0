"""Pins the oracle (oracle/hdnnp_oracle.c + oracle/runner.py) before it is
trusted as the checker: against the golden vectors written by the compiled
reference, and against what the mathematics fixes (closed forms, brute force,
finite differences, invariants). Also plants mistakes in the oracle to show
that these checks catch them. CPU only."""
from __future__ import annotations

import copy
import hashlib
import itertools
import os

import numpy as np
import pytest

from conftest import EVAL_CASES, POTENTIALS, golden, golden_structure


# ---------------------------------------------------------------- parser
@pytest.mark.parametrize("name", ["methylammonium", "perovskite"])
def test_parser_matches_reference_table(oracle, name):
    """read_potential == the reference's parsed handle (sorted SF order,
    scaling rows, weights, flags) -- exact."""
    pot = oracle.read_potential(POTENTIALS[name])
    t = golden("table_" + name)
    assert list(t["Z"]) == pot.elements
    assert list(t["nsf"]) == [len(pot.sfs[z]) for z in pot.elements]
    assert float(t["maxrc"]) == pot.maxrc
    assert bool(t["scalesf"]) == pot.scalesf and bool(t["centersf"]) == pot.centersf
    rows = []
    for z in pot.elements:
        for s in pot.sfs[z]:
            rows.append([s.name, s.els[0], s.els[1] if s.name == 3 else 0, s.rcut, s.eta,
                         s.rs if s.name == 2 else 0.0, s.lam if s.name == 3 else 0.0,
                         s.zeta if s.name == 3 else 0.0, s.sfmin, s.sfmax, s.sfmean])
    np.testing.assert_array_equal(np.array(rows), t["sf"])
    w = np.concatenate([np.concatenate([np.concatenate([W.T.reshape(-1), b]) for W, b in pot.layers[z]])
                        for z in pot.elements])
    np.testing.assert_array_equal(w, t["weights"])
    assert [a for a in str(t["acts"][0])] == pot.acts
    np.testing.assert_array_equal(t["energy_range"], [pot.emin, pot.emax])


# ---------------------------------------------------------------- evaluation
@pytest.mark.parametrize("case", EVAL_CASES)
def test_oracle_matches_reference(oracles, wl, case):
    d = golden("eval_" + case)
    s = golden_structure(wl, d)
    o = oracles[str(d["potential"])]
    r = o.evaluate(s, want_sf="G" in d, nthreads=1)
    E = float(d["energy"])
    assert abs(r["energy"] - E) <= 1e-12 * max(1.0, abs(E))
    np.testing.assert_allclose(r["force"], d["force"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(r["dEdLat"], d["dEdLat"], rtol=0,
                               atol=1e-12 * max(1.0, np.abs(d["dEdLat"]).max()) * s.nat ** 0.5)
    if "G" in d:
        w = d["G"].shape[1]
        np.testing.assert_allclose(r["G"][:, :w], d["G"], rtol=0, atol=1e-13)
        np.testing.assert_allclose(r["dEdG"][:, :w], d["dEdG"], rtol=0, atol=1e-13)


@pytest.mark.parametrize("case", EVAL_CASES)
def test_oracle_neighbors_bitexact(oracle, wl, case):
    d = golden("eval_" + case)
    s = golden_structure(wl, d)
    pot = POTENTIALS[str(d["potential"])]
    rc = oracle.read_potential(pot).maxrc
    ent = oracle.canonical(oracle.neighbor_entries(s, rc))
    assert ent.shape[0] == int(d["nneigh"])
    assert hashlib.sha256(ent.astype("<i4").tobytes()).hexdigest() == str(d["neighbors_sha256"])
    if "neighbors" in d:
        np.testing.assert_array_equal(ent, d["neighbors"])


def _wrap(pos, lat):
    inv = np.linalg.inv(lat.T)
    s = pos @ inv.T
    s = s - np.floor(s)
    return s @ lat


def _brute_neighbors(s, rc):
    """All (i, j, L) with |R_i - R_j - L.h| < rc over enough images (textbook
    all-pairs search, cf. the reference's unused buildNeighborList)."""
    if s.periodic:
        pos = _wrap(_wrap(s.pos, s.lat), s.lat)
        h = s.lat
        # image range from the plane spacings
        vol = abs(np.linalg.det(h))
        spans = [vol / np.linalg.norm(np.cross(h[(k + 1) % 3], h[(k + 2) % 3])) for k in range(3)]
        n = [int(np.ceil(rc / w)) + 1 for w in spans]
        images = list(itertools.product(*[range(-k, k + 1) for k in n]))
    else:
        pos = s.pos
        h = np.zeros((3, 3))
        images = [(0, 0, 0)]
    out, near = [], 0
    for L in images:
        sh = np.array(L) @ h
        d = pos[:, None, :] - (pos[None, :, :] + sh)
        r2 = np.sum(d * d, axis=2)
        near += np.sum(np.abs(r2 - rc * rc) < 1e-8)
        ii, jj = np.nonzero(r2 < rc * rc)
        for i, j in zip(ii, jj):
            if L == (0, 0, 0) and i == j:
                continue
            out.append((i, j, *L))
    return np.array(out, dtype=np.int32).reshape(-1, 5), near


@pytest.mark.parametrize("case", ["ma_cell1", "ma_tric2", "cluster30", "perov1", "perov132"])
def test_oracle_neighbors_bruteforce(oracle, wl, case):
    d = golden("eval_" + case)
    s = golden_structure(wl, d)
    ref, near = _brute_neighbors(s, 12.0)
    assert near == 0, "a pair sits within 1e-8 of the cutoff; the comparison would be ambiguous"
    got = oracle.canonical(oracle.neighbor_entries(s, 12.0))
    np.testing.assert_array_equal(got, oracle.canonical(ref))


# ---------------------------------------------------------------- closed forms
def _write_potential(dirpath, sf_lines, elements=("H",), nodes=(2,), acts=("l", "l"), scale=False):
    os.makedirs(dirpath, exist_ok=True)
    with open(os.path.join(dirpath, "input.nn"), "w") as f:
        f.write(f"number_of_elements {len(elements)}\nelements {' '.join(elements)}\n")
        f.write(f"global_hidden_layers_short {len(nodes)}\n")
        f.write("global_nodes_short " + " ".join(map(str, nodes)) + "\n")
        f.write("global_activation_short " + " ".join(acts) + "\n")
        for ln in sf_lines:
            f.write(ln + "\n")
        if scale:
            f.write("scale_symmetry_functions\ncenter_symmetry_functions\n")
    from conftest import load_oracle
    nsf = {e: sum(1 for ln in sf_lines if ln.split()[1] == e) for e in elements}
    z = {e: load_oracle().elem_sym_to_num(e) for e in elements}
    with open(os.path.join(dirpath, "scaling.data"), "w") as f:
        for i, e in enumerate(sorted(elements, key=lambda x: z[x]), start=1):
            for k in range(1, nsf[e] + 1):
                f.write(f"{i} {k} 0.0 1.0 0.0\n")
        f.write("0.0 1.0\n")
    rng = np.random.default_rng(3)
    for e in elements:
        sizes = [nsf[e], *nodes, 1]
        with open(os.path.join(dirpath, f"weights.{z[e]:03d}.data"), "w") as f:
            for a, b in zip(sizes[:-1], sizes[1:]):
                for v in rng.normal(size=a * b + b):
                    f.write(f"{float(v)!r}\n")


def test_cutoff_closed_form(oracle):
    rc = 12.0
    fc, dfc = oracle.cutoff(rc, rc)
    assert fc == 0.0 and dfc == 0.0
    fc, _ = oracle.cutoff(0.0 + 1e-300, rc)
    assert abs(fc - np.tanh(1.0) ** 3) < 1e-15
    for r in (0.7, 3.3, 8.1, 11.99):
        h = 1e-6
        fp = (oracle.cutoff(r + h, rc)[0] - oracle.cutoff(r - h, rc)[0]) / (2 * h)
        assert abs(oracle.cutoff(r, rc)[1] - fp / r) < 1e-8


def test_g2_two_atoms_closed_form(oracle, wl, tmp_path):
    eta, rs, rc, r = 0.031, 1.5, 12.0, 3.7
    _write_potential(str(tmp_path), [f"symfunction_short H 2 H {eta} {rs} {rc}"])
    o = oracle.Oracle(oracle.read_potential(str(tmp_path)))
    s = wl.Structure(np.array([[0.0, 0, 0], [r, 0, 0]]), np.array([1, 1], np.int32))
    G = o.evaluate(s, want_sf=True)["G"]
    want = np.exp(-eta * (r - rs) ** 2) * np.tanh(1 - r / rc) ** 3
    np.testing.assert_allclose(G[:, 0], [want, want], rtol=1e-14)


@pytest.mark.parametrize("lam,zeta", [(1.0, 1.0), (-1.0, 4.0), (1.0, 16.0), (1.0, 2.5)])
def test_g4_equilateral_closed_form(oracle, wl, tmp_path, lam, zeta):
    eta, rc, r = 0.02, 12.0, 4.1
    _write_potential(str(tmp_path), [f"symfunction_short H 2 H 0.01 0.0 {rc}",
                                     f"symfunction_short H 3 H H {eta} {lam} {zeta} {rc}"])
    o = oracle.Oracle(oracle.read_potential(str(tmp_path)))
    P = np.array([[0.0, 0, 0], [r, 0, 0], [r / 2, r * np.sqrt(3) / 2, 0]])
    s = wl.Structure(P, np.array([1, 1, 1], np.int32))
    G = o.evaluate(s, want_sf=True)["G"][:, 1]
    want = 2 ** (1 - zeta) * (1 + lam / 2) ** zeta * np.exp(-3 * eta * r * r) * np.tanh(1 - r / rc) ** 9
    np.testing.assert_allclose(G, want, rtol=1e-13)


# ---------------------------------------------------------------- derivatives
@pytest.mark.parametrize("case", ["perov1", "ma_tric2", "cluster30"])
def test_forces_are_minus_energy_gradient(oracles, wl, case):
    d = golden("eval_" + case)
    s = golden_structure(wl, d)
    o = oracles[str(d["potential"])]
    F = o.evaluate(s)["force"]
    rng = np.random.default_rng(5)
    h = 1e-5
    for _ in range(6):
        i, k = int(rng.integers(s.nat)), int(rng.integers(3))
        sp, sm = s.copy(), s.copy()
        sp.pos[i, k] += h
        sm.pos[i, k] -= h
        fd = -(o.evaluate(sp)["energy"] - o.evaluate(sm)["energy"]) / (2 * h)
        assert abs(fd - F[i, k]) < 2e-7, (i, k, fd, F[i, k])
    np.testing.assert_allclose(F.sum(axis=0), 0.0, atol=1e-12 * s.nat)


@pytest.mark.parametrize("case", ["perov1", "ma_tric2"])
def test_dEdLat_is_lattice_derivative(oracles, wl, case):
    """dEdLat[k][a] = dE/d(component a of lattice vector k) at fixed
    fractional coordinates."""
    d = golden("eval_" + case)
    s = golden_structure(wl, d)
    o = oracles[str(d["potential"])]
    L = o.evaluate(s)["dEdLat"]
    frac = s.pos @ np.linalg.inv(s.lat)
    h = 1e-5
    for k in range(3):
        for a in range(3):
            vals = []
            for sgn in (1, -1):
                t = s.copy()
                t.lat[k, a] += sgn * h
                t.pos = frac @ t.lat
                vals.append(o.evaluate(t)["energy"])
            fd = (vals[0] - vals[1]) / (2 * h)
            assert abs(fd - L[k, a]) < 2e-7, (k, a, fd, L[k, a])


def test_energy_invariances(oracles, wl):
    o = oracles["methylammonium"]
    s = wl.cluster(20, seed=4)
    E0 = o.evaluate(s)["energy"]
    rng = np.random.default_rng(1)
    R = wl._random_rotation(rng)
    t = s.copy()
    t.pos = s.pos @ R.T + np.array([1.3, -2.2, 0.4])
    assert abs(o.evaluate(t)["energy"] - E0) < 1e-10
    perm = rng.permutation(s.nat)
    u = wl.Structure(s.pos[perm], s.Z[perm])
    assert abs(o.evaluate(u)["energy"] - E0) < 1e-10
    p = golden_structure(wl, golden("eval_perov2"))
    op = oracles["perovskite"]
    Ep = op.evaluate(p)["energy"]
    q = p.copy()
    q.pos = q.pos + np.array([5.1, 0.3, -7.7])
    assert abs(op.evaluate(q)["energy"] - Ep) < 1e-9


# ---------------------------------------------------------------- planted mistakes
def test_planted_mistakes_are_caught(oracle, wl):
    """Plausible slips planted into copies of the oracle's model must break
    the golden comparison: the checks above can fail."""
    pot = oracle.read_potential(POTENTIALS["perovskite"])
    d = golden("eval_perov1")
    s = golden_structure(wl, d)
    E = float(d["energy"])

    def energy_with(mutate):
        p = copy.deepcopy(pot)
        mutate(p)
        return oracle.Oracle(p).evaluate(s)["energy"]

    z = pot.elements[0]
    g4 = [k for k, sf in enumerate(pot.sfs[z]) if sf.name == 3][0]
    mutations = {
        "lambda sign": lambda p: setattr(p.sfs[z][g4], "lam", -p.sfs[z][g4].lam),
        "swapped SF order": lambda p: p.sfs[z].insert(0, p.sfs[z].pop(1)),
        "no centring": lambda p: setattr(p, "centersf", False),
        "transposed layer": lambda p: p.layers[z].__setitem__(
            1, (p.layers[z][1][0].T.copy(), p.layers[z][1][1])),
    }
    for name, mut in mutations.items():
        assert abs(energy_with(mut) - E) > 1e-8, name


# ---------------------------------------------------------------- committee
def test_committee_oracle_pinned(oracle, wl, tmp_path):
    """The committee restatement (globalHdnnps.f90:31-70) against what its
    definition fixes: members that share the first potential's symmetry
    functions give the plain mean of their own evaluations and np.std of the
    energies; a member's own scaling.data is ignored; one member is the
    potential itself."""
    from conftest import committee_member
    base = POTENTIALS["methylammonium"]
    m2 = committee_member(base, str(tmp_path / "m2") + "/", seed=2)
    m3 = committee_member(base, str(tmp_path / "m3") + "/", seed=3)
    m3s = committee_member(base, str(tmp_path / "m3s") + "/", seed=3, scaling_factor=1.5)
    P = [oracle.read_potential(p) for p in (base, m2, m3)]
    s = golden_structure(wl, golden("eval_ma_cell8"))
    single = [oracle.Oracle(p).evaluate(s) for p in P]
    E = np.array([r["energy"] for r in single])
    c = oracle.committee(P, s)
    assert abs(c["energy"] - E.mean()) <= 1e-12 * abs(E.mean())
    np.testing.assert_allclose(c["force"], sum(r["force"] for r in single) / 3, rtol=0, atol=1e-12)
    np.testing.assert_allclose(c["dEdLat"], sum(r["dEdLat"] for r in single) / 3, rtol=0, atol=1e-12)
    assert abs(c["std"] - E.std()) <= 1e-12
    assert E.std() > 1e-6                        # the members do differ
    c2 = oracle.committee([P[0], oracle.read_potential(m3s)], s)
    assert c2["members"][1] == single[2]["energy"]
    alone = oracle.Oracle(oracle.read_potential(m3s)).evaluate(s)["energy"]
    assert abs(alone - single[2]["energy"]) > 1e-6    # the scaling would matter alone
    c1 = oracle.committee([P[0]], s)
    assert c1["energy"] == single[0]["energy"] and c1["std"] == 0.0
