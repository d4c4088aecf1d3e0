"""Randomised parity sweep: seeded structures of the shapes the fixtures do
not cover, evaluated on the GPU in FP64 and compared with the oracle
(E 1e-10 relative, F 1e-9 Ha/bohr, dEdLat 1e-9 max(1, |dEdLat|)); the same
structures as one batch equal the single evaluations bit for bit.

Kinds (perovskite potential, elements H C N I Pb):
- triclinic cells: lattice vectors of 12-30 bohr at 60-120 degree angles,
  20-80 atoms at least 1.5 bohr apart;
- small cells: edges of 5-10 bohr with 2-8 atoms, so that every atom sees
  many images of itself;
- clusters: 8-120 atoms, nonperiodic.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import PEROV_PARAMS

pytestmark = pytest.mark.gpu

ELEMS = np.array([1, 6, 7, 53, 82], np.int32)


def _lattice(rng, lo, hi):
    a, b, c = rng.uniform(lo, hi, 3)
    al, be, ga = np.radians(rng.uniform(60, 120, 3))
    v1 = np.array([a, 0.0, 0.0])
    v2 = np.array([b * np.cos(ga), b * np.sin(ga), 0.0])
    cx = c * np.cos(be)
    cy = c * (np.cos(al) - np.cos(be) * np.cos(ga)) / np.sin(ga)
    cz2 = c * c - cx * cx - cy * cy
    if cz2 < (0.3 * c) ** 2:
        return None
    return np.array([v1, v2, [cx, cy, np.sqrt(cz2)]])


def _fill(rng, lat, n, dmin):
    pos = []
    tries = 0
    while len(pos) < n and tries < 20000:
        tries += 1
        f = rng.random(3)
        p = f @ lat
        ok = True
        for q in pos:
            d = (p - q) @ np.linalg.inv(lat)
            d -= np.round(d)
            if np.linalg.norm(d @ lat) < dmin:
                ok = False
                break
        if ok:
            pos.append(p)
    return np.array(pos)


def _structure(wl, seed):
    rng = np.random.default_rng(1000 + seed)
    kind = seed % 3
    if kind == 0:
        lat = None
        while lat is None:
            lat = _lattice(rng, 12.0, 30.0)
        pos = _fill(rng, lat, int(rng.integers(20, 81)), 1.5)
        per = True
    elif kind == 1:
        lat = None
        while lat is None:
            lat = _lattice(rng, 5.0, 10.0)
        pos = _fill(rng, lat, int(rng.integers(2, 9)), 1.5)
        per = True
    else:
        c = wl.cluster(int(rng.integers(8, 121)), seed=int(rng.integers(1 << 30)), box=22.0, dmin=1.5)
        pos, lat, per = c.pos, np.zeros((3, 3)), False
    Z = rng.choice(ELEMS, size=len(pos)).astype(np.int32)
    return wl.Structure(np.ascontiguousarray(pos), Z, lat, per)


SEEDS = list(range(24))


@pytest.fixture(scope="module")
def sweep(pkg, wl):
    return [_structure(wl, k) for k in SEEDS]


@pytest.mark.parametrize("k", SEEDS)
def test_random_structure_against_oracle(pkg, oracles, sweep, k):
    s = sweep[k]
    h = pkg.parseRuNNerInput(PEROV_PARAMS, "fp64")
    E, F, L = h.energy_and_forces(s)
    r = oracles["perovskite"].evaluate(s)
    assert abs(E - r["energy"]) <= 1e-10 * max(1.0, abs(E))
    np.testing.assert_allclose(F, r["force"], rtol=0, atol=1e-9)
    if s.periodic:
        np.testing.assert_allclose(L, r["dEdLat"], rtol=0, atol=1e-9 * max(1.0, np.abs(r["dEdLat"]).max()))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_random_batch_equals_single(pkg, sweep, precision):
    h = pkg.parseRuNNerInput(PEROV_PARAMS, precision)
    single = [h.energy_and_forces(s) for s in sweep]
    Eb, Fb, Lb = h.energy_and_forces_batch(sweep)
    for k, s in enumerate(sweep):
        assert Eb[k] == single[k][0]
        assert np.array_equal(Fb[k], single[k][1])
        if s.periodic:
            assert np.array_equal(Lb[k], single[k][2])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_one_workgroup_sort_equals_radix_sort(pkg, precision):
    """A build of up to 8,192 atoms sorts atoms by cell and by element in one
    workgroup; a larger one uses the radix sort and the chunked element sort.
    Each structure here is below the limit alone and the batch of all three is
    above it. The results must be bit-identical both ways, so both sorts give
    the same atom orders."""
    wl = pkg.workloads
    structs = [wl.perovskite(6, 6, 6, seed=11), wl.methylammonium(), wl.perovskite(8, 8, 8, seed=12)]
    assert max(len(s.Z) for s in structs) <= 8192 < sum(len(s.Z) for s in structs)
    h = pkg.parseRuNNerInput(PEROV_PARAMS, precision)
    hm = pkg.parseRuNNerInput(PEROV_PARAMS, precision)
    single = [h.energy_and_forces(s) for s in structs]
    Eb, Fb, Lb = hm.energy_and_forces_batch(structs)
    for k, s in enumerate(structs):
        print(f"{precision} structure {k} nat={len(s.Z)}: E alone {single[k][0]!r} in batch {Eb[k]!r}", flush=True)
        assert Eb[k] == single[k][0]
        assert np.array_equal(Fb[k], single[k][1])
        if s.periodic:
            assert np.array_equal(Lb[k], single[k][2])


def _dense_cell_cluster(wl, nblob=300, nspread=1500, seed=31):
    """A nonperiodic structure whose cell grid holds ~2.5 atoms per cell except
    one cell with `nblob` atoms: a 1.5-bohr blob placed in the middle of one
    grid cell (cells are at least 0.6 rc = 7.2 bohr wide) inside a spread
    cluster of `nspread` atoms."""
    s = wl.cluster_scaled(nspread, seed=seed)
    lo, hi = s.pos.min(axis=0), s.pos.max(axis=0)
    length = (hi + 1e-4) - (lo - 1e-4)
    ng = np.floor(length / (0.6 * 12.0)) + 1
    centre = (lo - 1e-4) + (np.floor(ng / 2) + 0.5) * length / ng
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((nblob, 3))
    d *= (1.5 * rng.random(nblob) ** (1 / 3) / np.linalg.norm(d, axis=1))[:, None]
    far = np.linalg.norm(s.pos - centre, axis=1) > 4.0       # room for the blob
    pos = np.vstack([s.pos[far], centre + d])
    Z = np.concatenate([s.Z[far], rng.choice(np.array([1, 6, 7], np.int32), nblob)]).astype(np.int32)
    return wl.Structure(pos, Z, np.zeros((3, 3)), False)


def test_one_workgroup_sort_with_a_crowded_cell(pkg):
    """A cell holding more atoms than the one-workgroup sort ranks by scanning
    the cell (256) switches it to ranking in rounds of the workgroup's size.
    The orders must still be the radix sort's: the structure alone (one
    workgroup) and in a batch above 8,192 atoms (radix sort) give bit-identical
    results and the same neighbour entries."""
    wl = pkg.workloads
    s = _dense_cell_cluster(wl)
    big = [wl.perovskite(8, 8, 8, seed=41), wl.perovskite(4, 4, 4, seed=42)]
    assert s.nat <= 8192 < s.nat + sum(b.nat for b in big)
    h = pkg.parseRuNNerInput(PEROV_PARAMS, "fp64")
    e, f, _ = h.energy_and_forces(s)
    ent = h.neighbor_entries()
    per_atom = np.bincount(ent[:, 0], minlength=s.nat)
    print(f"crowded cell: nat={s.nat}, most neighbours {per_atom.max()}", flush=True)
    assert per_atom.max() >= 299
    hb = pkg.parseRuNNerInput(PEROV_PARAMS, "fp64")
    Eb, Fb, _ = hb.energy_and_forces_batch([big[0], s, big[1]])
    assert Eb[1] == e
    assert np.array_equal(Fb[1], f)
    eb = hb.neighbor_entries()
    off = big[0].nat
    mine = eb[(eb[:, 0] >= off) & (eb[:, 0] < off + s.nat)].copy()
    mine[:, :2] -= off
    key = lambda a: a[np.lexsort(a.T[::-1])]
    np.testing.assert_array_equal(key(mine), key(ent))
