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
