"""Staging capacity: neighbour counts that take each staging size of the
symmetry-function kernels (64, 96, 128, 256, 512, 1024 entries per atom),
and counts beyond what a build stages.

The structures are 3x3x3 periodic cells of a jittered simple cubic Pb
lattice, whose spacing sets the neighbours within the 12-bohr cutoff (about
62, 81, 104, 226, 424, 766 and 1241 per atom). Dense lattices put the
symmetry functions far outside the potential's scaling range and the MLP
saturates, so besides E and F the scaled G and dE/dG of every atom are
compared with the oracle (FP64: 1e-11 relative to max(1, |x|)).
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import PEROV_PARAMS

pytestmark = pytest.mark.gpu


def _lattice(wl, d, m=3, seed=0):
    rng = np.random.default_rng(seed)
    g = np.stack(np.meshgrid(*[np.arange(m)] * 3, indexing="ij"), -1).reshape(-1, 3) * float(d)
    pos = g + rng.normal(scale=0.05 * d, size=g.shape)
    return wl.Structure(pos, np.full(len(g), 82, np.int32), np.eye(3) * m * d, True)


def _close(a, b, rel):
    a, b = np.asarray(a), np.asarray(b)
    return np.all(np.abs(a - b) <= rel * np.maximum(1.0, np.abs(b)))


@pytest.mark.parametrize("d", [5.0, 4.4, 4.0, 3.2, 2.6])
def test_staging_sizes_against_oracle(pkg, wl, oracles, d):
    s = _lattice(wl, d)
    h = pkg.parseRuNNerInput(PEROV_PARAMS, "fp64")
    E, F, L = h.energy_and_forces(s)
    G, D, Ei = h.symfunctions(s.nat)
    sel = np.arange(s.nat, dtype=np.int32)
    Eo, Go, Do = oracles["perovskite"].atoms(s, sel)
    r = oracles["perovskite"].evaluate(s)
    print(f"d={d}: neighbours {len(h.neighbor_entries()) / s.nat:.0f} per atom, E={E:.10f}", flush=True)
    assert _close(G, Go, 1e-11) and _close(D, Do, 1e-11) and _close(Ei, Eo, 1e-11)
    assert abs(E - r["energy"]) <= 1e-10 * max(1.0, abs(E))
    np.testing.assert_allclose(F, r["force"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(L, r["dEdLat"], rtol=0, atol=1e-9 * max(1.0, np.abs(r["dEdLat"]).max()))


def test_largest_staging_size(pkg, wl, oracles):
    """~766 neighbours per atom: the mixed form stages 1024 entries; FP64,
    whose staging arrays are twice as wide, reports that it stages at most
    512 and stays usable."""
    s = _lattice(wl, 2.1)
    hm = pkg.parseRuNNerInput(PEROV_PARAMS, "mixed")
    E, F, _ = hm.energy_and_forces(s)
    G, _, _ = hm.symfunctions(s.nat)
    _, Go, _ = oracles["perovskite"].atoms(s, np.arange(s.nat, dtype=np.int32))
    assert np.isfinite(E) and np.all(np.isfinite(F))
    assert _close(G, Go, 1e-4)
    h = pkg.parseRuNNerInput(PEROV_PARAMS, "fp64")
    with pytest.raises(pkg.HdnnpError) as ei:
        h.energy_and_forces(s)
    assert ei.value.code == 6 and "at most 512" in str(ei.value)
    E2, _, _ = h.energy_and_forces(_lattice(wl, 4.0))
    assert np.isfinite(E2)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_beyond_staging_capacity(pkg, wl, precision):
    """~1241 neighbours per atom exceed every staging size: a capacity error
    that names the count, and the context stays usable."""
    h = pkg.parseRuNNerInput(PEROV_PARAMS, precision)
    with pytest.raises(pkg.HdnnpError) as ei:
        h.energy_and_forces(_lattice(wl, 1.8))
    assert ei.value.code == 6 and "neighbours" in str(ei.value)
    E, F, _ = h.energy_and_forces(_lattice(wl, 5.0))
    assert np.isfinite(E) and np.all(np.isfinite(F))
