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


def _dense_in_dilute(wl):
    c = wl.cluster(120, seed=11, box=14.0, dmin=1.6)
    rng = np.random.default_rng(12)
    gas = []
    while len(gas) < 40:
        p = rng.random(3) * 60.0
        if np.min(np.linalg.norm(c.pos + 23.0 - p, axis=1)) > 14.0 and all(np.linalg.norm(p - q) > 3.0 for q in gas):
            gas.append(p)
    pos = np.vstack([c.pos + 23.0, np.array(gas)])
    Z = np.concatenate([c.Z, rng.choice(np.array([1, 6, 7], np.int32), size=len(gas))]).astype(np.int32)
    return wl.Structure(pos, Z, np.eye(3) * 60.0, True)


def test_triplet_list_fallback_fp64(pkg, wl, oracles):
    """FP64 hands the value pass's triplets to the gradient pass, in per-atom
    lists sized from the mean density. A dense 120-atom C/N/H cluster in a
    60-bohr periodic box, with 40 dilute atoms around it, has ~100
    neighbours per cluster atom at a mean of ~4: the cluster atoms overflow
    their lists and are compacted by the gradient pass itself, while the
    dilute atoms use theirs. Both against the oracle."""
    s = _dense_in_dilute(wl)
    h = pkg.parseRuNNerInput(PEROV_PARAMS, "fp64")
    E, F, L = h.energy_and_forces(s)
    ent = h.neighbor_entries()
    cnt = np.bincount(np.asarray(ent)[:, 0], minlength=s.nat)
    print(f"neighbours: cluster mean {cnt[:120].mean():.0f}, gas mean {cnt[120:].mean():.1f}", flush=True)
    assert cnt[:120].mean() > 60 and cnt[120:].max() < 20
    r = oracles["perovskite"].evaluate(s)
    assert abs(E - r["energy"]) <= 1e-10 * max(1.0, abs(E))
    np.testing.assert_allclose(F, r["force"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(L, r["dEdLat"], rtol=0, atol=1e-9 * max(1.0, np.abs(r["dEdLat"]).max()))


def test_triplet_list_windows_mixed(pkg, wl, oracles):
    """The same structure in mixed precision. The FP32 value pass gives each
    atom whose candidate pairs fit its list one region per element-pair block
    (the dilute atoms); the cluster atoms (~5,000 candidate pairs against a
    list of 2,048) are evaluated in windows of the list, and the gradient
    pass compacts their triplets itself. Against the oracle within the mixed
    bounds, and against the FP64 evaluation."""
    s = _dense_in_dilute(wl)
    hm = pkg.parseRuNNerInput(PEROV_PARAMS, "mixed")
    Em, Fm, Lm = hm.energy_and_forces(s)
    cnt = np.bincount(np.asarray(hm.neighbor_entries())[:, 0], minlength=s.nat)
    print(f"neighbours: most {cnt.max()}, atoms with more than 64: {(cnt > 64).sum()}", flush=True)
    assert (cnt > 64).sum() >= 20 and cnt[120:].max() < 20   # pairs(65) = 2,080 > 2,048
    E, F, L = pkg.parseRuNNerInput(PEROV_PARAMS, "fp64").energy_and_forces(s)
    r = oracles["perovskite"].evaluate(s)
    rmse = float(np.sqrt(np.mean((Fm - r["force"]) ** 2)))
    print(f"mixed: dE/atom {abs(Em - r['energy']) / s.nat:.3e} F rmse {rmse:.3e} "
          f"dEdLat max {np.abs(Lm - r['dEdLat']).max():.3e}", flush=True)
    assert abs(Em - r["energy"]) / s.nat <= 1e-5 * max(1.0, abs(r["energy"]) / s.nat)
    assert rmse <= 1e-5 * max(1.0, float(np.sqrt(np.mean(r["force"] ** 2))))
    assert np.abs(Lm - L).max() <= 1e-5 * max(1.0, np.abs(L).max())
