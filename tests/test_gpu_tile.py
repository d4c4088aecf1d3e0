"""Tiles of the symmetry-function kernels (atoms per wavefront, 1/2/4) and
the autotuner that picks one per context: every tile reproduces the
reference's golden vectors and the oracle, batched results equal single
ones within a context, and the tuner keeps the fastest tile it measured."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import MA_PARAMS, PEROV_PARAMS, POTENTIALS, golden, golden_structure

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("tile", [2, 4])
@pytest.mark.parametrize("case", ["ma", "ma_cell1", "ma_tric2", "ma_cell8", "cluster30", "perov1", "perov2"])
def test_tile_matches_reference(pkg, wl, tile, case):
    g = golden("eval_" + case)
    s = golden_structure(wl, g)
    pot = POTENTIALS[str(g["potential"])]
    h = pkg.parseRuNNerInput(pot, "fp64")
    h.tile = tile
    assert h.tile == tile
    E, F, L = h.energy_and_forces(s)
    e = float(g["energy"])
    assert abs(E - e) <= 1e-10 * max(1.0, abs(e))
    np.testing.assert_allclose(F, g["force"], rtol=0, atol=1e-9)
    if s.periodic:
        np.testing.assert_allclose(L, g["dEdLat"], rtol=0, atol=1e-9 * max(1.0, np.abs(g["dEdLat"]).max()))
    hm = pkg.parseRuNNerInput(pot, "mixed")
    hm.tile = tile
    Em, Fm, _ = hm.energy_and_forces(s)
    assert np.sqrt(np.mean((Fm - F) ** 2)) <= 1e-5 and abs(Em - E) / s.nat <= 1e-5


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tile_batch_equals_single(pkg, wl, precision):
    h = pkg.parseRuNNerInput(PEROV_PARAMS, precision)
    h.tile = 2
    ss = [wl.cluster(8 + 7 * k, seed=k) for k in range(6)] + [wl.perovskite(1, 1, 2, seed=9)]
    E, F, L = h.energy_and_forces_batch(ss)
    for k, s in enumerate(ss):
        e, f, l = h.energy_and_forces(s)
        assert E[k] == e and np.array_equal(F[k], f), k


def test_autotune_keeps_the_fastest(pkg, wl):
    """On a batch of molecules (7 neighbours per atom) and on perovskite
    cells (~52) the tuner times every tile the sample's capacity takes and
    keeps the fastest; the tuned context still reproduces the reference."""
    rng = np.random.default_rng(1)
    mols = []
    for k in range(256):
        m = wl.methylammonium()
        m.pos = m.pos @ wl._random_rotation(rng).T + rng.normal(0, 0.05, m.pos.shape)
        mols.append(m)
    h = pkg.parseRuNNerInput(MA_PARAMS, "mixed")
    t, ms = h.autotune(mols)
    print("molecules:", t, ms, flush=True)
    assert t in (1, 2, 4) and h.tile == t and ms[t] > 0
    assert ms[t] == min(v for v in ms.values() if v > 0)
    g = golden("eval_ma")
    E, F, _ = h.energy_and_forces(golden_structure(wl, g))
    assert abs(E - float(g["energy"])) <= 1e-5 * 8 and np.sqrt(np.mean((F - g["force"]) ** 2)) <= 1e-5
    hp = pkg.parseRuNNerInput(PEROV_PARAMS, "mixed")
    tp, msp = hp.autotune([wl.perovskite(4, 4, 4, seed=2)])
    print("perovskite:", tp, msp, flush=True)
    assert hp.tile == tp and msp[tp] == min(v for v in msp.values() if v > 0)
