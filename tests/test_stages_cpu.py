"""The stage fixtures (tests/golden/stages_*.npz, written by the reference's
own hdnnpCalcNeighborLists / hdnnpCalcSymfunctions / hdnnpCalcNNs through
oracle/refdump) pinned before the GPU tests use them:

- the reference's two symmetry-function paths, calcSfsTogether .true.
  (calcAllSymmetryFunctions) and .false. (symmetryFunction2 / 3), agree;
- the stage routines give the energies, forces and dEdLat of
  hdnnpEnergyAndForces (the eval fixtures) and its neighbour sets;
- the G of the stage fixtures equal the eval fixtures' G;
- dsf(:, n) is dG/dR of the atom in slot n: the oracle's G under central
  finite differences of that atom's position (a cluster, where every atom
  is in a list once) and, for a periodic cell, the sum of dsf over the slots
  of one atom (its images) against the same difference;
- slot 1 of every list is the atom itself.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import POTENTIALS, golden, golden_structure

CASES = ["ma_tric2", "cluster30", "perov1"]


def rows(d, i):
    """(ineis, neis, G, dsf or None) of atom i of a stage fixture"""
    nn, ns = d["nneis"], d["nsf"]
    o = int(nn[:i].sum())
    og = int(ns[:i].sum())
    ineis = d["ineis"][o:o + nn[i]]
    neis = d["neis"][o:o + nn[i]]
    G = d["G"][og:og + ns[i]]
    dsf = None
    if i < int(d["ndsf"]):
        od = int((3 * ns[:i] * nn[:i]).sum())
        dsf = d["dsf"][od:od + 3 * ns[i] * nn[i]].reshape(ns[i], nn[i], 3)
    return ineis, neis, G, dsf


@pytest.mark.parametrize("case", CASES)
def test_reference_paths_agree(case):
    t, f = golden(f"stages_{case}_T"), golden(f"stages_{case}_F")
    assert abs(float(t["energy"]) - float(f["energy"])) <= 1e-12 * abs(float(t["energy"]))
    np.testing.assert_allclose(t["force"], f["force"], rtol=0, atol=1e-12)
    np.testing.assert_array_equal(t["ineis"], f["ineis"])
    np.testing.assert_array_equal(t["neis"], f["neis"])
    np.testing.assert_allclose(t["G"], f["G"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(t["dsf"], f["dsf"], rtol=0, atol=1e-11)


@pytest.mark.parametrize("case", CASES)
def test_stages_give_the_evaluation(case):
    st, ev = golden(f"stages_{case}_T"), golden(f"eval_{case}")
    E = float(ev["energy"])
    assert abs(float(st["energy"]) - E) <= 1e-12 * max(1.0, abs(E))
    np.testing.assert_allclose(st["force"], ev["force"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(st["dEdLat"], ev["dEdLat"], rtol=0, atol=1e-12 * max(1.0, np.abs(ev["dEdLat"]).max()))
    nat = len(st["nneis"])
    ent = []
    for i in range(nat):
        ineis, neis, G, _ = rows(st, i)
        assert ineis[0] == i + 1
        ent.extend([(i, j - 1) for j in ineis[1:]])
        if "G" in ev:
            np.testing.assert_allclose(G, ev["G"][i, :len(G)], rtol=0, atol=1e-14)
    assert len(ent) == int(ev["nneigh"])
    got = sorted(ent)
    want = sorted(map(tuple, ev["neighbors"][:, :2].tolist())) if "neighbors" in ev else got
    assert got == want


@pytest.mark.parametrize("case, atoms", [("cluster30", [0, 7, 19]), ("ma_tric2", [0, 5]), ("perov1", [0, 2])])
def test_dsf_is_the_gradient_of_G(oracles, wl, case, atoms):
    d = golden(f"stages_{case}_T")
    s = golden_structure(wl, d)
    orc = oracles[str(d["potential"])]
    h = 1e-5
    for i in atoms:
        ineis, _, G0, dsf = rows(d, i)
        assert dsf is not None
        for j in sorted(set((ineis - 1).tolist()))[:6]:
            want = dsf[:, ineis - 1 == j, :].sum(axis=1)       # all slots of atom j
            fd = np.zeros_like(want)
            for c in range(3):
                sp, sm = s.copy(), s.copy()
                sp.pos = s.pos.copy()
                sm.pos = s.pos.copy()
                sp.pos[j, c] += h
                sm.pos[j, c] -= h
                Gp = orc.evaluate(sp, want_sf=True)["G"][i, :len(G0)]
                Gm = orc.evaluate(sm, want_sf=True)["G"][i, :len(G0)]
                fd[:, c] = (Gp - Gm) / (2 * h)
            np.testing.assert_allclose(want, fd, rtol=0, atol=2e-6 * max(1.0, np.abs(fd).max()))
