"""On-device molecular dynamics (hdnnp_md_*): the reference example's velocity
Verlet (hdnnpExample.f90:58-72) against the reference trajectory, the Verlet
skin against rebuilding every step, and several systems stepped together
against each alone."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import MA_PARAMS, PEROV_PARAMS, golden

pytestmark = pytest.mark.gpu

DT = 0.5 * 41.34137333519408      # 0.5 fs in atomic time units


def _trajectory(pkg, wl, skin, nsteps):
    g = golden("md_methylammonium")
    h = pkg.parseRuNNerInput(MA_PARAMS, "fp64")
    md = pkg.MdSystems(h, [wl.methylammonium()], [g["p0"]], DT, skin=skin)
    out = []
    for _ in range(nsteps):
        ep, ek = md.step(1)
        out.append([ep[0], ek[0], ep[0] + ek[0]])
    return np.array(out), md


def test_md_matches_reference_trajectory(pkg, wl):
    """1001 steps, lists rebuilt at every step (skin 0): the reference run's
    energies per step to 1e-9 Ha and its final positions."""
    g = golden("md_methylammonium")
    steps, md = _trajectory(pkg, wl, 0.0, 1001)
    np.testing.assert_allclose(steps, g["steps"], rtol=0, atol=1e-9)
    pos, mom, _ = md.get()
    np.testing.assert_allclose(pos, g["pos_final"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(mom, g["p_final"], rtol=0, atol=1e-9)
    assert md.stats == (1001, 1002)


def test_md_skin_matches_rebuild(pkg, wl):
    """A 1-bohr skin reuses the lists and follows the same trajectory."""
    g = golden("md_methylammonium")
    steps, md = _trajectory(pkg, wl, 1.0, 1001)
    nsteps, builds = md.stats
    print(f"skin 1 bohr: {builds} list builds in {nsteps} steps; "
          f"max |dE| vs reference {np.abs(steps - g['steps']).max():.3e}", flush=True)
    assert builds < nsteps / 5
    np.testing.assert_allclose(steps, g["steps"], rtol=0, atol=1e-8)


def test_md_systems_together_equal_alone(pkg, wl):
    """Three systems stepped in one batch equal each stepped alone (skin 0),
    bit for bit; a periodic perovskite cell among them."""
    h = pkg.parseRuNNerInput(PEROV_PARAMS, "fp64")
    rng = np.random.default_rng(5)
    ss = [wl.perovskite(2, 1, 1, seed=31), wl.cluster(24, seed=32), wl.perovskite(1, 1, 1, seed=33)]
    ps = [rng.standard_normal((s.nat, 3)) * 2.0 for s in ss]
    md = pkg.MdSystems(h, ss, ps, DT, skin=0.0)
    ep, ek = md.step(20)
    pos, mom, f = md.get()
    off = np.concatenate([[0], np.cumsum([s.nat for s in ss])])
    for k, s in enumerate(ss):
        m1 = pkg.MdSystems(h, [s], [ps[k]], DT, skin=0.0)
        e1, k1 = m1.step(20)
        p1, q1, f1 = m1.get()
        assert e1[0] == ep[k] and k1[0] == ek[k]
        assert np.array_equal(p1, pos[off[k]:off[k + 1]]) and np.array_equal(f1, f[off[k]:off[k + 1]])


def test_md_periodic_skin_energy_conservation(pkg, wl):
    """Methylammonium in a periodic 20-bohr box (images beyond the cutoff),
    drifting across the cell boundary: with a 1-bohr skin the lists are reused
    across the wrap, the run follows the one that rebuilds every step, and the
    total energy is conserved. (Dense periodic boxes are outside what the
    example potential was fitted to, and the synthetic perovskite potential
    has random weights: neither is stable under MD.)"""
    g = golden("md_methylammonium")
    s = wl.methylammonium()
    s.periodic = True
    s.lat = np.eye(3) * 20.0
    s.pos = s.pos + np.array([18.5, 10.0, 10.0])          # near the x face
    m = pkg.atomic_masses(s.Z)
    p = g["p0"] + m[:, None] * np.array([2.5e-4, 1e-4, 0.0])   # centre-of-mass drift
    ha = pkg.parseRuNNerInput(MA_PARAMS, "fp64")
    hb = pkg.parseRuNNerInput(MA_PARAMS, "fp64")
    a = pkg.MdSystems(ha, [s], [p], DT, skin=1.0)
    b = pkg.MdSystems(hb, [s], [p], DT, skin=0.0)
    ea, eb = [], []
    for _ in range(400):
        x = a.step(1)
        y = b.step(1)
        ea.append(x[0][0] + x[1][0])
        eb.append(y[0][0] + y[1][0])
    pos, _, _ = a.get()
    print(f"builds with skin: {a.stats}; total-energy drift {ea[-1] - ea[0]:.3e}; "
          f"x of N: {pos[0, 0]:.3f}", flush=True)
    assert pos[:, 0].max() > 20.0                        # crossed the cell face
    assert a.stats[1] < 40
    np.testing.assert_allclose(ea, eb, rtol=0, atol=1e-9)
    assert abs(ea[-1] - ea[0]) < 1e-4


def test_md_shared_context_rebuilds(pkg, wl):
    """Two MD objects and plain evaluations interleaved on one context: each
    object notices that its lists were replaced, builds them again, and
    follows the trajectory it has alone on a context of its own."""
    g = golden("md_methylammonium")
    h = pkg.parseRuNNerInput(MA_PARAMS, "fp64")
    s = wl.methylammonium()
    a = pkg.MdSystems(h, [s], [g["p0"]], DT, skin=1.0)
    b = pkg.MdSystems(h, [s], [g["p0"] * 0.5], DT, skin=1.0)
    ea, eb = [], []
    for t in range(30):
        ea.append(a.step(1)[0][0])
        if t % 3 == 0:
            pkg.hdnnpEnergyAndForces(h, s)
        eb.append(b.step(1)[0][0])
    assert a.stats[1] == 31 and b.stats[1] == 31     # every step found foreign lists
    h1 = pkg.parseRuNNerInput(MA_PARAMS, "fp64")
    a1 = pkg.MdSystems(h1, [s], [g["p0"]], DT, skin=1.0)
    e1 = [a1.step(1)[0][0] for _ in range(30)]
    assert a1.stats[1] < 5
    np.testing.assert_allclose(ea, e1, rtol=0, atol=1e-9)
    np.testing.assert_allclose(ea, g["steps"][:30, 0], rtol=0, atol=1e-9)


def test_md_chunks_equal_single_steps(pkg, wl):
    """Steps on reused lists run in chunks that are checked at their end and
    undone (then stepped one at a time) when an atom left the skin: 300 steps
    in one call equal 300 calls of one step bit for bit, with the same list
    builds, on the repulsive perovskite potential at 300 K, and the total
    energy is conserved."""
    from conftest import PEROV_MD_PARAMS
    s = wl.perovskite(4, 2, 2, seed=71)
    m = pkg.atomic_masses(s.Z)
    rng = np.random.default_rng(8)
    p = np.sqrt(m * 300.0 * 3.166811563e-6)[:, None] * rng.standard_normal((s.nat, 3))
    out = []
    for single in (True, False):
        h = pkg.parseRuNNerInput(PEROV_MD_PARAMS, "mixed")
        md = pkg.MdSystems(h, [s], [p], DT, skin=1.0)
        e0 = sum(x[0] for x in md.step(0))
        if single:
            for _ in range(300):
                ep, ek = md.step(1)
        else:
            ep, ek = md.step(300)
        out.append((ep[0], ek[0], *md.get(), md.stats))
        print(f"single {single}: builds {md.stats[1]}, drift {ep[0] + ek[0] - e0:.3e}", flush=True)
        assert abs(ep[0] + ek[0] - e0) < 1e-3
    a, b = out
    assert a[0] == b[0] and a[1] == b[1] and a[5] == b[5]
    for x, y in zip(a[2:5], b[2:5]):
        assert np.array_equal(x, y)
    assert 5 < a[5][1] < 150          # lists were reused, and rebuilt


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_md_graph_replays_equal_direct_launches(pkg, wl, precision, monkeypatch):
    """All but the first step of a chunk are replays of a captured HIP graph
    of one step. 60 calls of 5 steps with graphs equal the same calls with
    every step launched directly, bit for bit after every call (each replay
    feeds the next, with the program's own synchronisations between)."""
    from conftest import PEROV_MD_PARAMS
    s = wl.perovskite(4, 2, 2, seed=72)
    m = pkg.atomic_masses(s.Z)
    p = np.sqrt(m * 300.0 * 3.166811563e-6)[:, None] * np.random.default_rng(9).standard_normal((s.nat, 3))
    runs = []
    for graphs in ("1", "0"):
        monkeypatch.setenv("HDNNP_MD_GRAPHS", graphs)
        h = pkg.parseRuNNerInput(PEROV_MD_PARAMS, precision)
        md = pkg.MdSystems(h, [s], [p], DT, skin=1.0)
        trace = []
        for _ in range(60):
            ep, ek = md.step(5)
            trace.append((ep[0], ek[0], md.get()[0].copy()))
        runs.append((trace, md.counters))
    (ta, ca), (tb, cb) = runs
    print(f"{precision}: with graphs {ca}; without {cb}", flush=True)
    assert ca["graph_replays"] > 100 and ca["graph_captures"] >= 1 and cb["graph_replays"] == 0
    assert ca["builds"] == cb["builds"] and ca["rollbacks"] == cb["rollbacks"]
    for x, y in zip(ta, tb):
        assert x[0] == y[0] and x[1] == y[1] and np.array_equal(x[2], y[2])


def test_md_failed_create_frees_its_buffers(pkg, wl):
    """An MD object whose first forces fail (an element the potential lacks)
    raises and gives its device buffers back: the per-step state saves of a
    24,576-atom cell are ~115 MB, so four leaked creations would show."""
    import torch
    h = pkg.parseRuNNerInput(PEROV_PARAMS, "mixed")
    s = wl.perovskite(16, 16, 8, seed=3)
    s.Z = s.Z.copy()
    s.Z[5] = 79
    p = np.zeros((s.nat, 3))
    m = np.ones(s.nat)
    with pytest.raises(pkg.HdnnpError):            # sizes the context's own buffers, which it keeps
        pkg.MdSystems(h, [s], [p], DT, skin=1.0, masses=m)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for _ in range(4):
        with pytest.raises(pkg.HdnnpError) as ei:
            pkg.MdSystems(h, [s], [p], DT, skin=1.0, masses=m)
        assert ei.value.code == 4 and "Z=79" in str(ei.value)
    torch.cuda.synchronize()
    assert free0 - torch.cuda.mem_get_info()[0] < (64 << 20)
    # the context still evaluates
    good = wl.perovskite(2, 2, 2, seed=4)
    E, F, _ = h.energy_and_forces(good)
    assert np.isfinite(E) and np.all(np.isfinite(F))
