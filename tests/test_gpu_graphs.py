"""Graph replay of small evaluations (hdnnp_set_graphs, off by default): an
evaluation of up to 8,192 atoms on the context's own stream is captured when
it repeats the launch of the evaluation before it and replayed afterwards. Every replayed
result equals a direct launch of the same input bit for bit (E, F, dEdLat),
for moved positions as in an MD loop, the host and the device entry points,
batches, both precisions, and after a capacity retry; a new shape is
launched directly and captured on its own repeat.
"""
from __future__ import annotations

import time

import numpy as np
import pytest

from conftest import MA_PARAMS, PEROV_PARAMS

pytestmark = pytest.mark.gpu


def _pair(pkg, params, precision):
    g = pkg.parseRuNNerInput(params, precision)
    d = pkg.parseRuNNerInput(params, precision)
    assert not g.graphs and not d.graphs          # off by default
    g.graphs = True
    assert g.graphs
    return g, d


def _same(a, b):
    (Ea, Fa, La), (Eb, Fb, Lb) = a, b
    assert Ea == Eb and np.array_equal(Fa, Fb)
    assert (La is None and Lb is None) or np.array_equal(La, Lb)


def _moved(s, k, sigma=0.02):
    t = s.copy()
    t.pos = s.pos + np.random.default_rng(100 + k).normal(0.0, sigma, s.pos.shape)
    return t


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_replay_equals_direct_host(pkg, wl, precision):
    """the reference's example (8-atom methylammonium) and a periodic
    perovskite cell through hdnnp_evaluate, positions moving every call"""
    for params, s0 in ((MA_PARAMS, wl.methylammonium()), (PEROV_PARAMS, wl.perovskite(2, 2, 1, seed=3))):
        g, d = _pair(pkg, params, precision)
        for k in range(6):
            s = _moved(s0, k)
            _same(g.energy_and_forces(s), d.energy_and_forces(s))
        c, r = g.graph_counts()
        print(f"{precision} nat={s0.nat}: captured {c}, replayed {r}", flush=True)
        assert c == 1 and r == 5, (c, r)
        assert d.graph_counts() == (0, 0)


def test_replay_equals_direct_device(pkg, wl):
    """hdnnp_evaluate_device with the same device arrays: replayed; new output
    array: a new launch (captured on its own repeat)"""
    import torch
    dev = torch.device("cuda:0")
    s0 = wl.perovskite(3, 3, 3, seed=8)
    g, d = _pair(pkg, PEROV_PARAMS, "mixed")
    Z = torch.from_numpy(s0.Z.astype(np.int32)).to(dev)
    pos = torch.empty((s0.nat, 3), dtype=torch.float64, device=dev)
    fg = torch.empty_like(pos)
    fd = torch.empty_like(pos)
    for k in range(5):
        s = _moved(s0, k)
        pos.copy_(torch.from_numpy(s.pos))
        Eg, _, Lg = g.evaluate_torch(Z, pos, True, s.lat, fg)
        Ed, _, Ld = d.evaluate_torch(Z, pos, True, s.lat, fd)
        assert Eg == Ed and torch.equal(fg, fd) and np.array_equal(Lg, Ld)
    assert g.graph_counts() == (1, 4)
    f2 = torch.empty_like(pos)
    for k in range(3):
        Eg, _, _ = g.evaluate_torch(Z, pos, True, s.lat, f2)
        assert Eg == Ed and torch.equal(f2, fd)
    assert g.graph_counts() == (2, 6)


def test_replay_batch_and_shape_change(pkg, wl):
    """a batch of small structures replays; another batch shape is launched
    directly, then captured; the first shape again (replayed while no buffer
    of the second had to grow)"""
    g, d = _pair(pkg, PEROV_PARAMS, "fp64")
    a = [wl.perovskite(1, 1, 2, seed=11), wl.cluster(20, seed=12), wl.perovskite(2, 1, 1, seed=13)]
    b = [wl.cluster(16, seed=14), wl.perovskite(1, 1, 2, seed=15)]
    for ss in (a, a, a, b, b, b, a):
        Eg, Fg, Lg = g.energy_and_forces_batch(ss)
        Ed, Fd, Ld = d.energy_and_forces_batch(ss)
        assert np.array_equal(Eg, Ed)
        for x, y in zip(Fg, Fd):
            assert np.array_equal(x, y)
        for x, y in zip(Lg, Ld):
            assert (x is None and y is None) or np.array_equal(x, y)
    c, r = g.graph_counts()
    assert c == 2 and r >= 4, (c, r)


def test_capacity_retry_and_switch_off(pkg, wl):
    """a crowded cell raises the neighbour capacity inside a replayed shape
    (the replay reports it, the evaluation is redone directly with room for
    it), and set_graphs(0) drops the graphs"""
    g, d = _pair(pkg, PEROV_PARAMS, "mixed")
    s0 = wl.perovskite(2, 2, 2, seed=21)
    for k in range(3):
        _same(g.energy_and_forces(_moved(s0, k)), d.energy_and_forces(_moved(s0, k)))
    assert g.graph_counts() == (1, 2)
    crowded = s0.copy()
    crowded.pos = s0.pos.copy()
    crowded.pos[:40] = s0.pos[0] + np.random.default_rng(5).normal(0.0, 2.5, (40, 3))
    rg, rd = g.energy_and_forces(crowded), d.energy_and_forces(crowded)
    assert np.isfinite(rd[0]) and np.isfinite(rd[1]).all()
    _same(rg, rd)
    assert g.graph_counts() == (1, 3)        # the replay that found the list too small
    g.graphs = False
    assert not g.graphs
    _same(g.energy_and_forces(s0), d.energy_and_forces(s0))
    assert g.graph_counts()[1] >= 1


def test_small_evaluation_latency(pkg, wl):
    """per-call latency of the 8-atom example and a 6^3 perovskite cell
    (2,592 atoms, device arrays), replayed and direct (printed; the bounds are
    loose so that a slow box does not fail the parity suite)"""
    import torch
    out = {}
    s = wl.methylammonium()
    for on in (True, False):
        h = pkg.parseRuNNerInput(MA_PARAMS, "fp64")
        h.graphs = on
        for _ in range(5):
            h.energy_and_forces(s)
        t = []
        for _ in range(200):
            a = time.perf_counter()
            h.energy_and_forces(s)
            t.append(time.perf_counter() - a)
        out[("ma", on)] = float(np.median(t)) * 1e6
    dev = torch.device("cuda:0")
    c = wl.perovskite(6, 6, 6)
    Z = torch.from_numpy(c.Z.astype(np.int32)).to(dev)
    pos = torch.from_numpy(c.pos).to(dev)
    f = torch.empty_like(pos)
    for on in (True, False):
        h = pkg.parseRuNNerInput(PEROV_PARAMS, "mixed")
        h.graphs = on
        for _ in range(5):
            h.evaluate_torch(Z, pos, True, c.lat, f)
        t = []
        for _ in range(100):
            a = time.perf_counter()
            h.evaluate_torch(Z, pos, True, c.lat, f)
            t.append(time.perf_counter() - a)
        out[("perov6", on)] = float(np.median(t)) * 1e6
    print("median us per call:", {f"{k[0]} graphs={k[1]}": round(v, 1) for k, v in out.items()}, flush=True)
    assert out[("ma", True)] < 2000 and out[("perov6", True)] < 5000


@pytest.mark.parametrize("graphs", [False, True])
def test_stage_timings_opt_in(pkg, wl, graphs):
    """hdnnp_set_stage_timings: off by default (timings all zero, results the
    same bit for bit), on: positive stage times (a replay puts its whole time
    in the first), off again: zeros"""
    s0 = wl.perovskite(2, 2, 1, seed=3)
    h = pkg.parseRuNNerInput(PEROV_PARAMS, "mixed")
    h.graphs = graphs
    ref = h.energy_and_forces(s0)
    assert np.array_equal(h.timings(), np.zeros(6))
    h.set_stage_timings(True)
    for k in range(4):                            # a capture and replays when graphs are on
        _same(h.energy_and_forces(s0), ref)
        t = h.timings()
        print(f"graphs={graphs} call {k}: stage ms {np.round(t, 4).tolist()}", flush=True)
        assert np.all(t >= 0) and t[0] > 0
        if not h.graph_counts()[1]:
            assert np.all(t[1:5] > 0)
    assert (h.graph_counts()[1] > 0) == graphs
    h.set_stage_timings(False)
    _same(h.energy_and_forces(s0), ref)
    assert np.array_equal(h.timings(), np.zeros(6))
