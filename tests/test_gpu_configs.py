"""The batched and served workloads at their stated sizes.

Batch: 4,096 independent structures in one evaluation, half perovskite
supercells n1 x n2 x n3 (n in 1..4, 12-768 atoms), half 8-64 atom C/N/H
clusters (441,635 atoms in all). A seeded sample of 64 structures equals
their evaluations alone bit for bit in both forms; a structure with golden
reference values placed inside the batch keeps them (FP64: E 1e-10
relative, F 1e-9 Ha/bohr, dEdLat 1e-9 max(1, |dEdLat|)); mixed precision
stays within force RMSE 1e-5 Ha/bohr, |dE|/nat 1e-5 Ha and stress
1e-5 Ha/bohr * rc * nat / V (8.6e-7 Ha/bohr^3 at the perovskite density) of
FP64 for either kind.

Serving: 64 MD sessions of 1,536-atom perovskite cells (8x4x4, 1-bohr skin)
step while a stream of evaluations with log-uniform 8-20,000 atoms, half
periodic, is submitted. Every evaluation equals its evaluation alone and two
sessions follow their trajectories alone, bit for bit.
"""
from __future__ import annotations

import threading

import numpy as np
import pytest

from conftest import PEROV_MD_PARAMS, PEROV_PARAMS, golden, golden_structure

pytestmark = pytest.mark.gpu

DT = 0.5 * 41.34137333519408
KT_300 = 300.0 * 3.166811563e-6


@pytest.fixture(scope="module")
def batch(wl):
    ss = wl.batch_mix(4096, seed=2024)
    d = golden("eval_perov2")
    ss[1000] = golden_structure(wl, d)        # an even slot: a periodic one is replaced
    return ss, d


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_4096_sample_equals_single(pkg, batch, precision):
    ss, d = batch
    h = pkg.parseRuNNerInput(PEROV_PARAMS, precision)
    print(f"batch of {len(ss)} structures, {sum(s.nat for s in ss)} atoms, {precision}", flush=True)
    E, F, L = h.energy_and_forces_batch(ss)
    sample = np.random.default_rng(9).choice(len(ss), 64, replace=False)
    single = pkg.parseRuNNerInput(PEROV_PARAMS, precision)
    for k in sorted(set(sample.tolist()) | {1000}):
        e, f, l = single.energy_and_forces(ss[k])
        assert E[k] == e and np.array_equal(F[k], f), k
        if ss[k].periodic:
            assert np.array_equal(L[k], l), k
    if precision == "fp64":
        Eref = float(d["energy"])
        assert abs(E[1000] - Eref) <= 1e-10 * max(1.0, abs(Eref))
        np.testing.assert_allclose(F[1000], d["force"], rtol=0, atol=1e-9)
        np.testing.assert_allclose(L[1000], d["dEdLat"], rtol=0, atol=1e-9 * max(1.0, np.abs(d["dEdLat"]).max()))


def test_batch_4096_mixed_error(pkg, batch):
    ss, _ = batch
    E64, F64, L64 = pkg.parseRuNNerInput(PEROV_PARAMS, "fp64").energy_and_forces_batch(ss)
    E32, F32, L32 = pkg.parseRuNNerInput(PEROV_PARAMS, "mixed").energy_and_forces_batch(ss)
    for periodic in (True, False):
        ks = [k for k, s in enumerate(ss) if bool(s.periodic) == periodic]
        dF = np.concatenate([(F32[k] - F64[k]).ravel() for k in ks])
        rmse = float(np.sqrt(np.mean(dF ** 2)))
        de = max(abs(E32[k] - E64[k]) / ss[k].nat for k in ks)
        # stress against its bound 1e-5 Ha/bohr * rc * nat / V, as a ratio
        ds = max(np.abs(pkg.dEdLat2Stress(ss[k].lat, L32[k]) - pkg.dEdLat2Stress(ss[k].lat, L64[k])).max()
                 / (1e-5 * 12.0 * ss[k].nat / abs(np.linalg.det(ss[k].lat))) for k in ks) if periodic else 0.0
        print(f"periodic {periodic}: {len(ks)} structures, F rmse {rmse:.3e} max {np.abs(dF).max():.3e}, "
              f"max dE/nat {de:.3e}, stress diff / bound {ds:.3f}", flush=True)
        assert rmse <= 1e-5 and de <= 1e-5 and ds <= 1.0


@pytest.mark.parametrize("precision", ["mixed"])
def test_serving_64_sessions_and_stream(pkg, wl, precision):
    rng = np.random.default_rng(12)
    cells = [wl.perovskite(8, 4, 4, seed=900 + k) for k in range(64)]
    moms = [np.sqrt(pkg.atomic_masses(s.Z) * KT_300)[:, None] * rng.standard_normal((s.nat, 3)) for s in cells]
    requests = wl.serving_requests(40, seed=5)
    assert min(s.nat for s in requests) < 64 and max(s.nat for s in requests) > 10000
    assert {bool(s.periodic) for s in requests} == {True, False}
    rounds, steps = 3, 4
    watched = (0, 37)

    sv = pkg.Server(PEROV_MD_PARAMS, precision, lanes=4, max_batch_atoms=65536)
    try:
        sessions = [sv.md_open(s, p, DT, skin=1.0) for s, p in zip(cells, moms)]
        print(f"{len(sessions)} sessions of {cells[0].nat} atoms open; {len(requests)} requests of "
              f"{min(s.nat for s in requests)}-{max(s.nat for s in requests)} atoms", flush=True)
        md_tickets = {sid: [] for sid in sessions}
        errors = []

        def md_client(sid):
            try:
                for _ in range(rounds):
                    md_tickets[sid].append(sv.md_step(sid, steps))
            except Exception as e:
                errors.append(repr(e))

        threads = [threading.Thread(target=md_client, args=(sid,)) for sid in sessions]
        for t in threads:
            t.start()
        tickets = []
        for s in requests:                       # a Poisson stream in arrival order
            tickets.append(sv.submit(s))
            threading.Event().wait(rng.exponential(0.002))
        for t in threads:
            t.join()
        assert not errors, errors
        got = [sv.result(t, timeout=300) for t in tickets]
        md_got = {}
        for q in watched:
            sid = sessions[q]
            res = [sv.md_result(t, timeout=300) for t in md_tickets[sid]]
            md_got[q] = (res, sv.md_get(sid))
        for q, sid in enumerate(sessions):
            if q not in watched:              # a ticket is waited for once
                for t in md_tickets[sid]:
                    sv.md_result(t, timeout=300)
        st = sv.stats()
        print({k: st[k] for k in ("evaluations", "eval_atoms", "md_requests", "md_steps", "batches", "mean_batch",
                                  "max_busy_lanes", "eval_p50_ms", "eval_p99_ms", "md_p50_ms", "md_p99_ms",
                                  "peak_reserved_bytes")}, flush=True)
        for sid in sessions:
            sv.md_close(sid)
    finally:
        sv.close()      # the lanes stop even when a check above fails
    assert st["evaluations"] == len(requests)
    assert st["md_requests"] == 64 * rounds and st["md_steps"] == 64 * rounds * steps

    h = pkg.parseRuNNerInput(PEROV_MD_PARAMS, precision)
    for k, s in enumerate(requests):
        E0, F0, D0 = h.energy_and_forces(s)
        E, F, D = got[k][:3]
        assert E == E0 and np.array_equal(F, F0) and np.array_equal(D, D0), (k, s.nat)
    for q in watched:
        hq = pkg.parseRuNNerInput(PEROV_MD_PARAMS, precision)
        m = pkg.MdSystems(hq, [cells[q]], [moms[q]], DT, skin=1.0)
        alone = [m.step(steps) for _ in range(rounds)]
        res, (pos, mom, f) = md_got[q]
        assert [r[0] for r in res] == [a[0][0] for a in alone], q
        assert [r[1] for r in res] == [a[1][0] for a in alone], q
        pos0, mom0, f0 = m.get()
        assert np.array_equal(pos, pos0) and np.array_equal(mom, mom0) and np.array_equal(f, f0), q
        m.close()
