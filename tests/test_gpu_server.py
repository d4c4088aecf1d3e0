"""Request server (hdnnp_server_*): continuous batching of evaluations and MD
sessions from several client threads. Every result equals the one obtained
alone (a context of its own, no other work) bit for bit; admission keeps the
reserved device memory within the budget; an error stays with its request."""
from __future__ import annotations

import threading

import numpy as np
import pytest

from conftest import PEROV_PARAMS, committee_member

pytestmark = pytest.mark.gpu

DT = 0.5 * 41.34137333519408


def _mix(wl, n, seed, nmax_cell=3):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        sub = int(rng.integers(1 << 30))
        if k % 2 == 0:
            c = rng.integers(1, nmax_cell + 1, size=3)
            out.append(wl.perovskite(int(c[0]), int(c[1]), int(c[2]), seed=sub))
        else:
            out.append(wl.cluster(int(rng.integers(8, 65)), seed=sub))
    return out


def _alone(pkg, precision, structures):
    h = pkg.parseRuNNerInput(PEROV_PARAMS, precision)
    return [h.energy_and_forces(s) for s in structures]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_server_mix_equals_alone(pkg, wl, precision):
    """40 evaluations submitted by three client threads while two MD
    sessions step: each evaluation equals its evaluation alone, each session
    follows its trajectory alone, and the lanes ran at the same time."""
    structs = _mix(wl, 40, seed=11)
    want = _alone(pkg, precision, structs)
    rng = np.random.default_rng(3)
    md_structs = [wl.perovskite(2, 2, 1, seed=41), wl.perovskite(1, 2, 2, seed=42)]
    md_p = [rng.standard_normal((s.nat, 3)) * 2.0 for s in md_structs]
    md_alone = []
    for s, p in zip(md_structs, md_p):
        h = pkg.parseRuNNerInput(PEROV_PARAMS, precision)
        m = pkg.MdSystems(h, [s], [p], DT, skin=1.0)
        e = [m.step(3) for _ in range(8)]
        md_alone.append(([x[0][0] for x in e], [x[1][0] for x in e], m.get()))

    sv = pkg.Server(PEROV_PARAMS, precision, lanes=3, max_batch_atoms=700)
    sessions = [sv.md_open(s, p, DT, skin=1.0) for s, p in zip(md_structs, md_p)]
    tickets = [None] * len(structs)

    def client(part):
        for k in part:
            tickets[k] = sv.submit(structs[k])

    md_tickets = [[], []]
    threads = [threading.Thread(target=client, args=(range(c, len(structs), 3),)) for c in range(3)]
    for t in threads:
        t.start()
    for _ in range(8):
        for q, sid in enumerate(sessions):
            md_tickets[q].append(sv.md_step(sid, 3))
    for t in threads:
        t.join()
    for k, t in enumerate(tickets):
        E, F, D, sd, lat = sv.result(t, timeout=120)
        E0, F0, D0 = want[k]
        assert E == E0 and np.array_equal(F, F0) and np.array_equal(D, D0), k
        assert sd == 0.0 and lat > 0
    for q, sid in enumerate(sessions):
        res = [sv.md_result(t, timeout=120) for t in md_tickets[q]]
        ep, ek, (pos0, mom0, f0) = md_alone[q]
        assert [r[0] for r in res] == ep and [r[1] for r in res] == ek
        pos, mom, f = sv.md_get(sid)
        assert np.array_equal(pos, pos0) and np.array_equal(mom, mom0) and np.array_equal(f, f0)
        sv.md_close(sid)
    st = sv.stats()
    print({k: st[k] for k in ("evaluations", "batches", "mean_batch", "max_busy_lanes", "eval_p50_ms",
                              "eval_p99_ms", "md_p50_ms", "md_p99_ms", "peak_reserved_bytes")}, flush=True)
    assert st["evaluations"] == len(structs) and st["md_requests"] == 16 and st["md_steps"] == 48
    assert st["max_busy_lanes"] >= 2
    assert st["eval_p50_ms"] <= st["eval_p99_ms"]
    sv.close()


def test_server_budget_admission(pkg, wl):
    """A small budget: requests wait for memory instead of failing, the
    reserved bytes never exceed the budget, requests and sessions larger than
    the whole budget are refused, and the results are unchanged."""
    budget = 28 << 20
    structs = [wl.perovskite(4, 4, 3, seed=60 + k) for k in range(6)] + _mix(wl, 10, seed=12)
    want = _alone(pkg, "mixed", structs)
    sv = pkg.Server(PEROV_PARAMS, "mixed", lanes=4, hbm_budget=budget, max_batch_atoms=600)
    tickets = [sv.submit(s) for s in structs]
    big = wl.perovskite(16, 16, 8, seed=1)
    with pytest.raises(pkg.HdnnpError) as ei:
        sv.submit(big)
    assert ei.value.code == 6 and "budget" in str(ei.value)
    with pytest.raises(pkg.HdnnpError) as ei:
        sv.md_open(big, np.zeros((big.nat, 3)), DT)
    assert ei.value.code == 6
    for k, t in enumerate(tickets):
        E, F, D, _, _ = sv.result(t, timeout=120)
        assert E == want[k][0] and np.array_equal(F, want[k][1]) and np.array_equal(D, want[k][2])
    st = sv.stats()
    print({k: st[k] for k in ("batches", "max_busy_lanes", "peak_reserved_bytes", "refused")}, flush=True)
    assert 0 < st["peak_reserved_bytes"] <= budget
    assert st["refused"] == 2
    sv.close()


def test_server_error_stays_with_its_request(pkg, wl):
    """An element the potential lacks fails that request only; the others
    of its batch are evaluated as alone."""
    structs = _mix(wl, 6, seed=13)
    want = _alone(pkg, "fp64", structs)
    bad = wl.cluster(20, seed=5)
    bad.Z = bad.Z.copy()
    bad.Z[3] = 79
    sv = pkg.Server(PEROV_PARAMS, "fp64", lanes=1)
    ts = [sv.submit(s) for s in structs[:3]] + [sv.submit(bad)] + [sv.submit(s) for s in structs[3:]]
    with pytest.raises(pkg.HdnnpError) as ei:
        sv.result(ts[3], timeout=120)
    assert ei.value.code == 4 and "Z=79" in str(ei.value)
    for k, t in enumerate(ts[:3] + ts[4:]):
        E, F, D, _, _ = sv.result(t, timeout=120)
        assert E == want[k][0] and np.array_equal(F, want[k][1]) and np.array_equal(D, want[k][2])
    with pytest.raises(pkg.HdnnpError):
        sv.result(ts[0])                      # a ticket is taken once
    sv.close()


def test_server_cancel(pkg, wl):
    """Queued requests cancelled behind a long one never run: their waits
    raise HDNNP_ERR_CANCELLED; every request not cancelled (cancel returned
    False because a lane had it, or never asked) equals its evaluation
    alone; a cancelled MD-step request leaves the session stepping from the
    state without it; a finished ticket cannot be cancelled, an unknown one
    is an argument error."""
    big = [wl.perovskite(8, 8, 8, seed=31 + k) for k in range(3)]   # ~6,000 atoms each, one batch each
    small = _mix(wl, 8, seed=19)
    structs = big + small
    want = _alone(pkg, "mixed", structs)
    sv = pkg.Server(PEROV_PARAMS, "mixed", lanes=1, max_batch_atoms=4096)
    ts = [sv.submit(s) for s in structs]
    cancelled = {k: sv.cancel(ts[k]) for k in (1, 2, 4, 6, 8, 10)}
    print("cancelled", cancelled, flush=True)
    assert any(cancelled.values())
    for k, t in enumerate(ts):
        if cancelled.get(k):
            with pytest.raises(pkg.HdnnpError) as ei:
                sv.result(t, timeout=120)
            assert ei.value.code == pkg.HDNNP_ERR_CANCELLED
        else:
            E, F, D, _, _ = sv.result(t, timeout=120)
            assert E == want[k][0] and np.array_equal(F, want[k][1]) and np.array_equal(D, want[k][2]), k
    t = sv.submit(small[0])
    sv.result(t, timeout=120)
    with pytest.raises(pkg.HdnnpError) as ei:
        sv.cancel(t)                          # waited for: the ticket is gone
    assert ei.value.code == 1
    # MD: a session behind a long evaluation, three queued requests of 2 steps
    cell = wl.perovskite(2, 2, 2, seed=40)
    rng = np.random.default_rng(4)
    mom = np.sqrt(pkg.atomic_masses(cell.Z) * 300.0 * 3.166811563e-6)[:, None] * rng.standard_normal((cell.nat, 3))
    sid = sv.md_open(cell, mom, DT, skin=1.0)
    sv.md_result(sv.md_step(sid, 0), timeout=120)
    tb = sv.submit(big[0])
    steps = [sv.md_step(sid, 2) for _ in range(3)]
    gone = [sv.cancel(steps[1])]
    print("MD step cancelled", gone, flush=True)
    kept = [steps[0]] + ([] if gone[0] else [steps[1]]) + [steps[2]]
    got = [sv.md_result(x, timeout=120) for x in kept]
    if gone[0]:
        with pytest.raises(pkg.HdnnpError) as ei:
            sv.md_result(steps[1], timeout=120)
        assert ei.value.code == pkg.HDNNP_ERR_CANCELLED
    sv.result(tb, timeout=120)
    pos, momf, f = sv.md_get(sid)
    h = pkg.parseRuNNerInput(PEROV_PARAMS, "mixed")
    m = pkg.MdSystems(h, [cell], [mom], DT, skin=1.0)
    m.step(0)
    alone = [m.step(2) for _ in kept]
    assert [g[0] for g in got] == [a[0][0] for a in alone]
    pos0, mom0, f0 = m.get()
    assert np.array_equal(pos, pos0) and np.array_equal(momf, mom0) and np.array_equal(f, f0)
    m.close()
    sv.md_close(sid)
    sv.close()


def test_server_failed_session_stays_with_itself(pkg, wl):
    """A session whose structure has an element the potential lacks fails
    its own requests only; a sound session and evaluations beside it run as
    alone, and the failed session closes."""
    bad = wl.cluster(20, seed=6)
    bad.Z = bad.Z.copy()
    bad.Z[2] = 79
    good = wl.cluster(24, seed=7)
    ev = _mix(wl, 3, seed=17)
    want = _alone(pkg, "fp64", ev)
    h = pkg.parseRuNNerInput(PEROV_PARAMS, "fp64")
    p = np.zeros((good.nat, 3))
    md = pkg.MdSystems(h, [good], [p], DT, skin=1.0)
    want_md = md.step(3)
    md.close()
    sv = pkg.Server(PEROV_PARAMS, "fp64", lanes=2)
    sb = sv.md_open(bad, np.zeros((bad.nat, 3)), DT, skin=1.0, masses=np.ones(bad.nat))
    sg = sv.md_open(good, p, DT, skin=1.0)
    ts = [sv.submit(s) for s in ev]
    with pytest.raises(pkg.HdnnpError) as ei:
        sv.md_result(sv.md_step(sb, 2), timeout=120)
    assert ei.value.code == 4 and "Z=79" in str(ei.value)
    ep, ek, _ = sv.md_result(sv.md_step(sg, 3), timeout=120)
    assert ep == want_md[0][0] and ek == want_md[1][0]
    for k, t in enumerate(ts):
        E, F, D, _, _ = sv.result(t, timeout=120)
        assert E == want[k][0] and np.array_equal(F, want[k][1]) and np.array_equal(D, want[k][2])
    sv.md_close(sb)
    sv.md_close(sg)
    sv.close()


def test_server_committee(pkg, wl, tmp_path):
    """A server of two potentials returns the committee mean and spread of a
    committee context."""
    m2 = str(tmp_path / "m2") + "/"
    committee_member(PEROV_PARAMS, m2, seed=7)
    structs = _mix(wl, 6, seed=14)
    h = pkg.Hdnnp([PEROV_PARAMS, m2], "fp64")
    sv = pkg.Server([PEROV_PARAMS, m2], "fp64", lanes=2)
    ts = [sv.submit(s) for s in structs]
    for s, t in zip(structs, ts):
        E0, F0, D0 = h.energy_and_forces(s)
        sd0 = h.committee(1)[0][0]
        E, F, D, sd, _ = sv.result(t, timeout=120)
        assert E == E0 and np.array_equal(F, F0) and np.array_equal(D, D0) and sd == sd0 and sd > 0
    sv.close()


def test_server_ingests_structure_files(pkg):
    """Structures read from the reference's file formats (input.data,
    POSCAR, FHI-aims, v_sim ascii) are evaluated by the server as by a
    context of their own."""
    import os
    from conftest import GOLD
    rd, d = pkg.readers, os.path.join(GOLD, "readers")
    ss = (rd.read_structures(os.path.join(d, "input.data"), "data") + [rd.read_poscar(os.path.join(d, "POSCAR"))]
          + [rd.read_fhiaims(os.path.join(d, "geometry.in")), rd.read_ascii(os.path.join(d, "cell_ang.ascii"))])
    want = _alone(pkg, "fp64", ss)
    sv = pkg.Server(PEROV_PARAMS, "fp64", lanes=2)
    ts = [sv.submit(s) for s in ss]
    for k, t in enumerate(ts):
        E, F, D, _, _ = sv.result(t, timeout=120)
        assert E == want[k][0] and np.array_equal(F, want[k][1]) and np.array_equal(D, want[k][2])
    sv.close()
