"""Benchmark: HDNNP energy + forces + stress throughput on one MI355X.

Workload (default): synthetic pseudocubic ABX3 perovskite, 16x16x16 supercell
= 49,152 atoms (H, C, N, I, Pb; a = 11.9 bohr, seeded jitter 0.2 bohr),
5-element potential with 40 G2 + 40 G4 per element (tanh^3 cutoff 12 bohr),
MLP 80-25-25-1. One step = one full evaluation (neighbour list, symmetry
functions, MLP forward/backward, forces, dE/dh) of the whole cell, with
positions and forces resident in device memory.

Both arithmetic modes are timed. `value` is the mixed-precision rate when its
force RMSE against the FP64 path is <= 1e-5 Ha/bohr, else the FP64 rate.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
    python bench.py --calibrate-copy   # plain 512 MiB device copy, for HBM counter calibration
    python bench.py --workload batch   # 4096 independent structures per evaluation
    python bench.py --workload serve   # request server: Poisson stream + 64 MD sessions
    python bench.py --workload molecule   # the example's 8-atom molecule: call latency, MD step

`roofline.traffic` (HBM bytes per evaluation of the three symmetry-function /
MLP kernels) is read from the rocprofv3 FETCH_SIZE / WRITE_SIZE summary in
profiles/, where it was measured for this workload; the counters cannot be
read from inside the timed run.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "high-dimensional-neural-network-potential_amd")
PARAMS = os.path.join(PKG_DIR, "data", "perovskite")
# same symmetry functions and network shapes with repulsive weights: velocity
# Verlet on it conserves energy (on the random weights it does not)
PARAMS_MD = os.path.join(PKG_DIR, "data", "perovskite_md")
# the same potential with the cosine cutoff (cutoff_type 1), its scaling from the oracle
PARAMS_COS = os.path.join(PKG_DIR, "data", "perovskite_cos")
PEAK_TFLOPS = {"fp64": 78.6, "mixed": 157.3}      # MI355X FP64 / FP32 vector peaks
FORCE_RMSE_BOUND = 1e-5                           # Ha/bohr
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "perovskite16_hbm_traffic.txt")


def measured_traffic(cells, mode):
    """HBM bytes per evaluation of k_sf_values + k_mlp + k_sf_grad from the
    committed counter summary (line `# traffic_per_evaluation_bytes <mode> <N>`),
    or None when the summary is absent or for another workload."""
    if cells != 16 or not os.path.exists(TRAFFIC_PROFILE):
        return None
    with open(TRAFFIC_PROFILE) as f:
        for ln in f:
            p = ln.split()
            if len(p) == 4 and p[1] == "traffic_per_evaluation_bytes" and p[2] == mode:
                return float(p[3])
    return None


def calibrate_copy():
    """A plain device-to-device copy of 512 MiB, three times: the known byte
    count against which FETCH_SIZE / WRITE_SIZE are calibrated."""
    import torch
    n = 64 << 20
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied 3 x {n * 8} bytes", flush=True)


def load_package():
    spec = importlib.util.spec_from_file_location(
        "hdnnp_amd", os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["hdnnp_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def algorithmic_flop(h, s, ent):
    """Frozen work figure of one evaluation (value + gradient):
    20 N_pair + 20 N_G2 + 50 N_trip + 40 N_G4 + nat * 4 * sum(nIn*nOut),
    integer counts over the rc-neighbour list (self excluded)."""
    nel = h.info.nelements
    zs = h.elements
    rank = {z: e for e, z in enumerate(zs)}
    el = np.array([rank[int(z)] for z in s.Z])
    cnt = np.zeros((s.nat, nel), dtype=np.int64)
    np.add.at(cnt, (ent[:, 0], el[ent[:, 1]]), 1)
    n_i = cnt.sum(axis=1)
    n_pair = int(n_i.sum())
    n_trip = int((n_i * (n_i - 1) // 2).sum())
    n_g2 = 0
    n_g4 = 0
    for e in range(nel):
        tab = h.sf_table(e)
        rows = cnt[el == e]
        for f in range(nel):
            g2 = int(np.sum((tab[:, 0] == 2) & (tab[:, 1] == zs[f])))
            n_g2 += g2 * int(rows[:, f].sum())
            for g in range(f, nel):
                m = (tab[:, 0] == 3) & (((tab[:, 1] == zs[f]) & (tab[:, 2] == zs[g])) |
                                        ((tab[:, 1] == zs[g]) & (tab[:, 2] == zs[f])))
                k = int(np.sum(m))
                if not k:
                    continue
                pairs = rows[:, f] * (rows[:, f] - 1) // 2 if f == g else rows[:, f] * rows[:, g]
                n_g4 += k * int(pairs.sum())
    nodes = [h.info.nodes[l] for l in range(h.info.nlayers)]
    sizes = [int(h.info.nsf[0])] + nodes
    mlp = s.nat * 4 * sum(a * b for a, b in zip(sizes[:-1], sizes[1:]))
    flop = 20 * n_pair + 20 * n_g2 + 50 * n_trip + 40 * n_g4 + mlp
    return flop, dict(N_pair=n_pair, N_G2=n_g2, N_trip=n_trip, N_G4=n_g4, mlp=mlp)


def host_cores():
    """The cores this process may use: its affinity mask, limited by the
    cgroup's CPU quota where one is set (cpu.max), with nproc and the CPU
    model for the record."""
    affinity = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
        if q != "max":
            quota = max(1, int(int(q) // int(period)))
    except (OSError, ValueError):
        pass
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = min(affinity, quota) if quota else affinity
    return usable, {"nproc": os.cpu_count(), "affinity": affinity, "cgroup_quota_cores": quota, "cpu_model": model}


def cpu_baseline(s, nrep=5):
    """The reference (oracle/_ref/refdump, built from the reference's sources)
    timed on every core this host gives the process (affinity and cgroup
    quota); falls back to the oracle's C restatement."""
    import resource
    threads, host = host_cores()
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_STACKSIZE="64M")

    def big_stack():
        # the reference keeps nat-sized automatic arrays on the stack
        soft, hard = resource.getrlimit(resource.RLIMIT_STACK)
        want = 8 << 30
        if hard != resource.RLIM_INFINITY:
            want = min(want, hard)
        resource.setrlimit(resource.RLIMIT_STACK, (want, hard))

    ref = os.path.join(ROOT, "oracle", "_ref", "refdump")
    with tempfile.TemporaryDirectory() as td:
        sf = os.path.join(td, "s.txt")
        sys.modules["hdnnp_amd"].workloads.write_structure(sf, s)
        if os.access(ref, os.X_OK):
            out = os.path.join(td, "t.txt")
            subprocess.run([ref, "time", PARAMS + "/", sf, str(nrep), out], check=True, env=env,
                           stdout=subprocess.DEVNULL, timeout=900, preexec_fn=big_stack)
            with open(out) as f:
                secs = [float(ln.split()[1]) for ln in f if ln.startswith("seconds")]
            kind = "reference"
        else:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import runner
            o = runner.Oracle(runner.read_potential(PARAMS))
            o.evaluate(s, nthreads=threads)
            secs = []
            for _ in range(nrep):
                t0 = time.perf_counter()
                o.evaluate(s, nthreads=threads)
                secs.append(time.perf_counter() - t0)
            kind = "port"
    t = float(np.median(secs))
    return {"value": s.nat / t, "unit": "atoms/s", "cores": threads, "kind": kind,
            "sample": f"the same {s.nat}-atom cell (tanh^3 cutoff, the reference's only one), FP64 E+F+dE/dh, "
                      f"1 warm-up + median of {nrep} calls of hdnnpEnergyAndForces, OMP_NUM_THREADS={threads} "
                      f"on {host['cpu_model']}",
            "seconds_per_eval": t, "seconds_all": secs, "host": host}


def run_mode(pkg, torch, s, precision, steps, warmup, params=PARAMS):
    h = pkg.parseRuNNerInput(params, precision)
    dev = torch.device("cuda:0")
    Z = torch.from_numpy(s.Z.astype(np.int32)).to(dev)
    pos = torch.from_numpy(np.ascontiguousarray(s.pos)).to(dev)
    force = torch.empty((s.nat, 3), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    for _ in range(warmup):
        h.evaluate_torch(Z, pos, True, s.lat, force)
    torch.cuda.synchronize()
    lat_ms, stage = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        a = time.perf_counter()
        E, _, L = h.evaluate_torch(Z, pos, True, s.lat, force)
        lat_ms.append((time.perf_counter() - a) * 1e3)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # the stage breakdown (and the roofline's kernel times): the same steps
    # again with HIP events between the stages, which the rate above runs
    # without (they are packets of their own on the stream)
    h.set_stage_timings(True)
    for _ in range(steps):
        h.evaluate_torch(Z, pos, True, s.lat, force)
        stage.append(h.timings())
    h.set_stage_timings(False)
    return h, dict(E=E, F=force.cpu().numpy(), dEdLat=L, wall=wall, lat_ms=np.array(lat_ms),
                   stage=np.array(stage))


def mixed_errors(pkg, s, rm, r64):
    """Errors of a mixed-precision result against the FP64 one of the same
    structure: force RMSE and max |dF| (Ha/bohr), |dE| per atom (Ha), and the
    largest stress difference, sigma = -dEdLat lat^T / det(lat) (Ha/bohr^3),
    absolute and relative to the largest FP64 stress component."""
    dF = rm["F"] - r64["F"]
    out = {"force_rmse": float(np.sqrt(np.mean(dF ** 2))), "force_max": float(np.abs(dF).max()),
           "energy_err_per_atom": abs(rm["E"] - r64["E"]) / s.nat}
    if s.periodic:
        sm = pkg.dEdLat2Stress(s.lat, rm["dEdLat"])
        s64 = pkg.dEdLat2Stress(s.lat, r64["dEdLat"])
        out["stress_err"] = float(np.abs(sm - s64).max())
        out["stress_err_rel"] = float(np.abs(sm - s64).max() / np.abs(s64).max())
        # the stress of a coherent force error at the 1e-5 Ha/bohr bound over the cutoff
        out["stress_bound"] = FORCE_RMSE_BOUND * 12.0 * s.nat / abs(np.linalg.det(np.asarray(s.lat).reshape(3, 3)))
    return out


def bench_batch(args):
    """Batched evaluation of many independent structures (half perovskite
    supercells n1 x n2 x n3, n in 1..4, half 8-64 atom C/N/H clusters): one
    hdnnp_evaluate_batch_device call per step, atoms resident in device memory
    (per-structure sizes, flags and lattices on the host)."""
    import torch
    pkg = load_package()
    ss = pkg.workloads.batch_mix(args.nstruct, seed=2024)
    nat_s = np.array([s.nat for s in ss], dtype=np.int32)
    nat = int(nat_s.sum())
    per = np.array([int(s.periodic) for s in ss], dtype=np.int32)
    lats = np.stack([np.asarray(s.lat, dtype=np.float64).reshape(3, 3) for s in ss])
    dev = torch.device("cuda:0")
    Z = torch.from_numpy(np.concatenate([s.Z for s in ss]).astype(np.int32)).to(dev)
    pos = torch.from_numpy(np.ascontiguousarray(np.concatenate([s.pos for s in ss]))).to(dev)
    force = torch.empty((nat, 3), dtype=torch.float64, device=dev)
    print(f"workload: batch of {len(ss)} structures, {nat} atoms", file=sys.stderr, flush=True)
    res = {}
    for mode in args.modes.split(","):
        h = pkg.parseRuNNerInput(PARAMS, mode)
        tile, tile_ms = h.autotune(ss)               # the context's tile, timed on this batch
        print(f"{mode}: autotuner keeps {tile} atom(s) per wavefront; SF-pass ms per tile {tile_ms}",
              file=sys.stderr, flush=True)
        for _ in range(args.warmup):
            h.evaluate_batch_torch(nat_s, Z, pos, per, lats, force)
        torch.cuda.synchronize()
        lat = []
        for _ in range(args.steps):
            a = time.perf_counter()
            E, _, L = h.evaluate_batch_torch(nat_s, Z, pos, per, lats, force)
            lat.append(time.perf_counter() - a)
        torch.cuda.synchronize()
        t = float(np.median(lat))
        h.set_stage_timings(True)                    # one more evaluation, with its stage times
        h.evaluate_batch_torch(nat_s, Z, pos, per, lats, force)
        stage = h.timings()
        h.set_stage_timings(False)
        res[mode] = {"atoms_per_s": nat / t, "structures_per_s": len(ss) / t, "ms_per_batch": t * 1e3,
                     "p99_ms": float(np.percentile(lat, 99)) * 1e3,
                     "autotune": {"atoms_per_wave": tile, "sf_ms_per_tile": tile_ms},
                     "stage_ms": np.round(stage, 4).tolist(), "F": force.cpu().numpy(), "E": E, "L": L}
        print(f"{mode}: {t * 1e3:.3f} ms per batch -> {nat / t:.4g} atoms/s, {len(ss) / t:.4g} structures/s; "
              f"stage ms {res[mode]['stage_ms']}", file=sys.stderr, flush=True)
    def kind_errors(mode):
        out = {}
        off = np.concatenate([[0], np.cumsum(nat_s)])
        for kind, periodic in (("periodic", True), ("cluster", False)):
            ks = [k for k in range(len(ss)) if bool(ss[k].periodic) == periodic]
            rows = np.concatenate([np.arange(off[k], off[k + 1]) for k in ks])
            dF = res[mode]["F"][rows] - res["fp64"]["F"][rows]
            e = {"structures": len(ks), "force_rmse": float(np.sqrt(np.mean(dF ** 2))),
                 "force_max": float(np.abs(dF).max()),
                 "energy_err_per_atom_max": float(max(abs(res[mode]["E"][k] - res["fp64"]["E"][k]) / ss[k].nat
                                                      for k in ks))}
            if periodic:
                ds = [np.abs(pkg.dEdLat2Stress(ss[k].lat, res[mode]["L"][k]) -
                             pkg.dEdLat2Stress(ss[k].lat, res["fp64"]["L"][k])).max() for k in ks]
                e["stress_err_max"] = float(max(ds))
            out[kind] = e
        return out

    rmse = None
    per_kind = None
    if "fp64" in res and "mixed" in res:
        rmse = float(np.sqrt(np.mean((res["mixed"]["F"] - res["fp64"]["F"]) ** 2)))
        per_kind = kind_errors("mixed")
        print(f"mixed vs fp64 per kind: {per_kind}", file=sys.stderr, flush=True)
    bf16 = None
    if "fp64" in res and "mixed-bf16" in res:
        bf16 = {"force_rmse": float(np.sqrt(np.mean((res["mixed-bf16"]["F"] - res["fp64"]["F"]) ** 2))),
                "per_kind": kind_errors("mixed-bf16")}
        print(f"mixed-bf16 vs fp64: {bf16}", file=sys.stderr, flush=True)
    head = "mixed" if ("mixed" in res and rmse is not None and rmse <= FORCE_RMSE_BOUND) else next(iter(res))
    out = {"metric": "atoms/s (E+F+stress), batched independent structures", "value": res[head]["atoms_per_s"],
           "unit": "atoms/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": res[head]["ms_per_batch"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64" if head == "fp64" else "f32 (symmetry functions, MLP); f64 (wrap, list, sums)",
           "data": "synthetic (seeded structures, seeded random-weight potential)",
           "config": {"workload": f"batch of {len(ss)} structures ({nat} atoms): perovskite n1xn2xn3 (n in 1..4) "
                                  f"and 8-64 atom clusters, atoms resident in device memory", "precision": head},
           "force_rmse_mixed_vs_fp64": rmse, "mixed_vs_fp64": per_kind, "bf16_mlp_vs_fp64": bf16,
           "modes": {m: {k: v for k, v in r.items() if k not in ("F", "E", "L")} for m, r in res.items()}}
    print(json.dumps(out), flush=True)


def bench_molecule(args):
    """The reference example's 8-atom methylammonium molecule with its own
    potential (the example's RuNNer files): latency of one host-array
    evaluation (hdnnp_evaluate) and time per velocity-Verlet step on the
    device (dt 0.5 fs, 140 K), without and with a 1-bohr Verlet skin."""
    import torch
    torch.cuda.is_available()
    pkg = load_package()
    s = pkg.workloads.methylammonium()
    params = os.path.join(ROOT, "tests", "golden", "methylammonium")
    rng = np.random.default_rng(3)
    kT = 140.0 * 3.166811563e-6
    mom = np.sqrt(pkg.atomic_masses(s.Z) * kT)[:, None] * rng.standard_normal((s.nat, 3))
    res = {}
    for mode in args.modes.split(","):
        h = pkg.parseRuNNerInput(params, mode)
        for _ in range(max(args.warmup, 20)):
            h.energy_and_forces(s)
        lat = []
        for _ in range(max(args.steps, 200)):
            a = time.perf_counter()
            h.energy_and_forces(s)
            lat.append((time.perf_counter() - a) * 1e3)
        r = {"call_p50_ms": float(np.percentile(lat, 50)), "call_p99_ms": float(np.percentile(lat, 99))}
        for skin in (0.0, 1.0):
            m = pkg.MdSystems(h, [s], [mom], 0.5 * 41.34137333519408, skin=skin)
            m.step(50)
            a = time.perf_counter()
            m.step(1000)
            r[f"md_step_ms_skin{skin:g}"] = (time.perf_counter() - a)
            m.close()
        res[mode] = r
        print(f"{mode}: call p50 {r['call_p50_ms'] * 1e3:.1f} us, p99 {r['call_p99_ms'] * 1e3:.1f} us; MD step "
              f"{r['md_step_ms_skin0'] * 1e3:.1f} us (no skin), {r['md_step_ms_skin1'] * 1e3:.1f} us (1-bohr skin)",
              file=sys.stderr, flush=True)
    head = "mixed" if "mixed" in res else next(iter(res))
    out = {"metric": "latency of one E+F evaluation (host arrays)", "value": res[head]["call_p50_ms"], "unit": "ms",
           "n_gpus": 1, "steps": max(args.steps, 200), "warmup": max(args.warmup, 20),
           "ms_per_step": res[head]["call_p50_ms"], "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64" if head == "fp64" else "f32 (symmetry functions, MLP); f64 (wrap, list, sums)",
           "data": "the reference example's molecule and potential",
           "config": {"workload": "methylammonium CH3NH3+ (8 atoms, nonperiodic), the example's RuNNer potential; "
                                  "1000 MD steps of dt 0.5 fs per MD figure", "precision": head},
           "modes": res}
    print(json.dumps(out), flush=True)


def serve_window(pkg, mode, pool, md_cells, args):
    """One serving run: a Poisson stream of evaluation requests drawn from
    `pool`, and MD sessions (one client thread each, args.md_steps steps per
    request) stepping for args.seconds. Returns the server's counters."""
    import threading
    sv = pkg.Server(PARAMS_MD, mode, lanes=args.lanes, max_batch_atoms=args.max_batch_atoms)
    rng = np.random.default_rng(5)
    for t in [sv.submit(s) for s in pool[:16]]:          # warm-up
        sv.result(t)
    kT = 300.0 * 3.166811563e-6                          # 300 K in Hartree
    sessions = []
    for k, s in enumerate(md_cells):
        m = pkg.atomic_masses(s.Z)
        p = np.sqrt(m * kT)[:, None] * rng.standard_normal((s.nat, 3))
        sessions.append(sv.md_open(s, p, 0.5 * 41.34137333519408, skin=1.0))
    for sid in sessions:
        sv.md_result(sv.md_step(sid, 0))                # opened
    stop = threading.Event()
    errors = []

    def md_client(sid):
        try:
            while not stop.is_set():
                sv.md_result(sv.md_step(sid, args.md_steps))
        except Exception as e:                            # reported after the run
            errors.append(repr(e))

    pending = []
    lock = threading.Lock()
    done_submitting = threading.Event()

    def collector():
        try:
            while True:
                with lock:
                    t = pending.pop(0) if pending else None
                if t is None:
                    if done_submitting.is_set():
                        return
                    time.sleep(0.0005)
                    continue
                sv.result(t)
        except Exception as e:
            errors.append(repr(e))

    md_threads = [threading.Thread(target=md_client, args=(sid,)) for sid in sessions]
    for th in md_threads:
        th.start()
    # capacity left to evaluations while the sessions step: a standing backlog
    # of len(pool) requests (one submitted per one completed) for 3 s after
    # 0.5 s to settle. Under a backlog the MD requests run only once aged, the
    # share they keep when evaluations queue, so an offered rate below this
    # capacity drains.
    backlog = [(sv.submit(s), s.nat) for s in pool]
    a = time.perf_counter()
    t_meas, atoms, k = None, 0, 0
    while True:
        t, n = backlog.pop(0)
        sv.result(t)
        now = time.perf_counter()
        if t_meas is not None:
            atoms += n
            if now - t_meas >= 3.0:
                break
        elif now - a >= 0.5:
            t_meas = now
        s = pool[k % len(pool)]
        backlog.append((sv.submit(s), s.nat))
        k += 1
    cap = atoms / (now - t_meas)
    for t, _ in backlog:
        sv.result(t)
    mean_nat = float(np.mean([s.nat for s in pool]))
    rate = args.load * cap / mean_nat
    threads = md_threads + [threading.Thread(target=collector)]
    sv.stats(reset=True)
    t0 = time.perf_counter()
    threads[-1].start()
    nsub, t_next = 0, t0
    while True:
        t_next += rng.exponential(1.0 / rate)
        if t_next - t0 >= args.seconds:
            break
        while True:
            now = time.perf_counter()
            if now >= t_next:
                break
            time.sleep(min(t_next - now, 0.001))
        t = sv.submit(pool[int(rng.integers(len(pool)))])
        with lock:
            pending.append(t)
        nsub += 1
    stop.set()
    done_submitting.set()
    for th in threads:
        th.join()
    st = sv.stats()
    wall = time.perf_counter() - t0
    for sid in sessions:
        sv.md_close(sid)
    sv.close()
    if errors:
        raise RuntimeError(f"serving run failed: {errors[:3]}")
    st.update({"wall_s": wall, "offered_req_per_s": rate, "eval_capacity_atoms_per_s": cap, "submitted": nsub,
               "atoms_per_s": (st["eval_atoms"] + st["md_atom_steps"]) / wall,
               "eval_atoms_per_s": st["eval_atoms"] / wall, "md_atom_steps_per_s": st["md_atom_steps"] / wall})
    return st


def bench_serve(args):
    """Serving: a Poisson stream of energy + forces + stress requests (atom
    counts log-uniform in 8..20,000, alternately periodic perovskite supercells
    and nonperiodic C/N/H clusters) at args.load of the evaluation capacity
    left by the sessions under a standing backlog, while 64 MD sessions of 1536-atom perovskite cells (8x4x4, 300 K,
    dt 0.5 fs, 1-bohr Verlet skin, state resident on the device) step
    continuously. The potential is the repulsive variant of the synthetic one
    (data/perovskite_md: same symmetry functions and network shapes, so the
    same work per atom), on which the sessions conserve energy. Reports served
    atoms/s (evaluated atoms + MD atom-steps) with p50/p99 latency."""
    import torch
    torch.cuda.is_available()
    pkg = load_package()
    pool = pkg.workloads.serving_requests(args.pool, seed=77)
    md_cells = [pkg.workloads.perovskite(8, 4, 4, seed=900 + k) for k in range(args.sessions)]
    print(f"workload: serving, pool of {len(pool)} requests (mean {np.mean([s.nat for s in pool]):.0f} atoms), "
          f"{len(md_cells)} MD sessions x {md_cells[0].nat if md_cells else 0} atoms", file=sys.stderr, flush=True)
    res = {}
    for mode in args.modes.split(","):
        st = serve_window(pkg, mode, pool, md_cells, args)
        res[mode] = st
        print(f"{mode}: {st['atoms_per_s']:.4g} atoms/s served ({st['evaluations']:.0f} evaluations, "
              f"p50 {st['eval_p50_ms']:.2f} ms, p99 {st['eval_p99_ms']:.2f} ms; {st['md_steps']:.0f} MD steps, "
              f"request p50 {st['md_p50_ms']:.2f} ms, p99 {st['md_p99_ms']:.2f} ms; mean batch "
              f"{st['mean_batch']:.1f}; lanes busy at once {st['max_busy_lanes']:.0f})", file=sys.stderr, flush=True)
    head = "mixed" if "mixed" in res else next(iter(res))
    r = res[head]
    out = {"metric": "atoms/s (E+F+stress) served; p50/p99 latency", "value": r["atoms_per_s"],
           "unit": "atoms/s", "n_gpus": 1, "steps": int(r["evaluations"]), "warmup": 16,
           "ms_per_step": r["eval_p50_ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64" if head == "fp64" else "f32 (symmetry functions, MLP); f64 (wrap, list, sums)",
           "data": "synthetic (seeded structures, seeded random-weight potential)",
           "config": {"workload": f"Poisson stream at {args.load:.0%} of the evaluation capacity left by the sessions, "
                                  f"requests of 8-20,000 "
                                  f"atoms (log-uniform, half periodic), plus {args.sessions} MD sessions of 1536 "
                                  f"atoms ({args.md_steps} steps per request, 300 K, dt 0.5 fs, skin 1 bohr); repulsive "
                                  f"variant of the synthetic potential (same shapes)",
                      "lanes": args.lanes, "max_batch_atoms": args.max_batch_atoms, "seconds": args.seconds,
                      "precision": head},
           "latency_ms": {"eval_p50": r["eval_p50_ms"], "eval_p99": r["eval_p99_ms"],
                          "md_request_p50": r["md_p50_ms"], "md_request_p99": r["md_p99_ms"]},
           "modes": res}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cells", type=int, default=16, help="supercell edge (16 -> 49,152 atoms)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-cosine", action="store_true", help="skip the cosine-cutoff variant")
    ap.add_argument("--no-bf16", action="store_true", help="skip the BF16-MLP variant (mixed-bf16)")
    ap.add_argument("--modes", default="fp64,mixed")
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--workload", default="cell", choices=["cell", "batch", "serve", "molecule"],
                    help="cell: one perovskite supercell (headline); batch: many independent structures; "
                         "serve: request server under a Poisson stream with MD sessions; molecule: the "
                         "reference example's 8-atom molecule (call latency, MD step)")
    ap.add_argument("--nstruct", type=int, default=4096, help="structures of the batch workload")
    ap.add_argument("--pool", type=int, default=128, help="serve: distinct request structures")
    ap.add_argument("--sessions", type=int, default=64, help="serve: MD sessions")
    ap.add_argument("--md-steps", type=int, default=10, help="serve: MD steps per session request")
    ap.add_argument("--seconds", type=float, default=15.0, help="serve: length of each timed run")
    ap.add_argument("--load", type=float, default=0.6, help="serve: offered load, fraction of capacity")
    ap.add_argument("--lanes", type=int, default=4, help="serve: server lanes")
    ap.add_argument("--max-batch-atoms", type=int, default=65536, help="serve: atoms per batch")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.workload in ("batch", "serve", "molecule"):
        if args.gpus != 1:
            print("bench.py runs on one GPU only", file=sys.stderr)
            sys.exit(2)
        {"batch": bench_batch, "serve": bench_serve, "molecule": bench_molecule}[args.workload](args)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); multi-GPU sharding of this "
              f"path is design-only in this build", file=sys.stderr)
        sys.exit(2)

    import torch
    pkg = load_package()
    s = pkg.workloads.perovskite(args.cells, args.cells, args.cells)
    print(f"workload: perovskite {args.cells}^3 = {s.nat} atoms", file=sys.stderr, flush=True)
    res = {}
    handles = {}
    for mode in args.modes.split(","):
        h, r = run_mode(pkg, torch, s, mode, args.steps, args.warmup)
        handles[mode] = h
        res[mode] = r
        print(f"{mode}: {args.steps} steps in {r['wall']:.3f} s -> {s.nat * args.steps / r['wall']:.4g} "
              f"atoms/s; stage ms {np.round(r['stage'].mean(axis=0), 3).tolist()}", file=sys.stderr, flush=True)

    h0 = next(iter(handles.values()))
    ent = h0.neighbor_entries()
    flop, counts = algorithmic_flop(h0, s, ent)
    rmse = None
    errors = None
    if "fp64" in res and "mixed" in res:
        rmse = float(np.sqrt(np.mean((res["mixed"]["F"] - res["fp64"]["F"]) ** 2)))
        errors = mixed_errors(pkg, s, res["mixed"], res["fp64"])
        print(f"mixed vs fp64: {errors}", file=sys.stderr, flush=True)
    # the cosine cutoff (cutoff_type 1) the configuration names, beside the
    # reference's tanh^3: same cell, same network, its own scaling
    cosine = {}
    if not args.no_cosine:
        cres = {}
        for mode in args.modes.split(","):
            _, r = run_mode(pkg, torch, s, mode, args.steps, args.warmup, params=PARAMS_COS)
            cres[mode] = r
            cosine[mode] = {"atoms_per_s": s.nat * args.steps / r["wall"], "ms_per_step": r["wall"] / args.steps * 1e3,
                            "stage_ms": np.round(r["stage"].mean(axis=0), 4).tolist()}
            print(f"cosine cutoff, {mode}: {cosine[mode]['atoms_per_s']:.4g} atoms/s", file=sys.stderr, flush=True)
        if "fp64" in cres and "mixed" in cres:
            cosine["mixed_vs_fp64"] = mixed_errors(pkg, s, cres["mixed"], cres["fp64"])
    # the MLP's forward layers in BF16 (FP32 accumulation), measured against
    # FP64 beside the FP32 form; it is not the headline unless within bound
    bf16 = None
    if not args.no_bf16 and "fp64" in res:
        _, rb = run_mode(pkg, torch, s, "mixed-bf16", args.steps, args.warmup)
        eb = mixed_errors(pkg, s, rb, res["fp64"])
        bf16 = {"atoms_per_s": s.nat * args.steps / rb["wall"], "ms_per_step": rb["wall"] / args.steps * 1e3,
                "stage_ms": np.round(rb["stage"].mean(axis=0), 4).tolist(), "vs_fp64": eb,
                "within_force_bound": eb["force_rmse"] <= FORCE_RMSE_BOUND}
        print(f"mixed-bf16: {bf16['atoms_per_s']:.4g} atoms/s; vs fp64 {eb}", file=sys.stderr, flush=True)
    head = "mixed" if ("mixed" in res and rmse is not None and rmse <= FORCE_RMSE_BOUND) else \
        ("fp64" if "fp64" in res else "mixed")
    r = res[head]
    ms_step = r["wall"] / args.steps * 1e3
    value = s.nat * args.steps / r["wall"]
    st = r["stage"].mean(axis=0)             # [wrap+cells, list, sf values, mlp, sf grad, gather+reduce]
    chain_ms = float(st[2] + st[3] + st[4])
    achieved = flop / (chain_ms * 1e-3) / 1e12
    out = {
        "metric": "atoms/s (E+F+stress) at <=1e-5 Ha/bohr force err vs FP64; p50/p99 latency",
        "value": value, "unit": "atoms/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else "f32 (symmetry functions, MLP); f64 (wrap, list, sums)",
        "data": "synthetic (seeded perovskite supercell, seeded random-weight potential)",
        "config": {"workload": f"perovskite {args.cells}^3 supercell, {s.nat} atoms, 5 elements, "
                               f"80 SFs/element (40 G2 + 40 G4), MLP 80-25-25-1, rc 12 bohr, "
                               f"cutoff tanh^3 (RuNNer cutoff_type 2, the reference's), "
                               f"E+F+dE/dh, precision {head}",
                   "atoms": s.nat, "precision": head, "cutoff_type": 2},
        "latency_ms": {"p50": float(np.percentile(r["lat_ms"], 50)),
                       "p99": float(np.percentile(r["lat_ms"], 99))},
        "force_rmse_mixed_vs_fp64": rmse,
        "mixed_vs_fp64": errors,
        "cosine_cutoff": cosine or None,
        "bf16_mlp": bf16,
        # the mixed form's rate over the FP64 form's on this GPU (both the same
        # fused two-pass design; the 20x aim assumed a plain FP64 port)
        "mixed_over_fp64": (s.nat * args.steps / res["mixed"]["wall"]) / (s.nat * args.steps / res["fp64"]["wall"])
        if ("mixed" in res and "fp64" in res) else None,
        "modes": {m: {"atoms_per_s": s.nat * args.steps / v["wall"],
                      "ms_per_step": v["wall"] / args.steps * 1e3,
                      "stage_ms": np.round(v["stage"].mean(axis=0), 4).tolist()} for m, v in res.items()},
        "roofline": {"bound": "valu", "kernel": "k_sf_values + k_mlp + k_sf_grad (one evaluation)",
                     "achieved": achieved, "peak": PEAK_TFLOPS[head], "unit": "TFLOP/s",
                     "frac": achieved / PEAK_TFLOPS[head], "traffic": measured_traffic(args.cells, head),
                     "flop_per_launch": flop, "counts": counts, "kernel_ms": chain_ms},
    }
    if not args.no_cpu_baseline:
        print("timing the CPU baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(s)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
