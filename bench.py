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


def cpu_baseline(s, nrep=2):
    """The reference (oracle/_ref/refdump, built from the reference's sources)
    timed on this host's cores; falls back to the oracle's C restatement."""
    import resource
    threads = min(16, len(os.sched_getaffinity(0)))
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
            "sample": f"the same {s.nat}-atom cell, FP64 E+F+dE/dh, 1 warm-up + median of {nrep} "
                      f"calls of hdnnpEnergyAndForces, OMP_NUM_THREADS={threads}",
            "seconds_per_eval": t}


def run_mode(pkg, torch, s, precision, steps, warmup):
    h = pkg.parseRuNNerInput(PARAMS, precision)
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
        stage.append(h.timings())
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return h, dict(E=E, F=force.cpu().numpy(), dEdLat=L, wall=wall, lat_ms=np.array(lat_ms),
                   stage=np.array(stage))


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
        res[mode] = {"atoms_per_s": nat / t, "structures_per_s": len(ss) / t, "ms_per_batch": t * 1e3,
                     "p99_ms": float(np.percentile(lat, 99)) * 1e3,
                     "stage_ms": np.round(h.timings(), 4).tolist(), "F": force.cpu().numpy()}
        print(f"{mode}: {t * 1e3:.3f} ms per batch -> {nat / t:.4g} atoms/s, {len(ss) / t:.4g} structures/s; "
              f"stage ms {res[mode]['stage_ms']}", file=sys.stderr, flush=True)
    rmse = None
    if "fp64" in res and "mixed" in res:
        rmse = float(np.sqrt(np.mean((res["mixed"]["F"] - res["fp64"]["F"]) ** 2)))
    head = "mixed" if ("mixed" in res and rmse is not None and rmse <= FORCE_RMSE_BOUND) else next(iter(res))
    out = {"metric": "atoms/s (E+F+stress), batched independent structures", "value": res[head]["atoms_per_s"],
           "unit": "atoms/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": res[head]["ms_per_batch"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64" if head == "fp64" else "f32 (symmetry functions, MLP); f64 (wrap, list, sums)",
           "data": "synthetic (seeded structures, seeded random-weight potential)",
           "config": {"workload": f"batch of {len(ss)} structures ({nat} atoms): perovskite n1xn2xn3 (n in 1..4) "
                                  f"and 8-64 atom clusters, atoms resident in device memory", "precision": head},
           "force_rmse_mixed_vs_fp64": rmse,
           "modes": {m: {k: v for k, v in r.items() if k != "F"} for m, r in res.items()}}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cells", type=int, default=16, help="supercell edge (16 -> 49,152 atoms)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--modes", default="fp64,mixed")
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--workload", default="cell", choices=["cell", "batch"],
                    help="cell: one perovskite supercell (headline); batch: many independent structures")
    ap.add_argument("--nstruct", type=int, default=4096, help="structures of the batch workload")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.workload == "batch":
        if args.gpus != 1:
            print("bench.py runs on one GPU only", file=sys.stderr)
            sys.exit(2)
        bench_batch(args)
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
    if "fp64" in res and "mixed" in res:
        rmse = float(np.sqrt(np.mean((res["mixed"]["F"] - res["fp64"]["F"]) ** 2)))
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
                               f"E+F+dE/dh, precision {head}",
                   "atoms": s.nat, "precision": head},
        "latency_ms": {"p50": float(np.percentile(r["lat_ms"], 50)),
                       "p99": float(np.percentile(r["lat_ms"], 99))},
        "force_rmse_mixed_vs_fp64": rmse,
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
