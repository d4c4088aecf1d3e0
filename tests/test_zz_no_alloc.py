"""hdnnp_evaluate_batch_device_noalloc, the entry point that allocates nothing:
its results equal hdnnp_evaluate_batch_device's bit for bit (the same
kernels on the same inputs; only the work buffers' addresses differ), on two
calls in a row, the second with a smaller input, in FP64 and mixed precision.
The call takes no device memory (the free memory the device reports does
not fall), the context's own evaluations are unchanged by it, and an input
that needs more scratch than the array holds is refused with
HDNNP_ERR_CAPACITY."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import PEROV_PARAMS

pytestmark = pytest.mark.gpu

HDNNP_ERR_CAPACITY = 6


def _inputs(torch, structs):
    dev = torch.device("cuda:0")
    nat = np.array([s.nat for s in structs], np.int32)
    Z = torch.from_numpy(np.concatenate([s.Z for s in structs]).astype(np.int32)).to(dev)
    pos = torch.from_numpy(np.ascontiguousarray(np.concatenate([s.pos.reshape(-1, 3) for s in structs]))).to(dev)
    per = np.array([int(s.periodic) for s in structs], np.int32)
    lat = np.stack([np.asarray(s.lat, np.float64).reshape(3, 3) for s in structs])
    return nat, Z, pos, per, lat


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_equals_batch_device(pkg, precision):
    import torch
    wl = pkg.workloads
    first = wl.batch_mix(40, seed=31) + [wl.perovskite(6, 6, 6, seed=32)]
    second = wl.batch_mix(7, seed=33)
    assert sum(s.nat for s in second) < sum(s.nat for s in first)
    h = pkg.parseRuNNerInput(PEROV_PARAMS, precision)
    for structs in (first, second):
        nat, Z, pos, per, lat = _inputs(torch, structs)
        print(f"{precision}: {len(structs)} structures, {int(nat.sum())} atoms", flush=True)
        # the existing entry point first (it also has the runtime load every
        # kernel's code object, which is device memory the runtime takes)
        E0, F0, L0 = h.evaluate_batch_torch(nat, Z, pos, per, lat)
        F0 = F0.cpu().numpy()
        force = torch.empty((int(nat.sum()), 3), dtype=torch.float64, device=pos.device)
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        E1, F1, L1 = h.evaluate_batch_torch(nat, Z, pos, per, lat, force, noalloc=True)
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        F1 = F1.cpu().numpy()
        print(f"free device memory before/after the call: {free0} / {free1}", flush=True)
        assert free1 >= free0 - (2 << 20)
        assert np.array_equal(E1, E0)
        assert np.array_equal(F1, F0)
        assert np.array_equal(L1, L0)
        # the context's own buffers came through the call untouched
        E2, F2, L2 = h.evaluate_batch_torch(nat, Z, pos, per, lat)
        assert np.array_equal(E2, E0) and np.array_equal(F2.cpu().numpy(), F0)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_counting_sort_path(pkg, precision):
    """A perovskite cell of 8,448 atoms (the largest both precisions take
    here): above the one-workgroup sort's 8,192 atoms, so the no-allocation
    call sorts by cell with the counting sort, from its fixed array; the
    results equal the allocating entry point's bit for bit."""
    import torch
    wl = pkg.workloads
    h = pkg.parseRuNNerInput(PEROV_PARAMS, precision)
    nat, Z, pos, per, lat = _inputs(torch, [wl.perovskite(8, 8, 11, seed=5)])
    print(f"{precision}: {int(nat.sum())} atoms", flush=True)
    E0, F0, L0 = h.evaluate_batch_torch(nat, Z, pos, per, lat)
    F0 = F0.cpu().numpy()
    force = torch.empty((int(nat.sum()), 3), dtype=torch.float64, device=pos.device)
    E1, F1, L1 = h.evaluate_batch_torch(nat, Z, pos, per, lat, force, noalloc=True)
    assert np.array_equal(E1, E0) and np.array_equal(F1.cpu().numpy(), F0) and np.array_equal(L1, L0)


def test_noalloc_refuses_what_does_not_fit(pkg):
    import torch
    wl = pkg.workloads
    h = pkg.parseRuNNerInput(PEROV_PARAMS, "fp64")
    s = wl.perovskite(16, 16, 8, seed=3)
    nat, Z, pos, per, lat = _inputs(torch, [s])
    force = torch.empty((s.nat, 3), dtype=torch.float64, device=pos.device)
    with pytest.raises(pkg.HdnnpError) as ei:
        h.evaluate_batch_torch(nat, Z, pos, per, lat, force, noalloc=True)
    print(ei.value, flush=True)
    assert ei.value.code == HDNNP_ERR_CAPACITY and "scratch" in str(ei.value)
    # still usable afterwards, with a structure that fits
    small = [wl.perovskite(2, 2, 2, seed=4)]
    nat, Z, pos, per, lat = _inputs(torch, small)
    E1, F1, _ = h.evaluate_batch_torch(nat, Z, pos, per, lat, noalloc=True)
    E0, F0, _ = h.evaluate_batch_torch(nat, Z, pos, per, lat)
    assert np.array_equal(E1, E0) and torch.equal(F1, F0)
