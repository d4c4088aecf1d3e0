"""Batched evaluation (hdnnp_evaluate_batch): many independent structures of
uneven size in one evaluation. Each structure's neighbour list, symmetry
functions, MLP rows, gradient rows and reduction chunks are the same as in an
evaluation of that structure alone, so the batch must reproduce the
single-structure results bit for bit (FP64 and mixed)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import PEROV_PARAMS, golden, golden_structure

pytestmark = pytest.mark.gpu


def _structures(wl):
    empty = wl.Structure(np.zeros((0, 3)), np.zeros(0, np.int32))
    return [wl.perovskite(1, 1, 1, seed=11), wl.cluster(8, seed=1), wl.perovskite(2, 1, 1, seed=12),
            wl.cluster(40, seed=2), empty, wl.perovskite(1, 2, 3, seed=13), wl.cluster(64, seed=3),
            golden_structure(wl, golden("eval_perov2"))]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_equals_single(pkg, wl, precision):
    h = pkg.parseRuNNerInput(PEROV_PARAMS, precision)
    ss = _structures(wl)
    print(f"batch of {len(ss)} ({sum(s.nat for s in ss)} atoms), {precision}", flush=True)
    E, F, L = h.energy_and_forces_batch(ss)
    for k, s in enumerate(ss):
        if s.nat == 0:
            assert E[k] == 0.0 and F[k].shape == (0, 3) and not L[k].any()
            continue
        e, f, l = h.energy_and_forces(s)
        assert E[k] == e, (k, E[k], e)
        assert np.array_equal(F[k], f), k
        if s.periodic:
            assert np.array_equal(L[k], l), k
        else:
            assert not L[k].any()


def test_batch_against_reference_golden(pkg, wl):
    """A structure with golden reference values keeps them inside a batch."""
    h = pkg.parseRuNNerInput(PEROV_PARAMS, "fp64")
    d = golden("eval_perov2")
    s = golden_structure(wl, d)
    E, F, L = h.energy_and_forces_batch([wl.cluster(20, seed=5), s, wl.perovskite(1, 1, 2, seed=6)])
    assert abs(E[1] - float(d["energy"])) <= 1e-10 * max(1.0, abs(float(d["energy"])))
    np.testing.assert_allclose(F[1], d["force"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(L[1], d["dEdLat"], rtol=0, atol=1e-9 * max(1.0, np.abs(d["dEdLat"]).max()))


def test_large_mixed_batch(pkg, wl):
    """300 structures of the batch recipe (perovskite supercells up to 4x4x4 and
    8-64 atom clusters); 30 of them re-evaluated alone."""
    h = pkg.parseRuNNerInput(PEROV_PARAMS, "mixed")
    ss = wl.batch_mix(300, seed=77)
    print(f"batch of {len(ss)} ({sum(s.nat for s in ss)} atoms)", flush=True)
    E, F, L = h.energy_and_forces_batch(ss)
    for k in np.random.default_rng(1).choice(len(ss), 30, replace=False):
        e, f, l = h.energy_and_forces(ss[k])
        assert E[k] == e and np.array_equal(F[k], f), k


def test_device_batch_equals_host_batch(pkg, wl):
    """hdnnp_evaluate_batch_device (atoms in device memory, bounding boxes of the
    nonperiodic structures computed on the device) gives the host batch's results."""
    import torch
    h = pkg.parseRuNNerInput(PEROV_PARAMS, "fp64")
    ss = [s for s in _structures(wl)]
    E, F, L = h.energy_and_forces_batch(ss)
    nat = np.array([s.nat for s in ss], dtype=np.int32)
    dev = torch.device("cuda:0")
    Z = torch.from_numpy(np.concatenate([s.Z for s in ss]).astype(np.int32)).to(dev)
    pos = torch.from_numpy(np.ascontiguousarray(np.concatenate([s.pos.reshape(-1, 3) for s in ss]))).to(dev)
    Ed, Fd, Ld = h.evaluate_batch_torch(nat, Z, pos, [int(s.periodic) for s in ss],
                                        np.stack([np.asarray(s.lat).reshape(3, 3) for s in ss]))
    assert np.array_equal(E, Ed)
    assert np.array_equal(np.concatenate(F), Fd.cpu().numpy())
    assert np.array_equal(L, Ld)


def test_mixed_batch_force_error_bound(pkg, wl):
    """Mixed precision over a batch of perovskite cells and C/N/H clusters:
    force RMSE against FP64 within 1e-5 Ha/bohr for either kind (the bound of
    the headline metric), energies within 1e-5 Ha per atom, and the stress of
    every periodic structure within 1e-5 Ha/bohr * rc * nat / V (the stress
    of a coherent force error at the force bound over the cutoff radius)."""
    ss = wl.batch_mix(400, seed=2024)
    E64, F64, L64 = pkg.parseRuNNerInput(PEROV_PARAMS, "fp64").energy_and_forces_batch(ss)
    E32, F32, L32 = pkg.parseRuNNerInput(PEROV_PARAMS, "mixed").energy_and_forces_batch(ss)
    for periodic in (True, False):
        ks = [k for k, s in enumerate(ss) if bool(s.periodic) == periodic]
        d = np.concatenate([(F32[k] - F64[k]).ravel() for k in ks])
        rmse = float(np.sqrt(np.mean(d ** 2)))
        de = max(abs(E32[k] - E64[k]) / ss[k].nat for k in ks)
        # stress against its bound 1e-5 Ha/bohr * rc * nat / V, as a ratio
        ds = max(np.abs(pkg.dEdLat2Stress(ss[k].lat, L32[k]) - pkg.dEdLat2Stress(ss[k].lat, L64[k])).max()
                 / (1e-5 * 12.0 * ss[k].nat / abs(np.linalg.det(ss[k].lat))) for k in ks) if periodic else 0.0
        print(f"periodic {periodic}: {len(ks)} structures, F rmse {rmse:.3e}, max dE/nat {de:.3e}, "
              f"stress diff / bound {ds:.3f}", flush=True)
        assert rmse <= 1e-5 and de <= 1e-5 and ds <= 1.0


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cell_counting_sort_equals_radix_sort(pkg, wl, precision):
    """A 12,288-atom perovskite cell alone is sorted into grid cells by the
    counting sort (above the one-workgroup sort's 8,192 atoms, under 24,576
    grid cells); next to a sparse periodic box of 35^3 = 42,875 grid cells the
    batch's grid exceeds 24,576 cells and takes the radix sort. Both give the
    atoms of a grid cell in index order, so the cell's results are the same
    bit for bit."""
    h = pkg.parseRuNNerInput(PEROV_PARAMS, precision)
    big = wl.perovskite(16, 16, 4, seed=21)
    sparse = wl.Structure(np.array([[10.0, 10.0, 10.0], [30.0, 10.0, 10.0]]), np.array([82, 53], np.int32),
                          np.eye(3) * 250.0, True)
    E1, F1, L1 = h.energy_and_forces(big)
    E, F, L = h.energy_and_forces_batch([big, sparse])
    print(f"{precision}: E alone {E1!r}, in the batch {E[0]!r}", flush=True)
    assert E[0] == E1 and np.array_equal(F[0], F1) and np.array_equal(L[0], L1)
    E2, F2, L2 = h.energy_and_forces(sparse)
    assert E[1] == E2 and np.array_equal(F[1], F2) and np.array_equal(L[1], L2)
