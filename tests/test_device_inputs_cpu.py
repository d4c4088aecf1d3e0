"""The torch wrappers of the device entry points (hdnnp_evaluate_device,
hdnnp_evaluate_batch_device) read Z as int32[nat] and pos / write force as
float64[nat][3] at the tensors' data pointers. Every other input must raise
ValueError before anything is launched. These checks run before the library
is called, so they need no GPU: a well-formed input reaches the last check,
that the tensors live in GPU memory."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import PEROV_PARAMS

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def h(pkg):
    return pkg.parseRuNNerInput(PEROV_PARAMS, "fp64")


def _good(nat=12):
    return (torch.zeros(nat, dtype=torch.int32), torch.zeros((nat, 3), dtype=torch.float64),
            torch.zeros((nat, 3), dtype=torch.float64))


LAT = np.eye(3) * 11.9


@pytest.mark.parametrize("bad, msg", [
    (lambda Z, p, f: (Z.to(torch.int64), p, f), "Z must be int32"),
    (lambda Z, p, f: (Z, p.to(torch.float32), f), "pos must be float64"),
    (lambda Z, p, f: (Z, p, f.to(torch.float32)), "force must be float64"),
    (lambda Z, p, f: (Z, p[:-1], f), "pos has shape"),
    (lambda Z, p, f: (Z, p, f[:-1]), "force has shape"),
    (lambda Z, p, f: (Z, p.reshape(-1), f), "pos has shape"),
    (lambda Z, p, f: (Z, torch.zeros((3, 12), dtype=torch.float64).t(), f), "pos must be contiguous"),
    (lambda Z, p, f: (Z, p, torch.zeros((3, 12), dtype=torch.float64).t()), "force must be contiguous"),
    (lambda Z, p, f: (Z, p.numpy(), f), "pos must be a torch tensor"),
    (lambda Z, p, f: (Z, p, f), "must be in GPU memory"),
])
def test_evaluate_torch_rejects(h, bad, msg):
    Z, p, f = bad(*_good())
    with pytest.raises(ValueError, match=msg):
        h.evaluate_torch(Z, p, True, LAT, f)


def test_evaluate_torch_rejects_bad_Z(h):
    with pytest.raises(ValueError, match="one-dimensional"):
        h.evaluate_torch(torch.zeros((12, 1), dtype=torch.int32), *_good()[1:2], True, LAT)
    with pytest.raises(ValueError, match="one-dimensional"):
        h.evaluate_torch(np.zeros(12, np.int32), *_good()[1:2], True, LAT)


def test_evaluate_device_needs_an_explicit_atom_count(h):
    with pytest.raises(TypeError):
        h.evaluate_device(0, 0, 0, True, LAT)          # nat is the first argument
    with pytest.raises(ValueError, match="nat < 0"):
        h.evaluate_device(-1, 0, 0, 0, True, LAT)
    with pytest.raises(ValueError, match="9 values"):
        h.evaluate_device(4, 0, 0, 0, True, np.eye(2))


@pytest.mark.parametrize("bad, msg", [
    (lambda nat, Z, p, f, per, lat: (nat, Z.to(torch.int64), p, f, per, lat), "Z must be int32"),
    (lambda nat, Z, p, f, per, lat: (nat, Z, p.to(torch.float32), f, per, lat), "pos must be float64"),
    (lambda nat, Z, p, f, per, lat: (np.array([5, 8]), Z, p, f, per, lat), "Z has shape"),
    (lambda nat, Z, p, f, per, lat: (nat, Z, p, f[:5], per, lat), "force has shape"),
    (lambda nat, Z, p, f, per, lat: (nat, Z, p, f, [1], lat), "periodic has 1 entries"),
    (lambda nat, Z, p, f, per, lat: (nat, Z, p, f, per, lat[:1]), "9 values per structure"),
    (lambda nat, Z, p, f, per, lat: (np.array([4, -1]), Z, p, f, per, lat), "negative atom count"),
    (lambda nat, Z, p, f, per, lat: (nat, Z, p, f, per, lat), "must be in GPU memory"),
])
def test_evaluate_batch_torch_rejects(h, bad, msg):
    Z, p, f = _good(12)
    args = bad(np.array([4, 8], np.int32), Z, p, f, [1, 0], np.stack([LAT, LAT]))
    nat, Z, p, f, per, lat = args
    with pytest.raises(ValueError, match=msg):
        h.evaluate_batch_torch(nat, Z, p, per, lat, f)
