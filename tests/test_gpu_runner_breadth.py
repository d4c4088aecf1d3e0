"""RuNNer keywords beyond the reference's on the GPU: every cutoff_type
0..8 and the scale_min_short / scale_max_short scaling, against the oracle
(pinned by tests/test_runner_breadth.py), in FP64 and mixed precision. Each
cutoff type gets the scaling its own symmetry functions have on a scaling
set (as a potential trained with it would), so that the mixed-precision
bound measures the arithmetic, not an input far outside the scaling."""
from __future__ import annotations

import numpy as np
import pytest

from test_runner_breadth import potential_copy, rescale

pytestmark = pytest.mark.gpu


def _compare(pkg, oracle, wl, d):
    o = oracle.Oracle(oracle.read_potential(d))
    for s in (wl.perovskite(2, 1, 1, seed=5), wl.cluster(30, seed=6)):
        r = o.evaluate(s)
        E, F, L = pkg.parseRuNNerInput(d, "fp64").energy_and_forces(s)
        fs = max(1.0, np.abs(r["force"]).max())
        assert abs(E - r["energy"]) <= 1e-10 * max(1.0, abs(r["energy"]))
        np.testing.assert_allclose(F, r["force"], rtol=0, atol=1e-9 * fs)
        np.testing.assert_allclose(L, r["dEdLat"], rtol=0, atol=1e-9 * max(1.0, np.abs(r["dEdLat"]).max()))
        Em, Fm, _ = pkg.parseRuNNerInput(d, "mixed").energy_and_forces(s)
        rmse = float(np.sqrt(np.mean((Fm - F) ** 2)))
        assert rmse <= 1e-5 * max(1.0, float(np.sqrt(np.mean(F ** 2))))
        assert abs(Em - E) / s.nat <= 1e-5 * max(1.0, abs(E) / s.nat)


@pytest.mark.parametrize("t", range(9))
def test_cutoff_types_against_oracle(pkg, wl, oracle, tmp_path, t):
    _compare(pkg, oracle, wl, rescale(potential_copy(tmp_path, cutoff_type=t), oracle, wl))


@pytest.mark.parametrize("center", [False, True])
def test_scale_range_against_oracle(pkg, wl, oracle, tmp_path, center):
    _compare(pkg, oracle, wl, potential_copy(tmp_path, center=center, scale_range=(-1.0, 1.0)))
