"""RuNNer keywords beyond the reference's on the GPU: every cutoff_type
0..8 and the scale_min_short / scale_max_short scaling, against the oracle
(pinned by tests/test_runner_breadth.py), in FP64 and mixed precision. Each
cutoff type gets the scaling its own symmetry functions have on a scaling
set (as a potential trained with it would), so that the mixed-precision
bound measures the arithmetic, not an input far outside the scaling."""
from __future__ import annotations

import os

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


def test_g5_against_oracle(pkg, wl, oracle, tmp_path):
    """Symmetry-function type 9 (G5, no fc(r_jk)) on the GPU against the
    oracle (pinned by tests/test_runner_breadth.py), FP64 and mixed."""
    from test_runner_breadth import g5_potential
    _compare(pkg, oracle, wl, g5_potential(tmp_path, oracle, wl))


def test_g5_batch_equals_single(pkg, wl, oracle, tmp_path):
    from test_runner_breadth import g5_potential
    d = g5_potential(tmp_path, oracle, wl)
    ss = [wl.perovskite(2, 1, 1, seed=5), wl.cluster(30, seed=6), wl.perovskite(1, 2, 2, seed=7)]
    for prec in ("fp64", "mixed"):
        h = pkg.parseRuNNerInput(d, prec)
        E, F, L = h.energy_and_forces_batch(ss)
        for k, s in enumerate(ss):
            e, f, l = h.energy_and_forces(s)
            assert E[k] == e and np.array_equal(F[k], f), (prec, k)


def test_g5_per_function_path(pkg, wl, oracle, tmp_path):
    """hdnnpCalcSymfunctions with calcSfsTogether = .false. takes G5 through
    the per-function kernel: its values equal the fused G5 kernel's, and its
    dG/dR rows the oracle's G under finite differences."""
    from test_runner_breadth import g5_potential
    d = g5_potential(tmp_path, oracle, wl)
    s = wl.cluster(16, seed=61)
    h = pkg.parseRuNNerInput(d, "fp64")
    nl = pkg.hdnnpCalcNeighborLists(h, s)
    fused = pkg.hdnnpCalcSymfunctions(h, s, nl, True)
    single = pkg.hdnnpCalcSymfunctions(h, s, nl, False)
    for a, b in zip(fused, single):
        np.testing.assert_allclose(a.sf, b.sf, rtol=0, atol=1e-11)
    o = oracle.Oracle(oracle.read_potential(d))
    hh = 1e-5
    for i in (0, 3):                             # atoms with 9 and 10 neighbours
        dsf = single[i].dsf                      # [s, slot, 3]
        for slot in (1, 4):
            j = int(nl[i].ineis[slot])
            fd = np.zeros((dsf.shape[0], 3))
            for c in range(3):
                sp, sm = s.copy(), s.copy()
                sp.pos = s.pos.copy()
                sm.pos = s.pos.copy()
                sp.pos[j, c] += hh
                sm.pos[j, c] -= hh
                fd[:, c] = (o.evaluate(sp, want_sf=True)["G"][i, :dsf.shape[0]] -
                            o.evaluate(sm, want_sf=True)["G"][i, :dsf.shape[0]]) / (2 * hh)
            np.testing.assert_allclose(dsf[:, slot, :], fd, rtol=0, atol=1e-6 * max(1.0, np.abs(fd).max()))
    E0, F0, _ = h.energy_and_forces(s)
    E1, F1, _ = pkg.hdnnpCalcNNs(h, s, nl, single)
    assert abs(E1 - E0) <= 1e-10 * abs(E0) and np.abs(F1 - F0).max() <= 1e-9


def test_blocks_with_many_g4_functions(pkg, wl, oracle, tmp_path):
    """The G4 functions of every centre element spread over three element
    pairs only, ~13 per block (the shipped potentials have at most 4): the
    FP32 passes keep 4 functions of a block in registers and take the others
    from LDS per triplet, which this case runs. Against the oracle in FP64 and
    mixed precision."""
    d = potential_copy(tmp_path)
    pairs = [("H", "H"), ("H", "C"), ("N", "I")]
    out, s = [], 0
    for ln in open(os.path.join(d, "input.nn")).read().splitlines():
        w = ln.split()
        if len(w) >= 8 and w[0] == "symfunction_short" and w[2] == "3":
            a, b = pairs[s % len(pairs)]
            s += 1
            ln = " ".join([w[0], w[1], w[2], a, b] + w[5:])
        out.append(ln)
    assert s >= 3 * 12
    open(os.path.join(d, "input.nn"), "w").write("\n".join(out) + "\n")
    _compare(pkg, oracle, wl, rescale(d, oracle, wl))


def test_general_lambda_model(pkg, wl, oracle, tmp_path):
    """A third of the G4 functions with lambda = +-0.5: the FP32 passes then
    take the general path (per-term powers, the compacted triplet list that
    pass A hands pass B in 16-bit entries) instead of the block sweep.
    Against the oracle in FP64 and mixed precision."""
    d = potential_copy(tmp_path)
    out, s = [], 0
    for ln in open(os.path.join(d, "input.nn")).read().splitlines():
        w = ln.split()
        if len(w) >= 8 and w[0] == "symfunction_short" and w[2] == "3":
            if s % 3 == 0:
                w[6] = "+0.5" if float(w[6]) > 0 else "-0.5"
                ln = " ".join(w)
            s += 1
        out.append(ln)
    open(os.path.join(d, "input.nn"), "w").write("\n".join(out) + "\n")
    _compare(pkg, oracle, wl, rescale(d, oracle, wl))
