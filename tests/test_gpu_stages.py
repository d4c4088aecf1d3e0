"""The reference's stage API (hdnnpCalcNeighborLists, hdnnpCalcSymfunctions,
hdnnpCalcNNs; hdnnps.f90:373-562) on the GPU against what the reference's
own routines wrote (tests/golden/stages_*.npz):

- through the Fortran module hdnnps, called in the order of
  globalHdnnpEnergyAndForces (globalHdnnps.f90:47-55) by the package's
  driver: the lists bit for bit (atom numbers and image positions, in the
  reference's order), scaled G and dG/dR <= 1e-11, E 1e-10 relative,
  F 1e-9 Ha/bohr, dEdLat 1e-9 max(1, |dEdLat|), for calcSfsTogether .true.
  and .false. (the per-function path); a second potential's networks on the
  first one's symmetry functions give that potential's own evaluation
  (<= 1e-10 relative / 1e-9);
- through the Python mirror, the same, plus: in FP64 the stages give
  hdnnpEnergyAndForces bit for bit, stale stages are rebuilt from the
  structure, and mixed precision stays within the force bound.
"""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from conftest import GOLD, PKG_DIR, POTENTIALS, ROOT, committee_member, golden, golden_structure
from test_stages_cpu import CASES, rows

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(PKG_DIR, "lib", "hdnnp_md")


def _parse(path):
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLD, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    return mg.parse_stages(path)


def _check(got, d, atol_sf=1e-11, atol_f=1e-9, exact_lists=True):
    E = float(d["energy"])
    assert abs(got["energy"] - E) <= 1e-10 * max(1.0, abs(E)), (got["energy"], E)
    np.testing.assert_allclose(got["force"], d["force"], rtol=0, atol=atol_f)
    np.testing.assert_allclose(got["dEdLat"], d["dEdLat"], rtol=0, atol=atol_f * max(1.0, np.abs(d["dEdLat"]).max()))
    np.testing.assert_array_equal(got["nneis"], d["nneis"])
    np.testing.assert_array_equal(got["ineis"], d["ineis"])
    if exact_lists:
        np.testing.assert_array_equal(got["neis"], d["neis"])
    np.testing.assert_allclose(got["G"], d["G"], rtol=0, atol=atol_sf)
    np.testing.assert_allclose(got["dsf"], d["dsf"], rtol=0, atol=atol_sf)


@pytest.mark.parametrize("flag", ["T", "F"])
@pytest.mark.parametrize("case", CASES)
def test_fortran_stages_match_reference(wl, tmp_path, case, flag):
    d = golden(f"stages_{case}_{flag}")
    s = golden_structure(wl, d)
    sf = str(tmp_path / "s.txt")
    out = str(tmp_path / "o.txt")
    wl.write_structure(sf, s)
    pdir = os.path.join(POTENTIALS[str(d["potential"])], "")
    print(f"case={case} calcSfsTogether={flag} nat={s.nat}", flush=True)
    r = subprocess.run([DRIVER, "stages", pdir, sf, flag, str(int(d["ndsf"])), out, "-"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    _check(_parse(out), d)


def test_fortran_second_potential_on_first_symmetry_functions(pkg, wl, tmp_path):
    """globalHdnnps.f90:53-55: hdnnpCalcNNs of member 2 on member 1's symmetry
    functions. Member 2 has member 1's symmetry functions and scaling and
    other weights, so this is member 2's own evaluation."""
    d = golden("stages_ma_tric2_T")
    s = golden_structure(wl, d)
    m2 = committee_member(POTENTIALS["methylammonium"], str(tmp_path / "m2") + "/", seed=3)
    sf = str(tmp_path / "s.txt")
    out = str(tmp_path / "o.txt")
    wl.write_structure(sf, s)
    r = subprocess.run([DRIVER, "stages", os.path.join(POTENTIALS["methylammonium"], ""), sf, "T",
                        str(int(d["ndsf"])), out, m2],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = _parse(out)
    _check(got, d)
    E2, F2, L2 = pkg.parseRuNNerInput(m2, "fp64").energy_and_forces(s)
    assert abs(got["energy2"] - E2) <= 1e-10 * abs(E2) and got["energy2"] != got["energy"]
    np.testing.assert_allclose(got["force2"], F2, rtol=0, atol=1e-9)
    np.testing.assert_allclose(got["dEdLat2"], L2, rtol=0, atol=1e-9 * max(1.0, np.abs(L2).max()))


def _python_stages(pkg, h, s, together):
    nl = pkg.hdnnpCalcNeighborLists(h, s)
    sfs = pkg.hdnnpCalcSymfunctions(h, s, nl, together)
    E, F, L = pkg.hdnnpCalcNNs(h, s, nl, sfs)
    return {"energy": E, "force": F, "dEdLat": L, "nneis": np.array([x.nNeis for x in nl]),
            "ineis": np.concatenate([x.ineis + 1 for x in nl]), "neis": np.concatenate([x.neis for x in nl]),
            "G": np.concatenate([c.sf for c in sfs]), "dsf": np.concatenate([c.dsf.ravel() for c in sfs])}, nl, sfs


@pytest.mark.parametrize("case", CASES)
def test_python_stages_match_reference_and_evaluation(pkg, wl, case):
    d = golden(f"stages_{case}_F")
    s = golden_structure(wl, d)
    h = pkg.parseRuNNerInput(POTENTIALS[str(d["potential"])], "fp64")
    for together in (True, False):
        got, nl, sfs = _python_stages(pkg, h, s, together)
        nd = int(d["ndsf"])
        od = int((3 * d["nsf"][:nd] * d["nneis"][:nd]).sum())
        got["dsf"] = got["dsf"][:od]
        _check(got, d)
    # the fused stages are hdnnpEnergyAndForces, bit for bit
    got, nl, sfs = _python_stages(pkg, h, s, True)
    E, F, L = h.energy_and_forces(s)
    print(f"{case}: stages E {got['energy']!r} evaluation E {E!r}", flush=True)
    assert got["energy"] == E and np.array_equal(got["force"], F) and np.array_equal(got["dEdLat"], L)
    # another evaluation on the handle replaces its lists: the stages are
    # built again from the structure
    h.energy_and_forces(wl.methylammonium() if str(d["potential"]) == "methylammonium" else
                        wl.perovskite(1, 1, 2, seed=3))
    E2, F2, L2 = pkg.hdnnpCalcNNs(h, s, nl, sfs)
    assert abs(E2 - E) <= 1e-12 * abs(E) and np.abs(F2 - F).max() <= 1e-12


def test_python_stages_mixed_precision(pkg, wl):
    d = golden("stages_perov1_T")
    s = golden_structure(wl, d)
    h = pkg.parseRuNNerInput(POTENTIALS["perovskite"], "mixed")
    got, _, _ = _python_stages(pkg, h, s, True)
    rmse = float(np.sqrt(np.mean((got["force"] - d["force"]) ** 2)))
    print(f"mixed stages: F rmse {rmse:.3e}, |dE|/nat {abs(got['energy'] - float(d['energy'])) / s.nat:.3e}",
          flush=True)
    assert rmse <= 1e-5 and abs(got["energy"] - float(d["energy"])) / s.nat <= 1e-5
    np.testing.assert_array_equal(got["ineis"], d["ineis"])
    np.testing.assert_allclose(got["G"], d["G"], rtol=1e-5, atol=1e-5)
