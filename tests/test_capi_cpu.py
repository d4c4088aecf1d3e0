"""C-ABI checks that need no GPU: exported symbols, the RuNNer parser behind
hdnnp_create (against the reference's parsed tables and the oracle parser on
edge-case files), error paths, dEdLat2Stress, workload generators."""
from __future__ import annotations

import ctypes
import os
import re
import shutil

import numpy as np
import pytest

from conftest import MA_PARAMS, PEROV_PARAMS, POTENTIALS, ROOT, golden


def test_library_exports_every_declared_symbol(pkg):
    with open(os.path.join(ROOT, "include", "hdnnp_capi.h")) as f:
        hdr = f.read()
    declared = set(re.findall(r"\b(hdnnp_[a-z_0-9]+)\s*\(", hdr))
    assert declared, "no declarations found"
    lib = pkg.load_library()
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert declared <= set(pkg.EXPORTS) | declared


@pytest.mark.parametrize("name", ["methylammonium", "perovskite"])
def test_capi_parser_matches_reference_table(pkg, name):
    h = pkg.parseRuNNerInput(POTENTIALS[name])
    t = golden("table_" + name)
    assert h.elements == list(t["Z"])
    assert [h.info.nsf[e] for e in range(h.info.nelements)] == list(t["nsf"])
    assert h.maxrc == float(t["maxrc"])
    rows = np.concatenate([h.sf_table(e) for e in range(h.info.nelements)])
    np.testing.assert_array_equal(rows, t["sf"])
    w = np.concatenate([h.weights(e) for e in range(h.info.nelements)])
    np.testing.assert_array_equal(w, t["weights"])
    assert h.info.act[:h.info.nlayers].decode() == str(t["acts"][0])
    assert (h.info.emin, h.info.emax) == tuple(t["energy_range"])
    h.close()


def _edge_case_potential(d):
    """Keyword order shuffled, elements not in Z order, 'd' exponents, a tab
    inside an ignored keyword line, SF keys equal at 5 decimals (stable order
    decides), comments, atom_energy."""
    os.makedirs(d, exist_ok=True)
    lines = [
        "# comment line",
        "global_hidden_layers_short 1",
        "number_of_elements 2",
        "elements O H",
        "global_nodes_short 3",
        "global_activation_short t l",
        "atom_energy O -75.1d0",
        "atom_energy  H  -0.5",
        "energy_threshold\t100.0d0",
        "symfunction_short O 3 H H 0.010000 -1.0 1.0 8.0",
        "symfunction_short O 2 H 0.0100001 0.0 8.0",
        "symfunction_short O 2 H 0.0100004 0.0 8.0",
        "symfunction_short O 2 O 0.5d-1 0.0 8.0",
        "symfunction_short O 3 H O 0.010000 1.0 2.0 8.0",
        "symfunction_short O 3 O H 0.010000 1.0 2.0 8.0",
        "symfunction_short H 2 O 0.3 1.0 8.0",
        "symfunction_short H 2 H 0.3 0.5 8.0",
        "#symfunction_short H 2 H 9.9 0.5 8.0",
        "symfunction_short H 3 O O 0.02 1.0 16.0 8.0",
        "scale_symmetry_functions",
    ]
    with open(os.path.join(d, "input.nn"), "w") as f:
        f.write("\n".join(lines) + "\n")
    with open(os.path.join(d, "scaling.data"), "w") as f:
        for k in range(1, 4):
            f.write(f"1 {k} 0.0 {k}.5 0.1\n")
        for k in range(1, 7):
            f.write(f"2 {k}\n 0.0 {k}.25 0.2\n")   # a record may span lines
        f.write("-1.0 1.0\n")
    rng = np.random.default_rng(0)
    for z, n in ((1, 3), (8, 6)):
        with open(os.path.join(d, f"weights.{z:03d}.data"), "w") as f:
            for v in rng.normal(size=n * 3 + 3 + 3 + 1):
                f.write(f"{float(v)!r} ignored trailing value\n")


def test_capi_parser_matches_oracle_on_edge_cases(pkg, oracle, tmp_path):
    d = str(tmp_path / "pot")
    _edge_case_potential(d)
    pot = oracle.read_potential(d)
    h = pkg.parseRuNNerInput(d)
    assert h.elements == pot.elements == [1, 8]
    for e, z in enumerate(pot.elements):
        rows = h.sf_table(e)
        want = [[s.name, s.els[0], s.els[1] if s.name == 3 else 0, s.rcut, s.eta,
                 s.rs if s.name == 2 else 0.0, s.lam if s.name == 3 else 0.0,
                 s.zeta if s.name == 3 else 0.0, s.sfmin, s.sfmax, s.sfmean] for s in pot.sfs[z]]
        np.testing.assert_array_equal(rows, np.array(want))
        w = np.concatenate([np.concatenate([W.T.reshape(-1), b]) for W, b in pot.layers[z]])
        np.testing.assert_array_equal(h.weights(e), w)
        assert h.info.atom_energy[e] == pot.atom_energy[z]
    # the two G2 of O on H with eta equal at 5 decimals keep file order
    o_rows = h.sf_table(1)
    assert list(o_rows[:3, 4]) == [0.0100001, 0.0100004, 0.05]
    assert h.info.scalesf == 1 and h.info.centersf == 0


@pytest.mark.parametrize("breakage,code", [
    ("no_input", 2), ("no_scaling", 2), ("bad_index", 2), ("too_many_sfs", 2),
    ("unknown_centre", 2), ("no_weights", 2), ("short_weights", 2),
])
def test_capi_parser_errors(pkg, tmp_path, breakage, code):
    d = str(tmp_path / "p")
    shutil.copytree(MA_PARAMS, d)
    if breakage == "no_input":
        os.remove(os.path.join(d, "input.nn"))
    elif breakage == "no_scaling":
        os.remove(os.path.join(d, "scaling.data"))
    elif breakage == "bad_index":
        with open(os.path.join(d, "scaling.data")) as f:
            lines = f.readlines()
        lines[3] = lines[3].replace("   4", "   5", 1)
        with open(os.path.join(d, "scaling.data"), "w") as f:
            f.writelines(lines)
    elif breakage == "too_many_sfs":
        with open(os.path.join(d, "input.nn"), "a") as f:
            for k in range(200):
                f.write(f"symfunction_short H 2 H {0.001 * (k + 1):.4f} 0.0 12.0\n")
    elif breakage == "unknown_centre":
        with open(os.path.join(d, "input.nn"), "a") as f:
            f.write("symfunction_short O 2 H 0.1 0.0 12.0\n")
    elif breakage == "no_weights":
        os.remove(os.path.join(d, "weights.006.data"))
    elif breakage == "short_weights":
        with open(os.path.join(d, "weights.006.data")) as f:
            lines = f.readlines()
        with open(os.path.join(d, "weights.006.data"), "w") as f:
            f.writelines(lines[:-5])
    with pytest.raises(pkg.HdnnpError) as ei:
        pkg.parseRuNNerInput(d)
    assert ei.value.code == code
    assert str(ei.value)


def test_dEdLat_to_stress_closed_form(pkg):
    rng = np.random.default_rng(2)
    lat = np.eye(3) * 11.9 + rng.normal(0, 0.5, (3, 3))     # row k = vector k
    dEdLat = rng.normal(size=(3, 3))                        # [k][a] = dE/dlat(a,k)
    st = pkg.dEdLat2Stress(lat, dEdLat)
    # Fortran: stress = -1/det(lat) * matmul(dEdLat, transpose(lat)) with
    # lat(a,k) = lat_c[k][a]; in C storage the result comes back transposed
    Lf, Df = lat.T, dEdLat.T
    want = -1.0 / np.linalg.det(Lf) * Df @ Lf.T
    np.testing.assert_allclose(st.T, want, rtol=1e-13, atol=1e-14)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_evaluate_without_device_fails_loudly(pkg, wl):
    h = pkg.parseRuNNerInput(MA_PARAMS)
    with pytest.raises(pkg.HdnnpError) as ei:
        h.energy_and_forces(wl.methylammonium())
    assert ei.value.code == 5


def test_workloads_are_deterministic(wl):
    a = wl.perovskite(2, 1, 3, seed=9)
    b = wl.perovskite(2, 1, 3, seed=9)
    assert a.nat == 72 and np.array_equal(a.pos, b.pos)
    assert list(a.Z[:12]) == [82, 53, 53, 53, 7, 6, 1, 1, 1, 1, 1, 1]
    np.testing.assert_allclose(np.diag(a.lat), [23.8, 11.9, 35.7])
    c = wl.cluster(40, seed=3)
    d2 = np.sum((c.pos[:, None] - c.pos[None]) ** 2, axis=2) + np.eye(40) * 100
    assert d2.min() >= 1.8 ** 2 and c.pos.min() >= 0 and c.pos.max() <= 20


def test_fortran_driver_is_built():
    exe = os.path.join(ROOT, "high-dimensional-neural-network-potential_amd", "lib", "hdnnp_md")
    assert os.access(exe, os.X_OK)
    mods = os.path.join(ROOT, "high-dimensional-neural-network-potential_amd", "lib", "hdnnps.mod")
    assert os.path.exists(mods)


def test_committee_rejects_mismatched_member(pkg):
    """Committee members need a network for every element of the first
    potential, taking its symmetry functions (checked at creation, no GPU)."""
    from conftest import MA_PARAMS, PEROV_PARAMS
    with pytest.raises(pkg.HdnnpError) as ei:
        pkg.Hdnnp([PEROV_PARAMS, MA_PARAMS], "fp64")
    assert ei.value.code == 3 and "committee member 2" in str(ei.value)
