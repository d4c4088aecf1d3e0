"""RuNNer keywords beyond the reference's (cutoff_type 0..8, the
scale_min_short / scale_max_short min-max scaling, and symmetry-function
type 9, the wide angular G5), in the oracle and the parser. The reference
evaluates none of them (it always uses cutoff type 2, ignores the scaling
range and skips other function types), so these are pinned by the
mathematics: closed forms at special points, G2 of two atoms, G5 of three
atoms (also where r_jk lies beyond the cutoff, which G4 cuts and G5 does
not), F = -grad E, and the extrema of a scaling set landing on
[scale_min, scale_max]. No GPU."""
from __future__ import annotations

import os
import shutil

import numpy as np
import pytest

from conftest import PEROV_PARAMS

RC = 12.0
TYPES = list(range(9))


def closed_form(t, r, rc=RC):
    """Cutoff functions in factored / trigonometric form (x = r / rc)."""
    x = r / rc
    if t == 0:
        return 1.0
    if t == 1:
        return np.cos(np.pi * x / 2) ** 2              # (cos(pi x) + 1) / 2
    if t == 2:
        return np.tanh(1 - x) ** 3
    if t == 3:
        return (np.tanh(1 - x) / np.tanh(1.0)) ** 3
    if t == 4:
        return np.exp(-x * x / (1 - x * x))            # exp(1 - 1/(1 - x^2))
    # smooth steps f(x) = 1 - S(x), S(x) = x^k p(x) with S(0) = 0, S(1) = 1,
    # the first k-1 derivatives zero at both ends, and S(x) + S(1-x) = 1:
    # f = 1 - S(x) below x = 1/2 and S(1 - x) above (no cancellation)
    S = {5: lambda u: u * u * (3 - 2 * u),
         6: lambda u: u ** 3 * (10 - 15 * u + 6 * u * u),
         7: lambda u: u ** 4 * (35 - 84 * u + 70 * u * u - 20 * u ** 3),
         8: lambda u: u ** 5 * (126 - 420 * u + 540 * u * u - 315 * u ** 3 + 70 * u ** 4)}[t]
    return 1 - S(x) if x < 0.5 else S(1 - x)


def potential_copy(tmp, cutoff_type=None, center=True, scale_range=None):
    d = os.path.join(str(tmp), f"pot_{cutoff_type}_{center}_{scale_range}")
    shutil.copytree(PEROV_PARAMS, d, copy_function=shutil.copyfile)
    out = []
    for ln in open(os.path.join(d, "input.nn")).read().splitlines():
        w = ln.split()
        if w and w[0] == "cutoff_type" and cutoff_type is not None:
            ln = f"cutoff_type {cutoff_type}"
        if w and w[0] == "center_symmetry_functions" and not center:
            continue
        out.append(ln)
    if scale_range is not None:
        out += [f"scale_min_short {scale_range[0]}", f"scale_max_short {scale_range[1]}"]
    open(os.path.join(d, "input.nn"), "w").write("\n".join(out) + "\n")
    return d + "/"


def rescale(d, oracle, wl):
    """scaling.data of potential directory d recomputed, as a training run
    would, for its own symmetry functions: min/max/mean of the unscaled G
    (the oracle's) over a perovskite cell and C/N/H clusters."""
    raw = d.rstrip("/") + "_raw/"
    shutil.copytree(d, raw, copy_function=shutil.copyfile)
    txt = open(os.path.join(raw, "input.nn")).read()
    txt = txt.replace("scale_symmetry_functions", "").replace("center_symmetry_functions", "")
    open(os.path.join(raw, "input.nn"), "w").write(txt)
    o = oracle.Oracle(oracle.read_potential(raw))
    rows = {z: [] for z in wl.ELEMENTS}
    for s in [wl.perovskite(2, 2, 2, seed=11)] + [wl.cluster(8 + 9 * k, seed=70 + k) for k in range(6)]:
        G = o.evaluate(s, want_sf=True)["G"]
        for z in wl.ELEMENTS:
            if np.any(s.Z == z):
                rows[z].append(G[s.Z == z][:, :80])
    with open(os.path.join(d, "scaling.data"), "w") as f:
        for iel, z in enumerate(wl.ELEMENTS, start=1):
            g = np.concatenate(rows[z])
            for k in range(80):
                mn, mx, me = float(g[:, k].min()), float(g[:, k].max()), float(g[:, k].mean())
                if not mx - mn > 0:
                    mx = mn + 1.0
                f.write(f"{iel} {k + 1} {mn!r} {mx!r} {me!r}\n")
        f.write("-1.0 1.0\n")
    return d


@pytest.mark.parametrize("t", TYPES)
def test_cutoff_special_points(oracle, t):
    f0 = oracle.cutoff(1e-300, RC, t)[0]
    assert abs(f0 - (np.tanh(1.0) ** 3 if t == 2 else 1.0)) < 1e-15
    assert oracle.cutoff(RC, RC, t) == (0.0, 0.0) and oracle.cutoff(RC + 1, RC, t) == (0.0, 0.0)
    half = oracle.cutoff(RC / 2, RC, t)[0]
    if t in (1, 5, 6, 7, 8):
        assert abs(half - 0.5) < 1e-14                 # symmetric smooth steps
    if t == 4:
        assert abs(half - np.exp(-1 / 3)) < 1e-15
    if t == 0:
        assert oracle.cutoff(RC * (1 - 1e-9), RC, t) == (1.0, 0.0)
        return
    # continuous and flat at rc, decreasing on (0, rc)
    fe, de = oracle.cutoff(RC * (1 - 1e-6), RC, t)
    assert abs(fe) < 1e-5 and abs(de * RC) < 1e-4
    rs = np.linspace(0.05, RC - 0.05, 200)
    fs = [oracle.cutoff(r, RC, t)[0] for r in rs]
    assert np.all(np.diff(fs) < 0)


@pytest.mark.parametrize("t", TYPES)
def test_cutoff_closed_forms_and_derivative(oracle, t):
    for r in (0.3, 2.9, 6.0, 9.4, 11.7):
        fc, dfc = oracle.cutoff(r, RC, t)
        assert abs(fc - closed_form(t, r)) < 1e-13
        h = 1e-6
        fp = (closed_form(t, r + h) - closed_form(t, r - h)) / (2 * h)
        assert abs(dfc * r - fp) < 1e-8


@pytest.mark.parametrize("t", TYPES)
def test_g2_two_atoms_each_cutoff(oracle, wl, tmp_path, t):
    from test_oracle import _write_potential
    eta, rs, r = 0.031, 1.5, 3.7
    _write_potential(str(tmp_path), [f"symfunction_short H 2 H {eta} {rs} {RC}"])
    with open(os.path.join(str(tmp_path), "input.nn"), "a") as f:
        f.write(f"cutoff_type {t}\n")
    o = oracle.Oracle(oracle.read_potential(str(tmp_path)))
    s = wl.Structure(np.array([[0.0, 0, 0], [r, 0, 0]]), np.array([1, 1], np.int32))
    G = o.evaluate(s, want_sf=True)["G"]
    want = np.exp(-eta * (r - rs) ** 2) * closed_form(t, r)
    np.testing.assert_allclose(G[:, 0], [want, want], rtol=1e-13)


@pytest.mark.parametrize("t", TYPES)
def test_forces_are_minus_gradient_each_cutoff(oracle, wl, tmp_path, t):
    o = oracle.Oracle(oracle.read_potential(potential_copy(tmp_path, cutoff_type=t)))
    s = wl.cluster(12, seed=40 + t)
    r = o.evaluate(s)
    h = 1e-5
    for i in (0, 5, 11):
        for a in range(3):
            sp, sm = s.copy(), s.copy()
            sp.pos[i, a] += h
            sm.pos[i, a] -= h
            fd = -(o.evaluate(sp)["energy"] - o.evaluate(sm)["energy"]) / (2 * h)
            assert abs(fd - r["force"][i, a]) < 1e-6 * max(1.0, abs(fd)), (t, i, a)


@pytest.mark.parametrize("center", [False, True])
def test_scale_range_maps_the_scaling_set(oracle, wl, tmp_path, center):
    """With scale_min_short -1 / scale_max_short 2, every symmetry function
    of the structure its scaling.data was made from spans [-1, 2] exactly
    (min-max form), or has its mean at -1 (centred form)."""
    s = wl.perovskite(2, 2, 1, seed=3)
    raw = potential_copy(tmp_path, center=False)
    text = open(os.path.join(raw, "input.nn")).read().replace("scale_symmetry_functions", "")
    open(os.path.join(raw, "input.nn"), "w").write(text)
    G = oracle.Oracle(oracle.read_potential(raw)).evaluate(s, want_sf=True)["G"]     # unscaled
    d = potential_copy(tmp_path, center=center, scale_range=(-1.0, 2.0))
    zs = sorted(set(int(z) for z in s.Z))
    with open(os.path.join(d, "scaling.data"), "w") as f:
        for iel, z in enumerate(sorted(set(wl.ELEMENTS)), start=1):
            g = G[s.Z == z][:, :80]
            for k in range(80):
                mn, mx = (g[:, k].min(), g[:, k].max()) if g.size else (0.0, 1.0)
                if mx - mn < 1e-300:
                    mx = mn + 1.0
                me = g[:, k].mean() if g.size else 0.0
                f.write(f"{iel} {k + 1} {float(mn)!r} {float(mx)!r} {float(me)!r}\n")
        f.write("-1.0 1.0\n")
    Gs = oracle.Oracle(oracle.read_potential(d)).evaluate(s, want_sf=True)["G"]
    for z in zs:
        g, gs = G[s.Z == z][:, :80], Gs[s.Z == z][:, :80]
        live = g.max(0) - g.min(0) > 1e-9 * np.maximum(1.0, np.abs(g).max(0))
        if center:
            np.testing.assert_allclose(gs.mean(0)[live], -1.0, atol=1e-9)
        else:
            np.testing.assert_allclose(gs.min(0)[live], -1.0, atol=1e-12)
            np.testing.assert_allclose(gs.max(0)[live], 2.0, atol=1e-12)


def test_scale_range_forces_are_minus_gradient(oracle, wl, tmp_path):
    o = oracle.Oracle(oracle.read_potential(potential_copy(tmp_path, center=False, scale_range=(-1.0, 1.0))))
    s = wl.cluster(10, seed=8)
    r = o.evaluate(s)
    h = 1e-5
    for i, a in ((0, 0), (4, 1), (9, 2)):
        sp, sm = s.copy(), s.copy()
        sp.pos[i, a] += h
        sm.pos[i, a] -= h
        fd = -(o.evaluate(sp)["energy"] - o.evaluate(sm)["energy"]) / (2 * h)
        assert abs(fd - r["force"][i, a]) < 1e-6 * max(1.0, abs(fd))


def test_parser_reads_runner_keywords(pkg, oracle, tmp_path):
    """The library's parser and the oracle's agree on the new keywords; bad
    values are model errors (code 3). No GPU: creation does not touch it."""
    d = potential_copy(tmp_path, cutoff_type=6, center=False, scale_range=(-0.5, 0.5))
    h = pkg.parseRuNNerInput(d)
    assert (h.info.cutoff_type, h.info.scale_range, h.info.scale_min, h.info.scale_max) == (6, 1, -0.5, 0.5)
    p = oracle.read_potential(d)
    assert (p.cutoff_type, p.scale_range, p.scale_min, p.scale_max) == (6, True, -0.5, 0.5)
    ref = pkg.parseRuNNerInput(PEROV_PARAMS)
    assert (ref.info.cutoff_type, ref.info.scale_range) == (2, 0)
    e = os.path.join(str(tmp_path), "g6")
    shutil.copytree(PEROV_PARAMS, e, copy_function=shutil.copyfile)
    open(os.path.join(e, "input.nn"), "a").write("symfunction_short H 6 H 0.01 1.0 12.0\n")
    with pytest.raises(pkg.HdnnpError) as ei:          # the reference would skip it silently
        pkg.parseRuNNerInput(e + "/")
    assert ei.value.code == 3 and "type 6" in str(ei.value)
    for bad in ("cutoff_type 9", "cutoff_type 2 0.1", "scale_min_short 1.0\nscale_max_short 0.5"):
        e = os.path.join(str(tmp_path), "bad" + str(abs(hash(bad))))
        shutil.copytree(PEROV_PARAMS, e, copy_function=shutil.copyfile)
        open(os.path.join(e, "input.nn"), "a").write(bad + "\n")
        with pytest.raises(pkg.HdnnpError) as ei:
            pkg.parseRuNNerInput(e + "/")
        assert ei.value.code == 3
        with pytest.raises(oracle.RunnerError):
            oracle.read_potential(e + "/")


def g5_potential(tmp, oracle, wl, scaled=True):
    """The perovskite potential with its 20 G4 functions of lambda = +1 made
    G5 (type 9, same parameters; they sort after the G4 ones), its scaling
    recomputed by the oracle for the new functions (or none: scaled=False)."""
    d = os.path.join(str(tmp), f"pot_g5_{scaled}")
    shutil.copytree(PEROV_PARAMS, d, copy_function=shutil.copyfile)
    out = []
    for ln in open(os.path.join(d, "input.nn")).read().splitlines():
        w = ln.split()
        if len(w) == 9 and w[0] == "symfunction_short" and w[2] == "3" and float(w[6]) == 1.0:
            w[2] = "9"
            ln = " ".join(w)
        if not scaled and w and w[0] in ("scale_symmetry_functions", "center_symmetry_functions"):
            continue
        out.append(ln)
    open(os.path.join(d, "input.nn"), "w").write("\n".join(out) + "\n")
    return rescale(d + "/", oracle, wl) if scaled else d + "/"


def _g5_raw(r, theta, eta, lam, zeta, rc=RC):
    """G5 of a centre with two neighbours at distance r and angle theta:
    2^(1-zeta) (1 + lambda cos theta)^zeta exp(-2 eta r^2) fc(r)^2."""
    b = max(0.0, 1.0 + lam * np.cos(theta))
    return 2.0 ** (1 - zeta) * b ** zeta * np.exp(-2 * eta * r * r) * closed_form(2, r, rc) ** 2


@pytest.mark.parametrize("r, deg", [(2.5, 60.0), (3.1, 90.0), (8.0, 100.0), (6.5, 150.0)])
def test_g5_three_atoms(oracle, wl, tmp_path, r, deg):
    """An H atom with two H neighbours at distance r and angle theta; at
    8 bohr and 100 degrees the neighbours are 12.26 bohr apart, beyond the
    cutoff: the G4 of the same parameters vanish there, G5 does not."""
    d = g5_potential(tmp_path, oracle, wl, scaled=False)
    p = oracle.read_potential(d)
    o = oracle.Oracle(p)
    th = np.radians(deg)
    s = wl.Structure(np.array([[0.0, 0, 0], [r, 0, 0], [r * np.cos(th), r * np.sin(th), 0]]),
                     np.array([1, 1, 1], np.int32))
    G = o.evaluate(s, want_sf=True)["G"][0]
    rjk = 2 * r * np.sin(th / 2)
    n5 = n4 = 0
    for k, sf in enumerate(p.sfs[1]):
        if sf.name == 9 and sf.els == (1, 1):
            assert abs(G[k] - _g5_raw(r, th, sf.eta, sf.lam, sf.zeta)) <= 1e-14 + 1e-12 * abs(G[k]), sf
            n5 += 1
        if sf.name == 3 and sf.els == (1, 1) and rjk >= RC:
            assert G[k] == 0.0
            n4 += 1
    assert n5 > 0 and (n4 > 0 or rjk < RC)


def test_g5_sorts_after_g4_in_both_parsers(pkg, oracle, wl, tmp_path):
    d = g5_potential(tmp_path, oracle, wl, scaled=False)
    h = pkg.parseRuNNerInput(d)
    p = oracle.read_potential(d)
    for e, z in enumerate(h.elements):
        rows = h.sf_table(e)
        types = rows[:, 0].astype(int)
        assert list(types) == [sf.name for sf in p.sfs[z]]
        assert (types == 9).sum() == 20 and types[-20:].tolist() == [9] * 20
        np.testing.assert_array_equal(rows[:, 4], [sf.eta for sf in p.sfs[z]])
        np.testing.assert_array_equal(rows[:, 6], [sf.lam if sf.name != 2 else 0.0 for sf in p.sfs[z]])


def test_g5_forces_are_minus_gradient(oracle, wl, tmp_path):
    o = oracle.Oracle(oracle.read_potential(g5_potential(tmp_path, oracle, wl)))
    for s in (wl.cluster(14, seed=51), wl.perovskite(1, 1, 1, seed=52)):
        r = o.evaluate(s)
        h = 1e-5
        for i, a in ((0, 0), (3, 1), (7, 2), (11, 0)):
            sp, sm = s.copy(), s.copy()
            sp.pos = s.pos.copy()
            sm.pos = s.pos.copy()
            sp.pos[i, a] += h
            sm.pos[i, a] -= h
            fd = -(o.evaluate(sp)["energy"] - o.evaluate(sm)["energy"]) / (2 * h)
            assert abs(fd - r["force"][i, a]) < 1e-6 * max(1.0, abs(fd)), (i, a)
        if s.periodic:   # dEdLat under a strain of the lattice at fixed fractional coordinates
            eps = 1e-6
            M = np.array([[0.3, 0.1, -0.2], [0.1, -0.4, 0.25], [0.05, 0.2, 0.1]])
            lat = np.asarray(s.lat).reshape(3, 3)
            want = float(np.sum(r["dEdLat"] * (lat @ M.T)))
            vals = []
            for sg in (1, -1):
                t = s.copy()
                t.lat = lat @ (np.eye(3) + sg * eps * M.T)
                t.pos = s.pos @ (np.eye(3) + sg * eps * M.T)
                vals.append(o.evaluate(t)["energy"])
            assert abs((vals[0] - vals[1]) / (2 * eps) - want) < 1e-6 * max(1.0, abs(want))
