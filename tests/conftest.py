"""Shared fixtures: the package (loaded from its hyphenated directory), the
oracle (test infrastructure) and the golden vectors."""
from __future__ import annotations

import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "high-dimensional-neural-network-potential_amd")
ORACLE_DIR = os.path.join(ROOT, "oracle")
GOLD = os.path.join(ROOT, "tests", "golden")
MA_PARAMS = os.path.join(GOLD, "methylammonium")
PEROV_PARAMS = os.path.join(PKG_DIR, "data", "perovskite")
PEROV_MD_PARAMS = os.path.join(PKG_DIR, "data", "perovskite_md")
POTENTIALS = {"methylammonium": MA_PARAMS, "perovskite": PEROV_PARAMS}


def load_package():
    if "hdnnp_amd" in sys.modules:
        return sys.modules["hdnnp_amd"]
    spec = importlib.util.spec_from_file_location(
        "hdnnp_amd", os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["hdnnp_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def load_oracle():
    if ORACLE_DIR not in sys.path:
        sys.path.insert(0, ORACLE_DIR)
    import runner  # noqa: E402
    return runner


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: takes more than a few seconds")


@pytest.fixture(scope="session")
def pkg():
    # torch ships its own HIP runtime; it has to come up before the library's
    # (see load_library), so on a GPU machine initialise it first
    if os.path.exists("/dev/kfd"):
        try:
            import torch
            torch.cuda.is_available()
        except ImportError:
            pass
    return load_package()


@pytest.fixture(scope="session")
def wl(pkg):
    return pkg.workloads


@pytest.fixture(scope="session")
def oracle():
    return load_oracle()


@pytest.fixture(scope="session")
def oracles(oracle):
    return {k: oracle.Oracle(oracle.read_potential(v)) for k, v in POTENTIALS.items()}


def golden(name):
    return np.load(os.path.join(GOLD, name + ".npz"))


def golden_structure(wl, d):
    return wl.Structure(np.array(d["pos"]), np.array(d["Z"], dtype=np.int32), np.array(d["lat"]),
                        bool(d["periodic"]))


EVAL_CASES = ["ma", "ma_cell1", "ma_tric2", "ma_cell8", "cluster30", "perov1", "perov2", "perov3",
              "perov132", "perov6"]


def committee_member(src, dst, seed, weight_noise=0.05, scaling_factor=None):
    """A committee member for tests: the potential in `src` with its network
    weights perturbed (w * (1 + weight_noise * N(0,1)), seeded) and, when
    scaling_factor is given, the min/max columns of scaling.data scaled by it.
    Written as RuNNer files to `dst` (a directory path ending in '/')."""
    import glob
    import shutil
    os.makedirs(dst, exist_ok=True)
    for f in glob.glob(os.path.join(src, "*")):
        shutil.copyfile(f, os.path.join(dst, os.path.basename(f)))
    rng = np.random.default_rng(seed)
    for f in sorted(glob.glob(os.path.join(dst, "weights.*.data"))):
        w = np.loadtxt(f)
        w = w * (1.0 + weight_noise * rng.standard_normal(w.shape))
        np.savetxt(f, w, fmt="%20.12f")
    if scaling_factor is not None:
        path = os.path.join(dst, "scaling.data")
        out = []
        for line in open(path).read().splitlines():
            p = line.split()
            if len(p) == 5:          # element, SF index, min, max, mean
                p[2] = repr(float(p[2]) * scaling_factor)
                p[3] = repr(float(p[3]) * scaling_factor)
                line = "  ".join(p)
            out.append(line)
        open(path, "w").write("\n".join(out) + "\n")
    return dst
