"""Accuracy of the FP64 exp and tanh polynomials of the symmetry-function
kernels (csrc/fmath.hpp), compiled for the host from the same header.

The header is built with g++ and checked against 40-digit values (mpmath):
exp on the range the kernels use (exp(-eta r^2), arguments from -80 to 0)
and beyond, including gradual underflow, overflow and NaN; tanh01 on [0, 1],
the argument 1 - r/rc of the tanh^3 cutoff (symmetryFunctions.f90:310-327).
The bounds are 1.5 ulp for exp and 2.5 ulp for tanh01 (the device library's
routines are below 1 ulp; the tanh^3 cutoff then carries at most ~1.5e-15
relative error, far inside the FP64 parity tolerance).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

mp = pytest.importorskip("mpmath")

from conftest import PKG_DIR

DRIVER = r"""
#include <cstdio>
#include <cstdlib>
#define __host__
#define __device__
#define __forceinline__ inline
#include "fmath.hpp"
int main(int argc, char **argv)
{
    const int mode = std::atoi(argv[1]);
    double x;
    while (std::scanf("%lf", &x) == 1)
        std::printf("%.17g\n", mode == 1 ? hdnnp::exp_f64(x) : mode == 2 ? hdnnp::tanh_f64(x) : hdnnp::tanh01(x));
    return 0;
}
"""


@pytest.fixture(scope="module")
def fm(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("fmath")
    (d / "drv.cpp").write_text(DRIVER)
    exe = d / "drv"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-I", os.path.join(PKG_DIR, "csrc"), str(d / "drv.cpp"), "-o",
                    str(exe)], check=True)

    def run(mode, xs):
        out = subprocess.run([str(exe), str(mode)], input="\n".join("%.17g" % v for v in xs), capture_output=True,
                             text=True, check=True).stdout.split()
        return np.array([float(v) for v in out])
    return run


def _ulps(got, xs, f):
    mp.mp.dps = 40
    err = []
    for g, x in zip(got, xs):
        t = f(mp.mpf(float(x)))
        err.append(float(abs(mp.mpf(float(g)) - t)) / np.spacing(float(t)))
    return np.array(err)


def test_exp_accuracy(fm):
    rng = np.random.default_rng(7)
    xs = np.concatenate([np.linspace(-80, 0, 2001), -rng.random(1500) * 700, rng.random(300) * 700,
                         [0.0, -1e-300, 1e-300, -0.34657359027997264, 0.34657359027997264]])
    e = _ulps(fm(1, xs), xs, mp.exp)
    print(f"exp_f64: max {e.max():.3f} ulp, mean {e.mean():.3f} ulp over {len(xs)} points")
    assert e.max() <= 1.5


def test_exp_special_values(fm):
    xs = [0.0, -708.5, -740.0, -745.0, -746.0, -1e5, -1e300, 709.7, 710.0, 1e5, 1e300]
    got = fm(1, xs)
    with np.errstate(over="ignore"):
        want = np.exp(xs)
    np.testing.assert_array_equal(got, want)     # denormals, 0 and inf exactly as libm
    assert np.isnan(fm(1, [float("nan")])[0])


def test_tanh01_accuracy(fm):
    rng = np.random.default_rng(8)
    xs = np.concatenate([np.linspace(0, 1, 2001)[1:], rng.random(1500), rng.random(300) * 1e-4, [1e-300, 1e-8]])
    e = _ulps(fm(0, xs), xs, mp.tanh)
    print(f"tanh01: max {e.max():.3f} ulp, mean {e.mean():.3f} ulp over {len(xs)} points")
    assert e.max() <= 2.5
    assert fm(0, [0.0])[0] == 0.0


def test_tanh_cubed_cutoff_against_libm(fm):
    """The cutoff (tanh(1 - r/rc))^3 itself, against numpy's tanh."""
    r = np.linspace(0.5, 11.999, 4000)
    x = 1 - r / 12.0
    t = fm(0, x)
    np.testing.assert_allclose(t ** 3, np.tanh(x) ** 3, rtol=4e-15, atol=0)   # ~3 x (1.4 + 1) ulp


def test_tanh_f64_accuracy(fm):
    """The MLP's tanh activation (neuralNetworks.f90:33-55) over any argument."""
    rng = np.random.default_rng(9)
    xs = np.concatenate([np.linspace(-30, 30, 3001), rng.standard_normal(1500) * 3, rng.random(300) * 1e-4 - 5e-5,
                         [1e-300, -1e-300, 18.0, -18.0, 19.5, 400.0, 401.0]])
    xs = xs[xs != 0]
    e = _ulps(fm(2, xs), xs, mp.tanh)
    print(f"tanh_f64: max {e.max():.3f} ulp, mean {e.mean():.3f} ulp over {len(xs)} points")
    assert e.max() <= 2.5
    got = fm(2, [0.0, -0.0, float("inf"), -float("inf"), 1e300, -1e300])
    np.testing.assert_array_equal(got, [0.0, -0.0, 1.0, -1.0, 1.0, -1.0])
    assert np.signbit(got[1])
    assert np.isnan(fm(2, [float("nan")])[0])
