"""The plain-FP64 power function of the mixed-precision Newton iteration
(phd-2d-gfm_amd/csrc/gfm_fastmath.hpp), built for the host with the same
rounding as the device build, against the C library's pow: relative error
within (2 + |y log x|) x 2.2e-16 x 2 over the ratios and exponents the
Riemann solver uses, special values as documented."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def powlib(tmp_path_factory):
    out = tmp_path_factory.mktemp("fastmath") / "libfastmath.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-Wno-unknown-pragmas",
                    "-I", os.path.join(ROOT, "phd-2d-gfm_amd", "csrc"),
                    os.path.join(ROOT, "tests", "fastmath_host.cpp"), "-o", str(out)], check=True)
    lib = ctypes.CDLL(str(out))
    D = ctypes.POINTER(ctypes.c_double)
    lib.pow_fast_n.argtypes = [ctypes.c_int, D, D, D]

    def f(x, y):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(np.broadcast_to(y, x.shape), dtype=np.float64)
        o = np.empty_like(x)
        lib.pow_fast_n(x.size, x.ctypes.data_as(D), y.ctypes.data_as(D), o.ctypes.data_as(D))
        return o
    return f


def check(f, x, y):
    """compare on the documented domain |y log x| < 700"""
    keep = np.abs(np.broadcast_to(y, x.shape) * np.log(x)) < 700
    x = x[keep]
    y = np.broadcast_to(y, keep.shape)[keep]
    ref = np.power(x, y)
    got = f(x, y)
    rel = np.abs(got - ref) / np.abs(ref)
    bound = 2 * 2.2e-16 * (2.0 + np.abs(y * np.log(x)))
    assert np.all(rel <= bound), (rel.max(), x[np.argmax(rel / bound)], y[np.argmax(rel / bound)])


def test_riemann_exponents(powlib):
    rng = np.random.default_rng(7)
    x = np.concatenate([10.0 ** rng.uniform(-8, 8, 20000), 1.0 + rng.uniform(-1e-3, 1e-3, 20000),
                        1.0 + rng.uniform(-1e-9, 1e-9, 5000), [1.0, 0.5, 2.0, np.nextafter(1.0, 2.0)]])
    for g in (1.05, 1.093, 1.4, 1.648, 1.67, 3.0, 4.4, 6.12):
        for y in ((g - 1) / (2 * g), 1.0 / g, 2 * g / (g - 1), 2.0 / (g - 1), 1.0 / ((g - 1) / (2 * g))):
            check(powlib, x, y)


def test_random_exponents(powlib):
    rng = np.random.default_rng(8)
    x = 10.0 ** rng.uniform(-12, 12, 100000)
    y = rng.uniform(-6, 6, 100000)
    check(powlib, x, y)


def test_special_values(powlib):
    got = powlib(np.array([0.0, -1.0, np.nan, 1.0]), 0.3)
    assert got[0] == 0.0 and np.isnan(got[1]) and np.isnan(got[2]) and got[3] == 1.0
