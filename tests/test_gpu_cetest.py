"""The reference's circular-explosion convergence test (phd-2d-gfm_amd/cetest.py, after
sim_circularexplosiontest.cpp) against the thesis's results: density L1 errors against a 4000^2
one-fluid reference at t = 0.25, decoded from CEtest_GFM.eps (three significant digits). The tests
use CFL 0.4 and LS9 without periodic reinitialisation; the thesis records neither.

- OGFM and RGFM reproduce the thesis within 5% at N = 50, 100, 200.
- MGFM reproduces the reference code, which discards the 6-iteration extrapolation of
  GFM_modified.cpp:202; the GPU matches the oracle. The thesis's MGFM errors are reproduced only
  when that extrapolation is applied to the new ghost states (the oracle built with
  OR_MGFM_EXTRAPOLATE_NEW_STATES, and the library with GFM_MGFM_EXTRAPOLATE=1), so the thesis
  runs evidently used a version that applied it.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import cetest as CT
import gfm_hip as G

HERE = os.path.dirname(os.path.abspath(__file__))
VARIANT = os.path.join(os.path.dirname(HERE), "oracle", "liboracle_mgfm_new.so")
CFL = 0.4


@pytest.fixture(scope="module")
def ref():
    print("CASE cetest reference 4000^2", flush=True)
    info, rho, steps, status = CT.reference(4000, CFL)
    assert status == 0
    return info, rho


def l1_of(rho_n, ref):
    n = rho_n.shape[0]
    xc = (np.arange(n) + 0.5) * 2.0 / n
    X, Y = np.meshgrid(xc, xc)
    return float(np.abs(rho_n - CT.bilinear(ref[0], ref[1], X, Y)).sum() * (2.0 / n) ** 2)


@pytest.mark.parametrize("gfm", ["OGFM", "riemannGFM"])
def test_thesis_convergence(ref, gfm):
    print("CASE cetest %s" % gfm, flush=True)
    for k, n in enumerate((50, 100, 200)):
        l1, steps, status = CT.gfm_error(gfm, n, CFL, ref[0], ref[1], 9, 0, 10)
        assert status == 0
        assert abs(l1 / CT.THESIS[gfm][k] - 1.0) <= 0.05, (gfm, n, l1, CT.THESIS[gfm][k])


def oracle_mgfm(oracle, path, n):
    L = oracle.lib(path)
    G = oracle.Grid(n, n, 6, 0.0, 0.0, 2.0 / n, 2.0 / n, 0, 0, 0, 0)
    E = oracle.Eos(1.4, 0.0, 1.4, 0.0)
    g1, g2, phi = np.zeros((n + 12, n + 12, 4)), np.zeros((n + 12, n + 12, 4)), np.zeros((n + 12, n + 12))
    L.or_ic_bubble(G, E, 1.0, 1.0, 0.4, oracle._p(oracle.arr(*CT.INSIDE)), oracle._p(oracle.arr(*CT.OUTSIDE)),
                   oracle._p(oracle.arr(*CT.OUTSIDE)), 1e300, oracle._p(g1), oracle._p(g2), oracle._p(phi))
    run = oracle.Run(3, 0, 9, 0, 10, CFL, 0.25, 100000)
    s = np.zeros(1, dtype=np.int32)
    t = np.zeros(1)
    assert L.or_twofluid_run(G, E, run, oracle._p(g1), oracle._p(g2), oracle._p(phi), oracle._p(s),
                             oracle._p(t)) == 0
    return np.where(phi <= 0.0, g1[..., 0], g2[..., 0])[6:6 + n, 6:6 + n]


def test_mgfm_reference_code_and_thesis(ref, oracle):
    print("CASE cetest MGFM", flush=True)
    n = 50
    gpu, steps, status = CT.gfm_error("MGFM", n, CFL, ref[0], ref[1], 9, 0, 10)
    assert status == 0
    as_coded = l1_of(oracle_mgfm(oracle, None, n), ref)
    assert abs(gpu / as_coded - 1.0) <= 1e-6          # the GPU is the reference code
    assert gpu > 2.0 * CT.THESIS["MGFM"][0]          # which is not what the thesis plots
    applied = l1_of(oracle_mgfm(oracle, VARIANT, n), ref)
    assert abs(applied / CT.THESIS["MGFM"][0] - 1.0) <= 0.02
    # the library's switch for that form (GFM_MGFM_EXTRAPOLATE=1) on the GPU
    os.environ["GFM_MGFM_EXTRAPOLATE"] = "1"
    try:
        for k, n in enumerate((50, 100, 200)):
            sw, _, status = CT.gfm_error("MGFM", n, CFL, ref[0], ref[1], 9, 0, 10)
            assert status == 0
            assert abs(sw / CT.THESIS["MGFM"][k] - 1.0) <= 0.05, (n, sw, CT.THESIS["MGFM"][k])
            if k == 0:
                assert abs(sw / applied - 1.0) <= 1e-6
    finally:
        del os.environ["GFM_MGFM_EXTRAPOLATE"]


def test_vof_gfm_thesis(ref):
    """GFM_VOF on the level set's PLIC interface against the thesis's VOF-GFM errors, which were run
    with the reference's own VOF tracker (PYVOF, absent here): within 15% at N = 50, 100, 200.  The
    level-set interface gives lower errors than the thesis's tracker (0.0412 and 0.0230 at N = 50 and
    100, 5% and 10% below), close to the thesis's MGFM (0.0413, 0.0236)."""
    print("CASE cetest VOFGFM", flush=True)
    for k, n in enumerate((50, 100, 200)):
        l1, steps, status = CT.gfm_error("VOFGFM", n, CFL, ref[0], ref[1], 9, 0, 10)
        print("VOFGFM N=%d L1=%.5f thesis=%.5f" % (n, l1, CT.THESIS["VOFGFM"][k]), flush=True)
        assert status == 0
        assert abs(l1 / CT.THESIS["VOFGFM"][k] - 1.0) <= 0.15, (n, l1, CT.THESIS["VOFGFM"][k])


def test_vof_gfm_on_clsvof_tracker_thesis(ref):
    """GFM_VOF on the CLSVOF tracker (a level set with a conservatively advected volume fraction)
    against the thesis's VOF-GFM errors, which came from a VOF tracker: within 15% at N = 50, 100, 200,
    with first-order convergence"""
    print("CASE cetest VOFGFM on CLSVOF", flush=True)
    got = []
    for k, n in enumerate((50, 100, 200)):
        l1, steps, status = CT.gfm_error("VOFGFM", n, CFL, ref[0], ref[1], 9, 0, 10, itm=G.ITM_CLSVOF)
        print("VOFGFM+CLSVOF N=%d L1=%.5f thesis=%.5f" % (n, l1, CT.THESIS["VOFGFM"][k]), flush=True)
        assert status == 0
        got.append(l1)
        assert abs(l1 / CT.THESIS["VOFGFM"][k] - 1.0) <= 0.15, (n, l1, CT.THESIS["VOFGFM"][k])
    assert got[0] > got[1] > got[2]
