"""Pins the oracle's exact Riemann solver (the reference links it from the
absent third-party exact_rs_stiffenedgas; call sites pure_RS_exact.cpp:29-60,
mixed_RS_exact.cpp:33-47) against published known answers, an independent
bisection root finder, the pressure-equation invariant and symmetries."""
import json
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def golden():
    with open(os.path.join(HERE, "golden", "toro_star_states.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("case", golden()["cases"], ids=lambda c: c["name"])
def test_toro_table_4_3(oracle, case):
    s, st, it = oracle.rs_star(1.4, 0.0, 1.4, 0.0, case["WL"], case["WR"])
    assert st == 0
    for got, want in zip(s, (case["p"], case["u"], case["rhoL"], case["rhoR"])):
        # table gives 6 significant digits (5 after rounding for small values)
        assert abs(got - want) <= 6e-6 * max(1.0, abs(want)) + 5e-6, (got, want)


def f_side(g, pinf, rho, u, p, ps):
    """Toro's pressure function with barred pressures, written independently"""
    A = 2.0 / ((g + 1.0) * rho)
    pb, psb = p + pinf, ps + pinf
    c = math.sqrt(g * pb / rho)
    if ps > p:
        return (ps - p) * math.sqrt(A / (psb + (g - 1.0) / (g + 1.0) * pb))
    return 2.0 * c / (g - 1.0) * ((psb / pb) ** ((g - 1.0) / (2.0 * g)) - 1.0)


def bisect_pstar(gL, pL_inf, gR, pR_inf, WL, WR):
    F = lambda ps: f_side(gL, pL_inf, WL[0], WL[1], WL[2], ps) + f_side(gR, pR_inf, WR[0], WR[1], WR[2], ps) + WR[1] - WL[1]
    lo = -min(pL_inf, pR_inf) + 1e-300
    hi = max(WL[2], WR[2]) + 1.0
    while F(hi) < 0:
        hi *= 2.0
    for _ in range(400):
        mid = 0.5 * (lo + hi)
        if F(mid) > 0:
            hi = mid
        else:
            lo = mid
        if hi - lo <= 1e-15 * abs(hi):
            break
    return 0.5 * (lo + hi)


EOS_PAIRS = [(1.4, 0.0, 1.4, 0.0), (1.4, 0.0, 1.667, 0.0), (4.4, 6e8, 1.4, 0.0), (7.15, 3.309e8, 1.4, 0.0),
             (1.093, 0.0, 1.4, 0.0)]


@pytest.mark.parametrize("eos", EOS_PAIRS)
def test_bisection_and_invariant(oracle, eos):
    gL, piL, gR, piR = eos
    rng = np.random.default_rng(7)
    scale = 1e5 if max(piL, piR) > 0 else 1.0
    for _ in range(60):
        WL = [rng.uniform(0.1, 5.0) * (1000 if piL > 0 else 1), rng.normal(0, 1) * math.sqrt(scale),
              rng.uniform(0.1, 5.0) * scale]
        WR = [rng.uniform(0.1, 5.0), rng.normal(0, 1) * math.sqrt(scale), rng.uniform(0.1, 5.0) * scale]
        s, st, it = oracle.rs_star(gL, piL, gR, piR, WL, WR)
        if st:
            continue
        ps_ref = bisect_pstar(gL, piL, gR, piR, WL, WR)
        assert abs(s[0] - ps_ref) <= 1e-9 * (abs(ps_ref) + min(piL, piR) + 1e-12 * scale), (s[0], ps_ref)
        # invariant f_L + f_R + du = 0 and u* formula
        fL = f_side(gL, piL, WL[0], WL[1], WL[2], s[0])
        fR = f_side(gR, piR, WR[0], WR[1], WR[2], s[0])
        cscale = math.sqrt(gL * (WL[2] + piL) / WL[0]) + math.sqrt(gR * (WR[2] + piR) / WR[0])
        assert abs(fL + fR + WR[1] - WL[1]) <= 1e-9 * cscale
        assert abs(s[1] - (0.5 * (WL[1] + WR[1]) + 0.5 * (fR - fL))) <= 1e-9 * cscale
        # mirror symmetry: swap sides and negate velocities -> u* -> -u*
        m, st2, _ = oracle.rs_star(gR, piR, gL, piL, [WR[0], -WR[1], WR[2]], [WL[0], -WL[1], WL[2]])
        assert abs(m[0] - s[0]) <= 1e-10 * (abs(s[0]) + min(piL, piR) + 1e-12 * scale)
        assert abs(m[1] + s[1]) <= 1e-9 * cscale
        assert abs(m[2] - s[3]) <= 1e-9 * s[3] and abs(m[3] - s[2]) <= 1e-9 * s[2]


def test_star_density_closed_forms(oracle):
    # shock side: Rankine-Hugoniot density ratio; rarefaction side: isentrope
    s, _, _ = oracle.rs_star(1.4, 0.0, 1.4, 0.0, [1.0, 0.0, 1.0], [0.125, 0.0, 0.1])
    ps = s[0]
    assert abs(s[2] - 1.0 * (ps / 1.0) ** (1 / 1.4)) <= 1e-14
    g6 = 0.4 / 2.4
    r = ps / 0.1
    assert abs(s[3] - 0.125 * (r + g6) / (g6 * r + 1)) <= 1e-14


def test_sample_sod_at_origin(oracle):
    # Sod at x/t = 0 lies in the left star region (u* > 0)
    WL, WR = [1.0, 0.0, 1.0], [0.125, 0.0, 0.1]
    s, _, _ = oracle.rs_star(1.4, 0.0, 1.4, 0.0, WL, WR)
    w = oracle.rs_sample(1.4, 0.0, 1.4, 0.0, WL, WR, s, 0.0)
    assert np.allclose(w, [s[2], s[1], s[0]], rtol=0, atol=0)
    # inside the left fan at x/t = -0.5 the closed-form fan state
    w = oracle.rs_sample(1.4, 0.0, 1.4, 0.0, WL, WR, s, -0.5)
    c = 2 / 2.4 * (math.sqrt(1.4) + 0.2 * (0.0 + 0.5))
    assert abs(w[1] - 2 / 2.4 * (math.sqrt(1.4) + 0.2 * 0.0 - 0.5)) <= 1e-14
    assert abs(w[0] - (c / math.sqrt(1.4)) ** 5) <= 1e-13


def test_pure_flux_identical_states_is_physical_flux(oracle):
    W = [1.3, 0.7, -0.2, 2.5]
    U = oracle.prims_to_cons(1.4, 0.0, W)
    F, st = oracle.pure_rs_flux(True, 1.4, 0.0, U, U)
    Fp = np.zeros(4)
    oracle.lib().or_flux_prims(1.4, 0.0, oracle._p(oracle.arr(*W)), oracle._p(Fp))
    assert st == 0 and np.array_equal(F, Fp)
    G, _ = oracle.pure_rs_flux(False, 1.4, 0.0, U, U)
    Wy = [W[0], W[2], W[1], W[3]]
    Fy = np.zeros(4)
    oracle.lib().or_flux_prims(1.4, 0.0, oracle._p(oracle.arr(*Wy)), oracle._p(Fy))
    assert np.array_equal(G, [Fy[0], Fy[2], Fy[1], Fy[3]])


def test_mixed_rs_two_material_tube(oracle):
    # ST1_x (sim_twofluid.cpp:367-370, :675-706): gamma 1.4 | 1.2, p 1e5 | 1e4
    UL = oracle.prims_to_cons(1.4, 0.0, [1.0, 0.0, 0.0, 1e5])
    UR = oracle.prims_to_cons(1.2, 0.0, [0.125, 0.0, 0.0, 1e4])
    out, st = oracle.mixed_rs(1.4, 0.0, 1.2, 0.0, UL, UR)
    ps = bisect_pstar(1.4, 0.0, 1.2, 0.0, [1.0, 0.0, 1e5], [0.125, 0.0, 1e4])
    assert st == 0 and abs(out[0] - ps) <= 1e-9 * ps


def test_eos_round_trip(oracle):
    rng = np.random.default_rng(11)
    for g, pinf in ((1.4, 0.0), (4.4, 6e8), (7.15, 3309.0)):
        for _ in range(50):
            W = [rng.uniform(0.1, 1000), rng.normal(), rng.normal(), rng.uniform(0.1, 10) * (1 + pinf)]
            U = oracle.prims_to_cons(g, pinf, W)
            W2 = oracle.cons_to_prims(g, pinf, U)
            assert np.allclose(W2, W, rtol=1e-9, atol=1e-9 * (1 + pinf))
