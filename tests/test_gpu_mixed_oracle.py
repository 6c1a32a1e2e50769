"""The mixed-precision whole step against the FP64 oracle on the benchmark workloads.

Mixed precision (FP32 state, reconstruction, Hancock step, flux differences, face sampling,
extrapolation and level set; FP64 Newton iterations, interfacial star states and sums) cannot match the
FP64 oracle to rounding.  What it must do is stay within FP32 accuracy of it on the same inputs while
the flow is not yet amplifying differences.  Per conserved variable k, composite real-fluid state (phi
<= 0: fluid 1), interior cells, after the stated steps:
    L1(k)   = mean |mixed - oracle| / mean |oracle|   (momenta: over the momentum magnitude)
    Linf(k) = max  |mixed - oracle| / max  |oracle|
and the interface: the cells whose side of the zero contour differs, as a fraction of the cells
on the phi <= 0 side.  Bounds, about four times the largest measured values (L1 7.2e-7, Linf 5.0e-6,
no cell on the other side): L1 <= 3e-6, Linf <= 2e-5, side <= 1e-3.

Workloads:
  - the four smallest instances of the 64-instance shocked-bubble batch (its exact seeded parameters),
    one batched context, GFM_riemann, 100 steps;
  - four instances of the 256-instance RMI batch (their seeded amplitude, mode count and Mach number)
    at 512x128, periodic in y, GFM_modified, 40 initial reinitialisation iterations, 100 steps;
  - the 4096x1152 shocked bubble's seeded bubble gas at 512x144, GFM_original, GFM_real and
    GFM_riemann, 100 steps.
Full-size runs of the same workloads, where no oracle is affordable, check properties instead
(test_gpu_fullsize.py).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import accuracy as A
import gfm_hip as G
import oracle_lib as O
import problems as P

SEED = 20261203


def run_mixed(cases, variant, steps, init_reinit=0):
    ctx = G.Context(G.methods(sim_type=G.SIM_TWOFLUID, fs=G.FS_MUSCL, gfm=variant, ls_order=5,
                              reinit_every=5, reinit_iters=10, precision=G.MIXED))
    for info, eos, _ in cases:
        ctx.add(info, eos, 0.6, 1e300, steps)
    ctx.commit()
    for k, (_, _, st) in enumerate(cases):
        ctx.upload(k, *st)
    if init_reinit:
        ctx.ls_reinit(init_reinit)
    ctx.step_graph(steps)
    ctx.sync()
    out = [(ctx.download(k), ctx.get_time(k), ctx.status(k)) for k in range(len(cases))]
    ctx.close()
    return out


def check(cases, variant, steps, bound_l1, bound_linf, bound_side, init_reinit=0):
    got = run_mixed(cases, variant, steps, init_reinit)
    ref = O.run_many(O.twofluid_run, [(info, eos, st, variant, steps, O.FS_MUSCL, 5, 5, 10, 0.6, init_reinit)
                                      for info, eos, st in cases])
    worst = {}
    for (info, _, _), (st, (t, n, _), status), (r1, r2, rp, ost, ot) in zip(cases, got, ref):
        assert ost == 0 and status == 0 and n == steps
        assert abs(t - ot) <= 1e-5 * ot                        # dt from FP32 wave speeds, FP64 sums
        e = A.compare(info, st, (r1, r2, rp))
        side = e["contour_L1_rel"]
        for k in A.NAMES:
            worst["L1_" + k] = max(worst.get("L1_" + k, 0.0), e["L1_" + k])
            worst["Linf_" + k] = max(worst.get("Linf_" + k, 0.0), e["Linf_" + k])
        worst["side"] = max(worst.get("side", 0.0), side)
    print("mixed vs oracle:", {k: "%.2e" % v for k, v in worst.items()}, flush=True)
    for k in A.NAMES:
        assert worst["L1_" + k] <= bound_l1, (k, worst)
        assert worst["Linf_" + k] <= bound_linf, (k, worst)
    assert worst["side"] <= bound_side, worst


def test_mixed_batch_subset_vs_oracle():
    print("CASE mixed batch subset", flush=True)
    cs = sorted(P.bubble_batch(64), key=lambda c: c["Nx"] * c["Ny"])[:4]
    cases = [P.shocked_bubble(**c) for c in cs]
    check(cases, G.GFM_RIEMANN, 100, bound_l1=3e-6, bound_linf=2e-5, bound_side=1e-3)


def test_mixed_rmi_subset_vs_oracle():
    print("CASE mixed rmi subset", flush=True)
    cs = P.rmi_batch(256)[:4]
    cases = [P.rmi(**dict(c, Nx=512, Ny=128)) for c in cs]
    check(cases, G.GFM_MODIFIED, 100, bound_l1=3e-6, bound_linf=2e-5, bound_side=1e-3, init_reinit=40)


@pytest.mark.parametrize("variant", [G.GFM_ORIGINAL, G.GFM_REAL, G.GFM_RIEMANN])
def test_mixed_large_reduced_vs_oracle(variant):
    print("CASE mixed large reduced variant=%d" % variant, flush=True)
    rng = np.random.default_rng(SEED)
    c = dict(bubble_gamma=float(rng.uniform(1.09, 1.67)), bubble_rho=float(rng.uniform(0.18, 5.0)))
    cases = [P.shocked_bubble(512, 144, **c)]
    check(cases, variant, 100, bound_l1=3e-6, bound_linf=2e-5, bound_side=1e-3)


def test_ghost_band_against_every_ghost_state(monkeypatch):
    """The mixed path's ghost band (GFM_GHOST_BAND, default 24 cells: the flux kernel does not update
    a fluid on tiles where it is a ghost farther than that from its real region) against the same steps
    with every ghost state updated (GFM_GHOST_BAND=0), on the batch subset: after 100 steps the
    real-fluid states agree far below the mixed path's own error against the oracle (bound 1e-8 L1
    and 1e-6 Linf per conserved variable, no cell on the other side of the interface), while fluid 1's
    ghost states more than 30 cells from the bubble differ, i.e. the band did skip work"""
    print("CASE ghost band", flush=True)
    cs = sorted(P.bubble_batch(64), key=lambda c: c["Nx"] * c["Ny"])[:4]
    cases = [P.shocked_bubble(**c) for c in cs]
    monkeypatch.delenv("GFM_GHOST_BAND", raising=False)
    band = run_mixed(cases, G.GFM_RIEMANN, 100)
    monkeypatch.setenv("GFM_GHOST_BAND", "0")
    full = run_mixed(cases, G.GFM_RIEMANN, 100)
    skipped = 0
    for (info, _, _), (a, ta, sa), (b, tb, sb) in zip(cases, band, full):
        assert sa == 0 and sb == 0 and ta[1] == tb[1] == 100
        e = A.compare(info, a, b)
        for k in A.NAMES:
            assert e["L1_" + k] <= 1e-8 and e["Linf_" + k] <= 1e-6, (k, e)
        assert e["contour_L1"] == 0.0
        # the band did skip work: some ghost states far from the interface kept older values
        g = info.numGC
        phi = b[2][g:g + info.Ny, g:g + info.Nx]
        far1 = phi > 30 * max(info.dx, info.dy)
        d1 = a[0][g:g + info.Ny, g:g + info.Nx][far1] != b[0][g:g + info.Ny, g:g + info.Nx][far1]
        skipped += int(d1.any())
    assert skipped == len(cases)
