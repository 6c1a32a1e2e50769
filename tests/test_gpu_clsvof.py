"""GPU parity of the CLSVOF interface tracker (settings ITM "CLSVOF", sim_twofluid.cpp:571-574) with
the oracle's restatement of the method (oracle/gfm_oracle.c, or_vof_advect / or_clsvof_correct;
the reference's tracker is an external class, so the oracle itself is pinned by
test_oracle_clsvof.py, not against the reference), through the C ABI.

Bars: z's initialisation from the level set to 1e-14; FP64 whole steps (GFM_riemann, GFM_VOF on
the tracker, GFM_modified on a y-periodic instance) against or_twofluid_run_z with the bars of
test_gpu_vof.py (L_inf <= 1e-8 x scale, L1 <= 1e-10 x scale per conserved variable, no cell
changes side) and z to 1e-8; the mixed path: status 0, z within [0, 1], the density within
1e-3 (L1, relative) of the FP64 oracle after 30 steps.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import gfm_hip as G
import oracle_lib as O
import problems as P


def ogrid(info):
    return O.Grid(info.Nx, info.Ny, info.numGC, info.x0, info.y0, info.dx, info.dy,
                  info.BC_L, info.BC_R, info.BC_B, info.BC_T)


def oeos(eos):
    return O.Eos(eos.gamma1, eos.pinf1, eos.gamma2, eos.pinf2)


def scaled_err(a, b):
    a = np.asarray(a).reshape(-1, a.shape[-1])
    b = np.asarray(b).reshape(a.shape)
    scale = np.maximum(np.abs(b).max(axis=0), 1e-300)
    return (np.abs(a - b).max(axis=0) / scale).max(), (np.abs(a - b).mean(axis=0) / scale).max()


def ctx_of(info, eos, variant, precision=G.FP64):
    ctx = G.Context(G.methods(sim_type=G.SIM_TWOFLUID, fs=G.FS_MUSCL, gfm=variant, ls_order=5,
                              reinit_every=5, reinit_iters=10, precision=precision, itm=G.ITM_CLSVOF))
    k = ctx.add(info, eos, 0.6, 1e300, 1 << 30)
    ctx.commit()
    return ctx, k


def oracle_run(info, eos, variant, state, nsteps):
    r1, r2, rp = (np.ascontiguousarray(a, dtype=np.float64).copy() for a in state)
    z = np.zeros_like(rp)
    O.lib().or_clsvof_init(ogrid(info), O._p(rp), O._p(z))
    run = O.Run(variant, 0, 5, 5, 10, 0.6, 1e300, nsteps, 0, 0, O.ITM_CLSVOF)
    s = np.zeros(1, dtype=np.int32)
    t = np.zeros(1)
    st = O.lib().or_twofluid_run_z(ogrid(info), oeos(eos), run, O._p(r1), O._p(r2), O._p(rp), O._p(z),
                                    O._p(s), O._p(t))
    return r1, r2, rp, z, st, float(t[0])


def test_clsvof_init_matches_oracle():
    print("CASE clsvof init", flush=True)
    info, eos, (g1, g2, phi) = P.shocked_bubble(128, 40)
    ctx, k = ctx_of(info, eos, G.GFM_RIEMANN)
    ctx.upload(k, g1, g2, phi)
    z = ctx.download_z(k)
    ctx.close()
    ref = np.zeros_like(phi)
    O.lib().or_clsvof_init(ogrid(info), O._p(phi.copy()), O._p(ref))
    assert np.abs(z - ref).max() <= 1e-14
    assert 0.0 < P.interior(info, z).sum() < P.interior(info, z).size


CASES = [("bubble", G.GFM_RIEMANN), ("bubble", G.GFM_VOFGFM), ("rmi", G.GFM_MODIFIED)]


@pytest.mark.parametrize("case,variant", CASES, ids=["bubble-riemann", "bubble-vofgfm", "rmi-mgfm"])
def test_clsvof_steps_match_oracle(case, variant):
    """40 FP64 steps; the bubble's shock reaches the interface within them, the RMI instance is
    periodic in y (the sweeps wrap through the ghost rows)"""
    print("CASE clsvof steps %s variant=%d" % (case, variant), flush=True)
    nsteps = 40
    info, eos, state = P.shocked_bubble(128, 40) if case == "bubble" else P.rmi(128, 32)
    ctx, k = ctx_of(info, eos, variant)
    ctx.upload(k, *state)
    ctx.step_graph(nsteps)
    ctx.sync()
    o1, o2, op = ctx.download(k)
    oz = ctx.download_z(k)
    t, n, _ = ctx.get_time(k)
    st = ctx.status(k)
    ctx.close()
    assert n == nsteps and st == 0
    r1, r2, rp, rz, ost, to = oracle_run(info, eos, variant, state, nsteps)
    assert ost == 0
    assert abs(t - to) <= 1e-12 * to
    for a, b in ((o1, r1), (o2, r2)):
        linf, l1 = scaled_err(P.interior(info, a), P.interior(info, b))
        assert linf <= 1e-8 and l1 <= 1e-10, (linf, l1)
    pi, pr = P.interior(info, op), P.interior(info, rp)
    assert np.array_equal(pi <= 0, pr <= 0)
    assert np.abs(pi - pr).max() <= 1e-8 * info.dx
    zi, zr = P.interior(info, oz), P.interior(info, rz)
    assert np.abs(zi - zr).max() <= 1e-8
    assert zi.min() >= 0.0 and zi.max() <= 1.0


def test_clsvof_mixed_precision_against_oracle():
    print("CASE clsvof mixed", flush=True)
    nsteps = 30
    info, eos, state = P.shocked_bubble(128, 40)
    ctx, k = ctx_of(info, eos, G.GFM_RIEMANN, precision=G.MIXED)
    ctx.upload(k, *state)
    ctx.step_graph(nsteps)
    ctx.sync()
    o1, o2, op = ctx.download(k)
    oz = ctx.download_z(k)
    st = ctx.status(k)
    ctx.close()
    assert st == 0
    r1, r2, rp, rz, ost, _ = oracle_run(info, eos, G.GFM_RIEMANN, state, nsteps)
    zi = P.interior(info, oz)
    assert zi.min() >= 0.0 and zi.max() <= 1.0
    assert np.abs(zi - P.interior(info, rz)).max() <= 1e-3
    real = P.interior(info, rp) <= 0
    rho = np.where(real, P.interior(info, o1)[..., 0], P.interior(info, o2)[..., 0])
    ref = np.where(real, P.interior(info, r1)[..., 0], P.interior(info, r2)[..., 0])
    assert np.abs(rho - ref).sum() / np.abs(ref).sum() <= 1e-3


def test_level_set_context_unchanged_by_the_tracker_switch():
    """the default tracker (ITM_LS) keeps its results: a context made with itm=ITM_LS steps to the
    bits of one made without the argument"""
    print("CASE itm ls default", flush=True)
    info, eos, state = P.shocked_bubble(96, 28)
    out = []
    for m in (G.methods(gfm=G.GFM_RIEMANN), G.methods(gfm=G.GFM_RIEMANN, itm=G.ITM_LS)):
        ctx = G.Context(m)
        k = ctx.add(info, eos, 0.6, 1e300, 1 << 30)
        ctx.commit()
        ctx.upload(k, *state)
        ctx.step_graph(10)
        ctx.sync()
        out.append(ctx.download(k))
        ctx.close()
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_clsvof_per_stage_level_set_matches_blocked(monkeypatch):
    """the one-sweep level-set kernels (GFM_NO_TEMPORAL_BLOCKING=1, as x-periodic contexts run) place
    the tracker's sweeps and coupling around their stages; 25 steps give the blocked path's bits"""
    print("CASE clsvof per-stage vs blocked", flush=True)
    info, eos, state = P.shocked_bubble(128, 40)
    out = []
    for nb in ("0", "1"):
        monkeypatch.setenv("GFM_NO_TEMPORAL_BLOCKING", nb)
        ctx, k = ctx_of(info, eos, G.GFM_RIEMANN)
        ctx.upload(k, *state)
        ctx.step_graph(25)
        ctx.sync()
        assert ctx.status(k) == 0
        out.append(ctx.download(k) + (ctx.download_z(k),))
        ctx.close()
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_clsvof_periodic_x_matches_oracle():
    """a bubble in a domain periodic in x and reflective in y (the per-stage kernels), 30 FP64 steps of
    a uniform stream carrying it across the periodic side, against the oracle"""
    print("CASE clsvof periodic x", flush=True)
    nsteps = 30
    info, eos, (g1, g2, phi) = P.shocked_bubble(96, 32)
    info.BC_L = info.BC_R = G.BC_PERIODIC
    info.BC_B = info.BC_T = G.BC_REFLECTIVE
    # every cell at the bubble's pressure and density pair, moving right at 0.3 of the gas sound speed
    g1 = np.array(g1, copy=True)
    g2 = np.array(g2, copy=True)
    for g, gam in ((g1, eos.gamma1), (g2, eos.gamma2)):
        rho = g[..., 0].mean()
        p = 1e5
        u = 0.3 * np.sqrt(gam * p / rho)
        g[..., 0] = rho
        g[..., 1] = rho * u
        g[..., 2] = 0.0
        g[..., 3] = p / (gam - 1.0) + 0.5 * rho * u * u
    state = (np.ascontiguousarray(g1), np.ascontiguousarray(g2), phi)
    ctx, k = ctx_of(info, eos, G.GFM_RIEMANN)
    ctx.upload(k, *state)
    ctx.step_graph(nsteps)
    ctx.sync()
    o1, o2, op = ctx.download(k)
    oz = ctx.download_z(k)
    st = ctx.status(k)
    ctx.close()
    assert st == 0
    r1, r2, rp, rz, ost, _ = oracle_run(info, eos, G.GFM_RIEMANN, state, nsteps)
    assert ost == 0
    for a, b in ((o1, r1), (o2, r2)):
        linf, l1 = scaled_err(P.interior(info, a), P.interior(info, b))
        assert linf <= 1e-8 and l1 <= 1e-10, (linf, l1)
    zi, zr = P.interior(info, oz), P.interior(info, rz)
    assert np.abs(zi - zr).max() <= 1e-8
    assert np.array_equal(P.interior(info, op) <= 0, P.interior(info, rp) <= 0)


def test_upload_z_restores_the_tracker():
    """gfm_upload_z sets the tracker's z (a checkpoint taken with gfm_download_z restores the steps'
    bits); level-set contexts refuse it"""
    print("CASE clsvof upload_z", flush=True)
    info, eos, state = P.shocked_bubble(96, 28)
    ctx, k = ctx_of(info, eos, G.GFM_RIEMANN)
    ctx.upload(k, *state)
    ctx.step_graph(12)
    ctx.sync()
    ck = (ctx.download(k), ctx.get_time(k)[:2], ctx.download_z(k))
    ctx.step_graph(10)
    ctx.sync()
    ref = ctx.download(k) + (ctx.download_z(k),)
    ctx.upload(k, *ck[0])
    ctx.upload_z(k, ck[2])
    ctx.set_time(k, *ck[1])
    assert np.array_equal(ctx.download_z(k), ck[2])
    ctx.step_graph(10)
    ctx.sync()
    got = ctx.download(k) + (ctx.download_z(k),)
    ctx.close()
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
    ls = G.Context(G.methods(gfm=G.GFM_RIEMANN))
    kk = ls.add(info, eos, 0.6, 1e300, 1 << 30)
    ls.commit()
    with pytest.raises(G.GfmError):
        ls.upload_z(kk, ck[2])
    ls.close()
