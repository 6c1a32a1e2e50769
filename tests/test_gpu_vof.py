"""GPU parity of the VOF-based ghost fluid methods (GFM_VOF.cpp:8-325, OGFM_VOF.cpp:8-274) with the FP64
oracle, through the C ABI (gfm_set_ghost_states, gfm_step_graph).

Bars as in test_gpu_parity.py (FP64 GPU against FP64 oracle; the GPU differs by FMA contraction in
the probe interpolation and the last-ulp behaviour of pow): one ghost fill <= 1e-10 x scale per
component over the whole padded grids and the velocity field, update boxes and mixed-cell counts exact;
60 steps: L_inf <= 1e-8 x scale, L1 <= 1e-10 x scale per conserved variable, no cell changes side.
The mixed path (FP32 state, FP64 interfacial states) against the oracle on the same inputs: the
bounds of test_gpu_mixed_oracle.py.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import accuracy as A
import gfm_hip as G
import oracle_lib as O
import problems as P

VOF = [G.GFM_VOFGFM, G.GFM_OGFMVOF]


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


def bubble_state(steps=30, Nx=160, Ny=48):
    """a shocked bubble advanced `steps` oracle steps (GFM_riemann), so that the shock has reached
    the interface"""
    info, eos, (g1, g2, phi) = P.shocked_bubble(Nx, Ny)
    run = O.Run(O.GFM_RIEMANN, 0, 5, 5, 10, 0.6, 1e300, steps)
    s = np.zeros(1, dtype=np.int32)
    t = np.zeros(1)
    assert O.lib().or_twofluid_run(ogrid(info), oeos(eos), run, O._p(g1), O._p(g2), O._p(phi), O._p(s),
                                   O._p(t)) == 0
    return info, eos, g1, g2, phi


def ctx_of(info, eos, variant, precision=G.FP64, slip=False, max_steps=1 << 30):
    ctx = G.Context(G.methods(sim_type=G.SIM_TWOFLUID, fs=G.FS_MUSCL, gfm=variant, ls_order=5,
                              reinit_every=5, reinit_iters=10, precision=precision))
    k = ctx.add(info, eos, 0.6, 1e300, max_steps)
    if slip:
        ctx.set_gfm_flags(use_slip_BCs=True)
    ctx.commit()
    return ctx, k


@pytest.mark.parametrize("variant,slip", [(G.GFM_VOFGFM, False), (G.GFM_VOFGFM, True), (G.GFM_OGFMVOF, False)])
def test_vof_ghost_states(oracle, variant, slip):
    print("CASE vof ghost_states variant=%d slip=%d" % (variant, slip), flush=True)
    info, eos, g1, g2, phi = bubble_state()
    ctx, k = ctx_of(info, eos, variant, slip=slip)
    ctx.upload(k, g1, g2, phi)
    ctx.set_ghost_states()
    ctx.sync()
    o1, o2, _ = ctx.download(k)
    vf = ctx.download_vfield(k)
    b1, b2 = ctx.bbox(k)
    n, _ = ctx.iface_info(k)
    st = ctx.status(k)
    ctx.close()
    r1, r2 = g1.copy(), g2.copy()
    bb1, bb2 = np.zeros(4, dtype=np.int32), np.zeros(4, dtype=np.int32)
    vo = np.zeros(phi.shape + (2,))
    ost = O.lib().or_set_ghost_states_flags(ogrid(info), oeos(eos), variant, O._p(phi), O._p(r1), O._p(r2),
                                            O._p(bb1), O._p(bb2), O._p(vo), 0, int(slip))
    assert ost == 0 and st == 0
    e1, _ = scaled_err(o1, r1)
    e2, _ = scaled_err(o2, r2)
    ev, _ = scaled_err(vf, vo)
    assert max(e1, e2) <= 1e-10, (e1, e2)
    assert ev <= 1e-10, ev
    assert list(bb1) == b1 and list(bb2) == b2
    if variant == G.GFM_OGFMVOF:
        assert b1 != [info.numGC, info.numGC + info.Ny, info.numGC, info.numGC + info.Nx]   # narrowed
    # the mixed cells of the interior grown by three cells, as the oracle's interface counts them
    H, W = phi.shape
    z = np.zeros((H, W))
    O.lib().or_vof_interface(ogrid(info), O._p(phi), O._p(z), O._p(np.zeros((H, W, 2))), O._p(np.zeros((H, W, 2))))
    g = info.numGC
    zr = z[g - 3:g + info.Ny + 3, g - 3:g + info.Nx + 3]
    assert n == int(np.count_nonzero((zr > 0) & (zr < 1))) and n > 0
    # the fill changed something within three cells of the interface, nothing beyond
    assert not np.array_equal(o1, g1) and not np.array_equal(o2, g2)


@pytest.mark.parametrize("variant", VOF)
def test_vof_steps(oracle, variant):
    print("CASE vof steps variant=%d" % variant, flush=True)
    nsteps = 60
    info, eos, g1, g2, phi = bubble_state(steps=20)
    ctx, k = ctx_of(info, eos, variant)
    ctx.upload(k, g1, g2, phi)
    ctx.step_graph(nsteps)
    ctx.sync()
    o1, o2, op = ctx.download(k)
    t, n, _ = ctx.get_time(k)
    st = ctx.status(k)
    ctx.close()
    assert n == nsteps
    run = O.Run(variant, 0, 5, 5, 10, 0.6, 1e300, nsteps)
    s = np.zeros(1, dtype=np.int32)
    to = np.zeros(1)
    r1, r2, rp = g1.copy(), g2.copy(), phi.copy()
    ost = O.lib().or_twofluid_run(ogrid(info), oeos(eos), run, O._p(r1), O._p(r2), O._p(rp), O._p(s), O._p(to))
    assert ost == 0 and st == 0
    assert abs(t - to[0]) <= 1e-12 * to[0]
    for a, b in ((o1, r1), (o2, r2)):
        linf, l1 = scaled_err(P.interior(info, a), P.interior(info, b))
        assert linf <= 1e-8 and l1 <= 1e-10, (linf, l1)
    pi, pr = P.interior(info, op), P.interior(info, rp)
    assert np.array_equal(pi <= 0, pr <= 0)
    assert np.abs(pi - pr).max() <= 1e-8 * np.abs(pr).max()


@pytest.mark.parametrize("variant", VOF)
def test_vof_steps_periodic_y(oracle, variant):
    """single-mode Richtmyer-Meshkov instance, transmissive in x and periodic in y: the mixed cells of
    the three ghost rows beyond each periodic side are solved too (GFM_VOF.cpp:73-75), 40 steps"""
    print("CASE vof steps periodic variant=%d" % variant, flush=True)
    nsteps = 40
    info, eos, (g1, g2, phi) = P.rmi(128, 32)
    ctx, k = ctx_of(info, eos, variant)
    ctx.upload(k, g1, g2, phi)
    ctx.step_graph(nsteps)
    ctx.sync()
    o1, o2, op = ctx.download(k)
    t, n, _ = ctx.get_time(k)
    st = ctx.status(k)
    ctx.close()
    assert n == nsteps and st == 0
    run = O.Run(variant, 0, 5, 5, 10, 0.6, 1e300, nsteps)
    s = np.zeros(1, dtype=np.int32)
    to = np.zeros(1)
    r1, r2, rp = g1.copy(), g2.copy(), phi.copy()
    assert O.lib().or_twofluid_run(ogrid(info), oeos(eos), run, O._p(r1), O._p(r2), O._p(rp), O._p(s),
                                   O._p(to)) == 0
    assert abs(t - to[0]) <= 1e-12 * to[0]
    for a, b in ((o1, r1), (o2, r2)):
        linf, l1 = scaled_err(P.interior(info, a), P.interior(info, b))
        assert linf <= 1e-8 and l1 <= 1e-10, (linf, l1)
    pi, pr = P.interior(info, op), P.interior(info, rp)
    assert np.array_equal(pi <= 0, pr <= 0)


@pytest.mark.parametrize("variant", VOF)
def test_vof_schedules_agree(variant, monkeypatch):
    """the dataflow solve of the mixed cells and its plain-rounds form give the same bits, and a
    batch of two instances on two streams gives each instance's single-context bits"""
    print("CASE vof schedules variant=%d" % variant, flush=True)
    info, eos, g1, g2, phi = bubble_state(steps=20)
    info2, eos2, st2 = P.shocked_bubble(128, 40, bubble_gamma=1.4, bubble_rho=3.0)
    out = []
    for plain in ("0", "1"):
        monkeypatch.setenv("GFM_IFACE_PLAIN_ROUNDS", plain)
        ctx, k = ctx_of(info, eos, variant)
        ctx.upload(k, g1, g2, phi)
        ctx.step_graph(30)
        ctx.sync()
        assert ctx.status(k) == 0
        out.append(ctx.download(k))
        ctx.close()
    monkeypatch.delenv("GFM_IFACE_PLAIN_ROUNDS")
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)
    monkeypatch.setenv("GFM_STEP_GROUPS", "2")
    ctx = G.Context(G.methods(sim_type=G.SIM_TWOFLUID, fs=G.FS_MUSCL, gfm=variant, ls_order=5, reinit_every=5,
                              reinit_iters=10, precision=G.FP64))
    ka = ctx.add(info, eos, 0.6)
    kb = ctx.add(info2, eos2, 0.6)
    ctx.commit()
    ctx.upload(ka, g1, g2, phi)
    ctx.upload(kb, *st2)
    ctx.step_graph(30)
    ctx.sync()
    assert ctx.status(ka) == 0 and ctx.status(kb) == 0
    got = ctx.download(ka)
    ctx.close()
    for a, b in zip(got, out[0]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("variant", VOF)
def test_vof_mixed_vs_oracle(variant):
    """the mixed path on a reduced shocked bubble, 100 steps, against the FP64 oracle on the same
    inputs: the bounds of the other variants' mixed-precision parity"""
    print("CASE vof mixed variant=%d" % variant, flush=True)
    steps = 100
    info, eos, st = P.shocked_bubble(256, 72, bubble_gamma=1.4, bubble_rho=3.0)
    ctx = G.Context(G.methods(sim_type=G.SIM_TWOFLUID, fs=G.FS_MUSCL, gfm=variant, ls_order=5, reinit_every=5,
                              reinit_iters=10, precision=G.MIXED))
    k = ctx.add(info, eos, 0.6, 1e300, steps)
    ctx.commit()
    ctx.upload(k, *st)
    ctx.step_graph(steps)
    ctx.sync()
    got, (t, n, _), status = ctx.download(k), ctx.get_time(k), ctx.status(k)
    ctx.close()
    r1, r2, rp, ost, ot = O.twofluid_run(info, eos, st, variant, steps)
    assert ost == 0 and status == 0 and n == steps
    assert abs(t - ot) <= 1e-5 * ot
    e = A.compare(info, got, (r1, r2, rp))
    print("vof mixed vs oracle:", {kk: "%.2e" % v for kk, v in e.items()}, flush=True)
    for name in A.NAMES:
        assert e["L1_" + name] <= 3e-6, (name, e)
        assert e["Linf_" + name] <= 2e-5, (name, e)
    assert e["contour_L1_rel"] <= 1e-3, e


def test_vof_rejects_what_the_reference_asserts():
    info, eos, st = P.shocked_bubble(64, 24)
    ctx = G.Context(G.methods(sim_type=G.SIM_TWOFLUID, gfm=G.GFM_OGFMVOF))
    ctx.add(info, eos, 0.6)
    with pytest.raises(RuntimeError):
        ctx.set_gfm_flags(use_slip_BCs=True)                   # OGFM_VOF.cpp:20-21
    ctx.close()
