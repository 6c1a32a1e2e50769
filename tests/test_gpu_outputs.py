"""The output diagnostics of sim_twofluid on the GPU: the output hold (the step after which the
reference writes its periodic output, sim_twofluid.cpp:79-83), the perturbation-amplitude tracers
(:1831-1887) and z (:199), each against the oracle or the reference's rule on the same inputs.

- Hold: a run that stops wherever t - last > interval and is resumed by the host stops exactly at the
  steps that rule picks from an unheld run's times, and ends bit for bit where the unheld run ends
  (holding changes nothing but when the host gets to look).
- Tracers: every step's recorded (t, x, y, u, v) and the RK4 move against or_tracers_step applied to
  that step's downloaded velocity field (FP64 and the mixed-precision path, whose FP32 velocity field
  converts to FP64 exactly): 1e-13 of scale.  Graph replays record the same history as eager steps.
- z: gfm_download_z against or_volume_fraction on the downloaded phi (FP64: 1e-14; FP32 storage:
  one FP32 rounding, 1e-7).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import gfm_hip as G
import oracle_lib as O
import problems as P

AIR = [1.0, 0.0, 0.0, 1.0]
POST = [1.411, -0.39, 0.0, 1.628]
SF6 = [5.04, 0.0, 0.0, 1.0]
TRACERS = [[2.88, 1.0], [2.92, 0.5]]


def ogrid(info):
    return O.Grid(info.Nx, info.Ny, info.numGC, info.x0, info.y0, info.dx, info.dy,
                  info.BC_L, info.BC_R, info.BC_B, info.BC_T)


def rmi_pert(Nx=256, Ny=64):
    """the reference's RMI_pert case (sim_twofluid.cpp:430-442, :1348-1453): reflective in y,
    interface x = 2.9 - 0.02 cos(2 pi y), shock at x = 3.2"""
    info = G.SimInfo(Nx, Ny, 8, 0.0, 0.0, 4.0 / Nx, 1.0 / Ny, G.BC_TRANSMISSIVE, G.BC_TRANSMISSIVE,
                     G.BC_REFLECTIVE, G.BC_REFLECTIVE)
    eos = G.Eos(1.4, 0.0, 1.093, 0.0)
    return info, eos, G.ic_rmi(info, eos, 2.9, 0.02, 1, AIR, POST, SF6, 3.2)


def ctx_of(info, eos, precision, max_steps, tracers=None, gfm=G.GFM_RIEMANN):
    ctx = G.Context(G.methods(sim_type=G.SIM_TWOFLUID, fs=G.FS_MUSCL, gfm=gfm, ls_order=5,
                              reinit_every=5, reinit_iters=10, precision=precision))
    k = ctx.add(info, eos, 0.5, 4.0, max_steps)
    if tracers is not None:
        ctx.set_tracers(k, tracers, keep_y=True)
    ctx.commit()
    return ctx, k


@pytest.mark.parametrize("graph", [True, False])
def test_output_hold_follows_reference_rule(graph):
    print("CASE output hold graph=%d" % graph, flush=True)
    n = 60
    info, eos, st = P.shocked_bubble(160, 44)
    free, k = ctx_of(info, eos, G.FP64, n)
    free.upload(k, *st)
    free.step_graph(n)
    free.compute_dt()                       # records t_n
    free.sync()
    times = free.mass_history(k, n + 1)[:, 0]
    ref_state = free.download(k)
    free.close()
    ivl = times[n] / 4.3
    expect, last = [], 0.0
    for s in range(1, n + 1):
        if times[s] - last > ivl:
            expect.append(s)
            last = times[s]
    assert len(expect) >= 3
    ctx, k = ctx_of(info, eos, G.FP64, n)
    ctx.upload(k, *st)
    ctx.set_output_hold(k, 0.0, ivl)
    got, last, nstep, prev = [], 0.0, 0, -1
    while nstep != prev:
        prev = nstep
        (ctx.step_graph if graph else ctx.step)(7)
        t, nstep, _ = ctx.get_time(k)
        if t - last > ivl:
            assert t == times[nstep]
            got.append(nstep)
            last = t
            ctx.set_output_hold(k, last, ivl)
            prev = -1
    assert nstep == n and got == expect, (got, expect)
    out = ctx.download(k)
    for a, b in zip(out, ref_state):
        assert np.array_equal(a, b)
    ctx.close()


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED])
def test_tracers_vs_oracle(oracle, precision):
    print("CASE tracers precision=%d" % precision, flush=True)
    n = 30
    info, eos, st = rmi_pert()
    ctx, k = ctx_of(info, eos, precision, n, TRACERS)
    ctx.upload(k, *st)
    ctx.ls_reinit(40)
    Gr = ogrid(info)
    pos = np.array(TRACERS, dtype=np.float64)
    p0 = pos.copy()
    entries = np.zeros((n + 1, 9))
    t_prev = 0.0
    for s in range(1, n + 1):
        ctx.step(1)
        t, nstep, dt = ctx.get_time(k)
        assert nstep == s and ctx.status(k) == 0
        vf = np.ascontiguousarray(ctx.download_vfield(k))
        e = np.zeros(9)
        oracle.lib().or_tracers_step(Gr, O._p(vf), 2, O._p(pos), O._p(p0), 1, t, dt, O._p(e))
        entries[s] = e
        assert t > t_prev
        t_prev = t
    h = ctx.tracer_history(k, n + 1)
    assert np.all(h[0] == 0.0)
    assert np.array_equal(h[1:, 0], entries[1:, 0])
    assert np.allclose(h[1:, [1, 5]], entries[1:, [1, 5]], rtol=1e-13, atol=0)       # x
    assert np.array_equal(h[1:, [2, 6]], np.tile([1.0, 0.5], (n, 1)))                 # kept rows
    scale = np.abs(entries[1:, [3, 4, 7, 8]]).max()
    assert np.abs(h[1:, [3, 4, 7, 8]] - entries[1:, [3, 4, 7, 8]]).max() <= 1e-13 * scale
    # the first line of -pertamp.dat: the points as placed, amplitude 0.02
    assert h[1, 1] == 2.88 and h[1, 5] == 2.92
    ctx.close()
    # graph replays: the same history
    g, k = ctx_of(info, eos, precision, n + 1, TRACERS)
    g.upload(k, *st)
    g.ls_reinit(40)
    g.step_graph(n)
    g.sync()
    assert np.array_equal(g.tracer_history(k, n + 1), h)
    # resetting the tracers after commit restarts them from their new positions
    g.set_tracers(k, [[2.5, 1.0], [2.6, 0.5]], keep_y=True)
    g.step_graph(1)
    g.sync()
    h2 = g.tracer_history(k, n + 2)
    assert h2[n + 1, 1] == 2.5 and h2[n + 1, 5] == 2.6
    g.close()


@pytest.mark.parametrize("precision,tol", [(G.FP64, 1e-14), (G.MIXED, 1e-7)])
def test_z_vs_oracle(oracle, precision, tol):
    print("CASE z precision=%d" % precision, flush=True)
    info, eos, st = P.shocked_bubble(160, 44)
    ctx, k = ctx_of(info, eos, precision, 1 << 20)
    ctx.upload(k, *st)
    ctx.step_graph(20)
    ctx.sync()
    z = ctx.download_z(k)
    _, _, phi = ctx.download(k)
    ref = np.zeros_like(phi)
    oracle.lib().or_volume_fraction(ogrid(info), O._p(np.ascontiguousarray(phi)), O._p(ref))
    assert np.abs(z - ref).max() <= tol
    assert ref.min() >= 0.0 and ref.max() <= 1.0
    cut = ((ref > 0) & (ref < 1)).sum()
    assert cut > 50
    # the step after a z download is the step without it (z uses a stage buffer)
    ctx.step_graph(5)
    ctx.sync()
    a = ctx.download(k)
    c2, k2 = ctx_of(info, eos, precision, 1 << 20)
    c2.upload(k2, *st)
    c2.step_graph(25)
    c2.sync()
    for x, y in zip(a, c2.download(k2)):
        assert np.array_equal(x, y)
    ctx.close()
    c2.close()


def test_tracer_and_hold_errors():
    info, eos, _ = P.shocked_bubble(64, 20)
    c = G.Context(G.methods(sim_type=G.SIM_TWOFLUID, gfm=G.GFM_RIEMANN))
    k = c.add(info, eos, 0.6)
    with pytest.raises(G.GfmError, match="0 to 4 tracers"):
        c.set_tracers(k, np.zeros((5, 2)))
    c.set_tracers(k, [[0.1, 0.02]])
    c.commit()
    with pytest.raises(G.GfmError, match="fixed at gfm_ctx_commit"):
        c.set_tracers(k, [[0.1, 0.02], [0.2, 0.03]], keep_y=False)
    c.close()
    one = G.Context(G.methods(sim_type=G.SIM_ONEFLUID))
    k = one.add(*P.circular_explosion(21)[:2], 0.6)
    with pytest.raises(G.GfmError, match="two-fluid"):
        one.set_tracers(k, [[0.1, 0.1]])
    one.close()
