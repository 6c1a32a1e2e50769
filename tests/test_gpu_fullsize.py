"""Full-size runs of the benchmark workloads in mixed precision, checked through properties that hold at
any size (the FP64 oracle cannot run these sizes in test time; test_gpu_mixed_oracle.py compares the
same workloads with it at reduced size or on subsets).

After 100 steps:
  - no status bit (no unphysical state, Riemann-solver failure or interfacial overflow);
  - for an instance whose shock (at x_s0 - M c0 t, c0 the pre-shock sound speed) is still more than
    4 cells from the interface: the shocked gas (the fluid at the right boundary) gains exactly the
    mass the constant post-shock state carries in through the transmissive right boundary,
    rho |u| Ly t (the boundary face's exact Riemann flux of two equal states; no wave has reached
    either boundary), to 1e-3 of that inflow; and the other fluid (the bubble, or the heavy gas of the
    RMI), at rest in uniform pressure, keeps its mass to 1e-6 relative;
  - an instance the shock has struck is only required to stay physical: the ghost fluid method does
    not conserve mass across a shocked interface, and the bubble-mass measure of the batch's struck
    instances moves by up to 7% within these 100 steps on the coarsest grids;
  - the single bubble on the mirror-symmetric domain (reflective top and bottom, bubble on the
    mid-line) stays mirror-symmetric: rho and E at row i equal those at row Ny-1-i, rho v equals
    minus its mirror, to 1e-4 of each component's scale (of the momentum magnitude for both momenta;
    FP32 rounding is not symmetric).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import gfm_hip as G
import problems as P

SEED = 20261203
STEPS = 100


def run(cases, variant, init_reinit=0):
    ctx = G.Context(G.methods(sim_type=G.SIM_TWOFLUID, fs=G.FS_MUSCL, gfm=variant, ls_order=5,
                              reinit_every=5, reinit_iters=10, precision=G.MIXED))
    for info, eos, _ in cases:
        ctx.add(info, eos, 0.6, 1e300, STEPS)
    ctx.commit()
    for k, (_, _, st) in enumerate(cases):
        ctx.upload(k, *st)
    if init_reinit:
        ctx.ls_reinit(init_reinit)
    ctx.step_graph(STEPS)
    ctx.sync()
    return ctx


def inflow_state(info, g1, g2, phi):
    """the fluid at the right boundary (0: fluid 1) and its post-shock density and velocity there (the
    last interior column)"""
    g = info.numGC
    i, j = g + info.Ny // 2, g + info.Nx - 1
    f = 0 if phi[i, j] <= 0.0 else 1
    U = (g1, g2)[f][i, j]
    return f, U[0], U[1] / U[0]


def untouched(info, meta, t):
    """the shock, entering from x_s0 at M c0, is still more than 4 cells from the interface"""
    xs0, mach, c0, x_iface = meta
    return xs0 - mach * c0 * t - x_iface > 4 * info.dx


def check_budget(ctx, cases, metas, label):
    worst1, worst2, n_untouched = 0.0, 0.0, 0
    for k, ((info, eos, (g1, g2, phi)), meta) in enumerate(zip(cases, metas)):
        assert ctx.status(k) == 0, (label, k, ctx.status(k))
        h = ctx.mass_history(k, STEPS)
        t, n, dt = ctx.get_time(k)
        assert n == STEPS
        m = ctx.mass(k)            # masses at the start of the last step: t = h[-1, 0]
        fin, rho_in, u_in = inflow_state(info, g1, g2, phi)
        fo = 1 - fin
        Ly = info.Ny * info.dy
        inflow = rho_in * abs(u_in) * Ly * h[-1, 0]
        d2 = (h[-1, 1 + fin] - h[0, 1 + fin] - inflow) / inflow
        d1 = (h[-1, 1 + fo] - h[0, 1 + fo]) / h[0, 1 + fo]
        if untouched(info, meta, h[-1, 0]):
            n_untouched += 1
            worst1, worst2 = max(worst1, abs(d1)), max(worst2, abs(d2))
            assert abs(d2) <= 1e-3, (label, k, d2)
            assert abs(d1) <= 1e-6, (label, k, d1)
        assert abs(m[1] - h[-1, 2]) <= 1e-12 * abs(m[1])
    print(label, "%d of %d instances untouched: other fluid's mass change %.2e, inflow budget error %.2e"
          % (n_untouched, len(cases), worst1, worst2), flush=True)
    return n_untouched


def bubble_meta(c):
    """shock start, Mach number, pre-shock sound speed (air, rho 1, p 1e5) and the interface's right edge
    (problems.shocked_bubble)"""
    return 0.225, c.get("mach", 1.22), np.sqrt(1.4 * 1e5), 0.175 + c.get("R", 0.025)


def rmi_meta(c):
    """problems.rmi: shock from x = 3.2 into air (rho 1, p 1), interface x = 2.9 - a sin(.)"""
    amp = c.get("amp_over_lambda", 0.05) * 1.0 / c.get("modes", 1)
    return 3.2, c.get("mach", 1.2), np.sqrt(1.4), 2.9 + amp


def test_large_fullsize_properties():
    print("CASE large 4096x1152 mixed GFM_riemann", flush=True)
    rng = np.random.default_rng(SEED)
    c = dict(Nx=4096, Ny=1152, bubble_gamma=float(rng.uniform(1.09, 1.67)), bubble_rho=float(rng.uniform(0.18, 5.0)))
    cases = [P.shocked_bubble(**c)]
    ctx = run(cases, G.GFM_RIEMANN)
    assert check_budget(ctx, cases, [bubble_meta(dict(c, mach=1.22, R=0.025))], "large") == 1
    info = cases[0][0]
    g1, g2, phi = ctx.download(0)
    side = (phi <= 0.0)[..., None]
    u = P.interior(info, np.where(side, g1, g2))
    m = u[::-1]
    scale = np.abs(u).reshape(-1, 4).max(0)
    scale[1] = scale[2] = np.hypot(u[..., 1], u[..., 2]).max()
    err = np.abs(np.stack([u[..., 0] - m[..., 0], u[..., 1] - m[..., 1], u[..., 2] + m[..., 2],
                           u[..., 3] - m[..., 3]], -1)).reshape(-1, 4).max(0) / scale
    print("mirror asymmetry per component", err, flush=True)
    assert (err <= 1e-4).all(), err
    pi = P.interior(info, phi) <= 0.0
    assert np.count_nonzero(pi != pi[::-1]) == 0
    ctx.close()


def test_batch_fullsize_properties():
    print("CASE batch of 64 mixed GFM_riemann", flush=True)
    cs = P.bubble_batch(64)
    cases = [P.shocked_bubble(**c) for c in cs]
    ctx = run(cases, G.GFM_RIEMANN)
    assert check_budget(ctx, cases, [bubble_meta(c) for c in cs], "batch") > 0
    ctx.close()


def test_rmi_fullsize_properties():
    print("CASE rmi batch of 256 mixed GFM_modified", flush=True)
    cs = P.rmi_batch(256)
    cases = [P.rmi(**c) for c in cs]
    ctx = run(cases, G.GFM_MODIFIED, init_reinit=40)
    assert check_budget(ctx, cases, [rmi_meta(c) for c in cs], "rmi") > 0
    ctx.close()


def test_long_interface_flags_in_global_memory():
    """A single 16384x4096 Richtmyer-Meshkov instance with a steep 16-mode interface (amplitude twice
    the wavelength; about 65536 interfacial cells, 16 per row at 1024x256 on the host) from the
    device initial conditions: its interfacial list is longer than the ordered solve's LDS flag array
    (49152 entries), so that solve keeps its flags in global memory. After 20 steps, before the shock
    (about 700 cells away) reaches the interface: no status bit, the whole list solved, and the
    barrier-round schedule gives the same FP64 mass history bit for bit."""
    steps = 20
    hist = []
    for plain in ("0", "1"):
        print("CASE long interface rmi 16384x4096 modes 16 mixed GFM_riemann plain_rounds=%s" % plain, flush=True)
        old = os.environ.get("GFM_IFACE_PLAIN_ROUNDS")
        os.environ["GFM_IFACE_PLAIN_ROUNDS"] = plain
        try:
            ctx = G.Context(G.methods(sim_type=G.SIM_TWOFLUID, fs=G.FS_MUSCL, gfm=G.GFM_RIEMANN, ls_order=5,
                                      reinit_every=5, reinit_iters=10, precision=G.MIXED))
            info, eos, desc = P.rmi_ic(Nx=16384, Ny=4096, amp_over_lambda=2.0, modes=16)
            ctx.add(info, eos, 0.6, 1e300, steps)
            ctx.commit()
        finally:
            if old is None:
                del os.environ["GFM_IFACE_PLAIN_ROUNDS"]
            else:
                os.environ["GFM_IFACE_PLAIN_ROUNDS"] = old
        ctx.ic_device([(0, desc)])
        ctx.step_graph(steps)
        ctx.sync()
        assert ctx.status(0) == 0, ctx.status(0)
        n, _ = ctx.iface_info(0)
        print("interfacial entries", n, flush=True)
        assert n > 49152
        assert ctx.get_time(0)[1] == steps
        h = ctx.mass_history(0, steps)
        assert np.isfinite(h).all() and (h[:, 1:] > 0).all()
        hist.append(h)
        ctx.close()
    assert np.array_equal(hist[0], hist[1])
