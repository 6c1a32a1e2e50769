"""Pins of the oracle's VOF-based ghost fluid methods (GFM_VOF.cpp:8-325, OGFM_VOF.cpp:8-274) and of the
VOF interface they read (or_vof_interface), against what the mathematics and the reference's rules fix.

The reference takes z, the mixed-cell normal and the interface location from trackers of the absent
PhD-Interface-tracking-methods library (sim_twofluid.cpp:571-589); the oracle derives them from the level
set (the PLIC line through each cell centre with phi's value and unit normal there), so these check that
restatement against independent geometry, and the ghost fills against closed forms:

1. Geometry on an exact plane phi = n . x - c (its centred gradient is exact): z is the area of the cell
   on the side phi <= 0, computed here by clipping the cell polygon (Sutherland-Hodgman) and the shoelace
   formula; the interface location is the midpoint of the clipped polygon's edge on the line.  On a
   circle the volume fractions sum to the circle's area (O(h^2)) and the normals point outward.
2. A sheared contact (equal pressure, no normal velocity, each fluid its own density, gas and tangential
   velocity), vertical and horizontal, with the mixed column on either side of z = 1/2: one ghost fill
   gives the closed forms of the reference's rules.  GFM_VOF: the star state is the contact itself
   (p* = p, u* n = 0, each star density its own fluid's), the tangential velocity of both star states is
   the majority fluid's (no-slip, GFM_VOF.cpp:166-177) or each fluid's own (slip, :161-165), and the pure
   cells within three cells of the mixed cells take the other fluid's star state (:221-287).  OGFM_VOF:
   every ghost takes the real fluid's velocity and pressure and its own fluid's density (isentropic from
   its own probe, :154-167, :224-227, :257-260).  Cells beyond three of a mixed cell keep their state;
   the velocity field is the real fluid's (zero outside the ring of three for OGFM_VOF, :171-180);
   OGFM_VOF's real-cell boxes are the fluids' extents grown by two (:45-93).
3. The two-material shock tubes ST1_x and ST2_y (sim_twofluid.cpp:367-382, :675-742) with the initial
   interface off the cell faces (a face-aligned interface has no mixed cell, so the VOF methods have
   nothing to fill): GFM_VOF converges on both to the exact two-material Riemann solution with the zero
   contour at the exact contact; OGFM_VOF on ST1_x, and on ST2_y it produces an unphysical state, as the
   original GFM does at that impedance ratio.
4. A uniform flow of one gas stays uniform under both.
"""
import numpy as np
import pytest

import oracle_lib as O
import test_oracle_pins as P


def clip(poly, a, b, c):
    """the part of polygon `poly` with a x + b y <= c (Sutherland-Hodgman against one line)"""
    out = []
    n = len(poly)
    for k in range(n):
        p, q = poly[k], poly[(k + 1) % n]
        fp, fq = a * p[0] + b * p[1] - c, a * q[0] + b * q[1] - c
        if fp <= 0:
            out.append(p)
        if (fp < 0 < fq) or (fq < 0 < fp):
            t = fp / (fp - fq)
            out.append((p[0] + t * (q[0] - p[0]), p[1] + t * (q[1] - p[1])))
    return out


def area(poly):
    if len(poly) < 3:
        return 0.0
    x = np.array([p[0] for p in poly])
    y = np.array([p[1] for p in poly])
    return 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))


def interface(G, phi):
    H, W = G.shape
    z, nrm, loc = np.zeros((H, W)), np.zeros((H, W, 2)), np.zeros((H, W, 2))
    O.lib().or_vof_interface(G, O._p(phi), O._p(z), O._p(nrm), O._p(loc))
    return z, nrm, loc


def centres(G):
    H, W = G.shape
    X, Y = np.meshgrid(G.x0 + (np.arange(W) - G.g + 0.5) * G.dx, G.y0 + (np.arange(H) - G.g + 0.5) * G.dy)
    return X, Y


@pytest.mark.parametrize("seed", range(6))
def test_plane_fraction_and_chord_midpoint(oracle, seed):
    rng = np.random.default_rng(seed)
    G = O.Grid(12, 10, 6, 0.0, 0.0, 0.1, 0.07, 0, 0, 0, 0)
    th = rng.uniform(0, 2 * np.pi)
    if seed == 0:
        th = 0.5 * np.pi          # an axis-aligned normal: one of a, b vanishes
    a, b = np.cos(th), np.sin(th)
    X, Y = centres(G)
    c = a * 0.6 + b * 0.35 + rng.uniform(-0.02, 0.02)
    phi = np.ascontiguousarray(a * X + b * Y - c)
    z, nrm, loc = interface(G, phi)
    H, W = G.shape
    nmix = 0
    for i in range(1, H - 1):
        for j in range(1, W - 1):
            cx, cy = X[i, j], Y[i, j]
            hx, hy = 0.5 * G.dx, 0.5 * G.dy
            cell = [(cx - hx, cy - hy), (cx + hx, cy - hy), (cx + hx, cy + hy), (cx - hx, cy + hy)]
            part = clip(cell, a, b, c)
            assert z[i, j] == pytest.approx(area(part) / (G.dx * G.dy), abs=1e-12)
            assert nrm[i, j] == pytest.approx([a, b], abs=1e-12)
            if 1e-9 < z[i, j] < 1 - 1e-9:
                nmix += 1
                on = [p for p in part if abs(a * p[0] + b * p[1] - c) <= 1e-12]
                on = sorted(set((round(p[0], 12), round(p[1], 12)) for p in on))
                assert len(on) == 2, on
                mid = 0.5 * (np.array(on[0]) + np.array(on[1]))
                assert loc[i, j] == pytest.approx(mid, abs=1e-11)
    assert nmix >= 10


def test_circle_volume_and_normals(oracle):
    N, R, xc, yc = 128, 0.3, 0.5, 0.47
    G = O.Grid(N, N, 6, 0.0, 0.0, 1.0 / N, 1.0 / N, 0, 0, 0, 0)
    X, Y = centres(G)
    phi = np.ascontiguousarray(np.hypot(X - xc, Y - yc) - R)
    z, nrm, loc = interface(G, phi)
    assert z.sum() * G.dx * G.dy == pytest.approx(np.pi * R * R, rel=2e-4)
    mixed = (z > 0) & (z < 1)
    r = np.hypot(loc[..., 0] - xc, loc[..., 1] - yc)
    assert np.abs(r[mixed] - R).max() <= 0.05 / N                     # locations on the circle
    out = np.stack([X - xc, Y - yc], -1) / np.hypot(X - xc, Y - yc)[..., None]
    assert np.abs(nrm[mixed] - out[mixed]).max() <= 1e-3               # outward, into fluid 2


def shear_setup(axis, frac, eos):
    """fluid 1 below/left of a planar interface, fluid 2 beyond; both at rest normal to it with
    pressure 2; tangential velocities 0.3 (fluid 1) and -0.5 (fluid 2); the mixed cells hold a
    fraction `frac` of fluid 1; every state more than five cells from the interface on the other
    fluid's side is a distinct sentinel"""
    G = O.Grid(20, 16, 8, 0.0, 0.0, 0.05, 0.0625, 0, 0, 0, 0)
    H, W = G.shape
    X, Y = centres(G)
    s = X if axis == 0 else Y
    h = G.dx if axis == 0 else G.dy
    pos = 0.5 + frac * h          # 0.5 is a cell face: the interface cuts the next cell at `frac`
    phi = np.ascontiguousarray(s - pos)
    W1 = np.zeros(4); W2 = np.zeros(4)
    W1[0], W1[3] = 1.7, 2.0
    W2[0], W2[3] = 0.4, 2.0
    W1[2 - axis] = 0.3
    W2[2 - axis] = -0.5
    sent = O.arr(7.0, -3.0, 2.5, 3.0)
    g1 = np.zeros((H, W, 4)); g2 = np.zeros((H, W, 4))
    far1 = phi > 5 * h        # fluid 1 as a ghost beyond five cells
    far2 = phi < -5 * h
    U1, U2 = O.prims_to_cons(eos[0], eos[1], W1), O.prims_to_cons(eos[2], eos[3], W2)
    g1[:] = U1; g2[:] = U2
    g1[far1] = O.prims_to_cons(eos[0], eos[1], sent)
    g2[far2] = O.prims_to_cons(eos[2], eos[3], sent)
    return G, O.Eos(*eos), phi, np.ascontiguousarray(g1), np.ascontiguousarray(g2), W1, W2, sent, h


def dilate(mask, r):
    """cells within Chebyshev distance r of a cell of `mask`"""
    out = np.zeros_like(mask)
    H, W = mask.shape
    for di in range(-r, r + 1):
        for dj in range(-r, r + 1):
            out[max(di, 0):H + min(di, 0), max(dj, 0):W + min(dj, 0)] |= \
                mask[max(-di, 0):H + min(-di, 0), max(-dj, 0):W + min(-dj, 0)]
    return out


def prims(eos, g, pinf, U):
    out = np.zeros_like(U)
    for idx in np.ndindex(U.shape[:-1]):
        out[idx] = O.cons_to_prims(g, pinf, np.ascontiguousarray(U[idx]))
    return out


@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("frac", [0.3, 0.7])
@pytest.mark.parametrize("variant,slip", [(O.GFM_VOF, 0), (O.GFM_VOF, 1), (O.GFM_OGFM_VOF, 0)])
def test_sheared_contact_closed_form(oracle, axis, frac, variant, slip):
    eos = (1.4, 0.0, 1.67, 0.1)
    G, E, phi, g1, g2, W1, W2, sent, h = shear_setup(axis, frac, eos)
    H, W = G.shape
    g = G.g
    bb1, bb2 = np.zeros(4, dtype=np.int32), np.zeros(4, dtype=np.int32)
    vf = np.zeros((H, W, 2))
    st = O.lib().or_set_ghost_states_flags(G, E, variant, O._p(phi), O._p(g1), O._p(g2), O._p(bb1), O._p(bb2),
                                           O._p(vf), 0, slip)
    assert st == 0
    P1, P2 = prims(E, E.gamma1, E.pinf1, g1), prims(E, E.gamma2, E.pinf2, g2)
    z, _, _ = interface(G, phi)
    t = 2 - axis                                           # the tangential velocity component
    # majority-fluid tangential velocity of the star states (no-slip), or each fluid's own (slip)
    tmaj = W1[t] if frac > 0.5 else W2[t]
    if variant == O.GFM_VOF:
        star1_t, star2_t = (W1[t], W2[t]) if slip else (tmaj, tmaj)
    interior = np.zeros((H, W), bool)
    interior[g:g + G.Ny, g:g + G.Nx] = True
    d = phi / h                                            # signed distance in cells
    mixed = (z > 0) & (z < 1)
    assert mixed[interior].sum() == (G.Ny if axis == 0 else G.Nx)
    near = dilate(mixed, 3)
    ring1 = interior & (z == 0) & near                     # fluid 1 ghosts filled (within 3 of mixed)
    ring2 = interior & (z == 1) & near
    assert ring1.sum() == 3 * (G.Ny if axis == 0 else G.Nx) and ring2.sum() == ring1.sum()

    def expect(Pk, mask, rho, tang):
        want = np.zeros(4)
        want[0], want[3], want[t] = rho, 2.0, tang
        np.testing.assert_allclose(Pk[mask], np.broadcast_to(want, Pk[mask].shape), rtol=1e-10, atol=1e-12)

    mix_i = interior & mixed
    if variant == O.GFM_VOF:
        expect(P1, ring1, W1[0], star1_t)                  # fluid 1 ghost = fluid 1 star state
        expect(P2, ring2, W2[0], star2_t)
        if frac > 0.5:
            expect(P1, mix_i, W1[0], W1[t])                # real: own state (isentropic at p* = p)
            expect(P2, mix_i, W2[0], star2_t)
        else:
            expect(P2, mix_i, W2[0], W2[t])
            expect(P1, mix_i, W1[0], star1_t)
    else:
        expect(P1, ring1, W1[0], W2[t])                    # ghost: real fluid's u, v, p; own density
        expect(P2, ring2, W2[0], W1[t])
        if frac > 0.5:
            expect(P1, mix_i, W1[0], W1[t])
            expect(P2, mix_i, W2[0], W1[t])
        else:
            expect(P2, mix_i, W2[0], W2[t])
            expect(P1, mix_i, W1[0], W2[t])
    # beyond three cells of the mixed cells nothing changes
    far1 = interior & ~near & (d > 0) & (d <= 5.0)         # fluid 1's ghosts beyond the ring
    far2 = interior & ~near & (d < 0) & (d >= -5.0)
    assert far1.any() and far2.any()
    expect(P1, far1, W1[0], W1[t])
    expect(P2, far2, W2[0], W2[t])
    deep = interior & (d > 5.5)
    np.testing.assert_allclose(P1[deep], np.broadcast_to(sent, P1[deep].shape), rtol=1e-12)
    # velocity field: the real fluid's
    real1 = z >= 0.5
    want_v = np.where(real1[..., None], P1[..., 1:3], P2[..., 1:3])
    inner = np.zeros((H, W), bool)
    inner[g - 3:g + G.Ny + 3, g - 3:g + G.Nx + 3] = True
    np.testing.assert_allclose(vf[interior], want_v[interior], rtol=1e-12, atol=1e-14)
    if variant == O.GFM_OGFM_VOF:
        assert np.all(vf[~inner] == 0.0)
        # real cells: each fluid's extent grown by two, clamped to the domain
        ii, jj = np.nonzero(z >= 0.5)
        want1 = [max(ii.min() - 2, g), min(ii.max() + 2, g + G.Ny), max(jj.min() - 2, g), min(jj.max() + 2, g + G.Nx)]
        ii, jj = np.nonzero(z < 0.5)
        want2 = [max(ii.min() - 2, g), min(ii.max() + 2, g + G.Ny), max(jj.min() - 2, g), min(jj.max() + 2, g + G.Nx)]
        assert list(bb1) == want1 and list(bb2) == want2
        assert (want1[3] < g + G.Nx) if axis == 0 else (want1[1] < g + G.Ny)
    else:
        assert list(bb1) == [g, g + G.Ny, g, g + G.Nx] and list(bb2) == list(bb1)


def test_ogfm_vof_rejects_slip(oracle):
    eos = (1.4, 0.0, 1.67, 0.1)
    G, E, phi, g1, g2, *_ = shear_setup(0, 0.3, eos)
    H, W = G.shape
    bb = np.zeros(4, dtype=np.int32)
    st = O.lib().or_set_ghost_states_flags(G, E, O.GFM_OGFM_VOF, O._p(phi), O._p(g1), O._p(g2), O._p(bb),
                                           O._p(bb.copy()), O._p(np.zeros((H, W, 2))), 0, 1)
    assert st == 4


SHIFT = {"ST1_x": 0.0037, "ST2_y": 0.0148}   # the interface off the cell faces at N = 100, 200, 400


@pytest.mark.parametrize("case,variant", [("ST1_x", O.GFM_VOF), ("ST2_y", O.GFM_VOF), ("ST1_x", O.GFM_OGFM_VOF)])
def test_two_material_tube_converges(oracle, case, variant):
    out = [P.tube_run(case, N, variant, shift=SHIFT[case]) for N in (100, 200, 400)]
    assert all(o[0] == 0 for o in out)
    e = [o[1] for o in out]
    assert e[2] <= 0.5 * e[0] and e[2] <= 3e-3, e
    assert all(abs(o[2]) <= 0.5 for o in out), [o[2] for o in out]


def test_ogfm_vof_fails_on_stiff_tube(oracle):
    st, _, _ = P.tube_run("ST2_y", 100, O.GFM_OGFM_VOF, shift=SHIFT["ST2_y"])
    assert st & 1


@pytest.mark.parametrize("variant", [O.GFM_VOF, O.GFM_OGFM_VOF])
def test_uniform_flow_stays_uniform(oracle, variant):
    g = 8
    G = O.Grid(48, 40, g, 0.0, 0.0, 1.0 / 48, 1.0 / 40, 0, 0, 0, 0)
    E = O.Eos(1.4, 0.0, 1.4, 0.0)
    H, W = G.shape
    U = O.prims_to_cons(1.4, 0.0, O.arr(1.3, 0.3, 0.2, 2.0))
    g1 = np.ascontiguousarray(np.broadcast_to(U, (H, W, 4)).copy())
    g2 = g1.copy()
    X, Y = centres(G)
    phi = np.ascontiguousarray(np.hypot(X - 0.45, Y - 0.5) - 0.2)
    p0 = phi.copy()
    R = O.Run(variant, O.FS_MUSCL, 5, 5, 10, 0.5, 1e300, 20)
    s = np.zeros(1, dtype=np.int32)
    t = np.zeros(1)
    assert O.lib().or_twofluid_run(G, E, R, O._p(g1), O._p(g2), O._p(phi), O._p(s), O._p(t)) == 0
    for a in (g1, g2):
        dev = np.abs(a[g:g + 40, g:g + 48] - U).reshape(-1, 4).max(0) / np.abs(U).max()
        assert dev.max() <= 1e-13, dev
    assert not np.array_equal(phi <= 0, p0 <= 0)
