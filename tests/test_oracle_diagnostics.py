"""Pins of the oracle's output diagnostics against what the mathematics fixes.

1. or_tracers_step (output_perturbation_amplitude's tracers, sim_twofluid.cpp:1861-1885; the solved
   velocity field's get_velocity and time_step_RK4 are external and restated as bilinear
   interpolation and classical RK4).  In a linear velocity field w' = A w + b, which bilinear
   interpolation reproduces exactly, one RK4 step of dt is exactly
       w + dt (I + dt A / 2 + (dt A)^2 / 6 + (dt A)^3 / 24) (A w + b),
   so a wrong stage weight, a dropped stage, a half step taken as a whole one or u and v exchanged
   moves the result by far more than rounding.  A general A (u depends on y, v on x) catches a
   transposed operand.  keep_y puts y back on its initial value (:1882-1884).
2. or_volume_fraction (z, sim_twofluid.cpp:199; the level set's get_z is external): for a planar
   phi of any orientation the plane through each cell centre is the interface itself, so z must be the
   exact area fraction of the cell on phi's negative side, computed here by clipping the cell's
   rectangle with the half plane (Sutherland-Hodgman) and the shoelace formula.  Axis-aligned and
   diagonal normals exercise the degenerate branches.  For a circle the fractions sum to its area
   with a second-order error (the plane approximation of a curved interface).
"""
import math

import numpy as np
import pytest

import oracle_lib as O


def test_tracer_rk4_linear_field(oracle):
    G = O.Grid(20, 16, 4, -1.0, -0.5, 0.1, 0.0625, 0, 0, 0, 0)
    H, W = G.shape
    A = np.array([[1.3, -0.7], [0.45, -0.9]])
    b = np.array([0.2, -0.35])
    j, i = np.meshgrid(np.arange(W), np.arange(H))
    x = G.x0 + (j - G.g + 0.5) * G.dx
    y = G.y0 + (i - G.g + 0.5) * G.dy
    vf = np.stack([A[0, 0] * x + A[0, 1] * y + b[0], A[1, 0] * x + A[1, 1] * y + b[1]], axis=-1)
    vf = np.ascontiguousarray(vf)
    p0 = np.array([[0.3, 0.2], [-0.45, 0.1], [0.61, -0.33]])
    dt, t = 0.05, 1.25
    M = np.eye(2) + dt * A / 2 + (dt * A) @ (dt * A) / 6 + np.linalg.matrix_power(dt * A, 3) / 24
    for keep_y in (0, 1):
        pos = p0.copy()
        entry = np.zeros(1 + 4 * len(p0))
        oracle.lib().or_tracers_step(G, O._p(vf), len(p0), O._p(pos), O._p(p0), keep_y, t, dt, O._p(entry))
        assert entry[0] == t
        for k, w in enumerate(p0):
            vel = A @ w + b
            assert np.allclose(entry[1 + 4 * k:5 + 4 * k], [w[0], w[1], vel[0], vel[1]], rtol=1e-13, atol=1e-15)
            exact = w + dt * (M @ vel)
            assert abs(pos[k, 0] - exact[0]) <= 1e-14, (pos[k], exact)
            assert abs(pos[k, 1] - (w[1] if keep_y else exact[1])) <= 1e-14, (pos[k], exact)
    # two steps compose: the field is frozen, so two steps of dt are M applied twice
    pos = p0.copy()
    entry = np.zeros(1 + 4 * len(p0))
    for _ in range(2):
        oracle.lib().or_tracers_step(G, O._p(vf), len(p0), O._p(pos), O._p(p0), 0, t, dt, O._p(entry))
    for k, w in enumerate(p0):
        w1 = w + dt * (M @ (A @ w + b))
        w2 = w1 + dt * (M @ (A @ w1 + b))
        assert np.allclose(pos[k], w2, rtol=0, atol=1e-14)


def clipped_area(cx, cy, dx, dy, n, c0):
    """area of the rectangle centred at (cx, cy) where n . p + c0 <= 0"""
    poly = [(cx - dx / 2, cy - dy / 2), (cx + dx / 2, cy - dy / 2), (cx + dx / 2, cy + dy / 2),
            (cx - dx / 2, cy + dy / 2)]
    f = lambda p: n[0] * p[0] + n[1] * p[1] + c0
    out = []
    for k in range(4):
        a, b = poly[k], poly[(k + 1) % 4]
        fa, fb = f(a), f(b)
        if fa <= 0:
            out.append(a)
        if (fa < 0 < fb) or (fb < 0 < fa):
            s = fa / (fa - fb)
            out.append((a[0] + s * (b[0] - a[0]), a[1] + s * (b[1] - a[1])))
    if len(out) < 3:
        return 0.0
    return 0.5 * abs(sum(out[k][0] * out[(k + 1) % len(out)][1] - out[(k + 1) % len(out)][0] * out[k][1]
                         for k in range(len(out))))


@pytest.mark.parametrize("theta", [0.0, math.pi / 2, math.pi / 4, 0.3, 2.0, -1.1, math.pi])
def test_volume_fraction_planar_exact(oracle, theta):
    G = O.Grid(24, 18, 3, 0.0, 0.0, 0.1, 0.07, 0, 0, 0, 0)
    H, W = G.shape
    n = np.array([math.cos(theta), math.sin(theta)])
    p0 = np.array([1.23, 0.61])
    j, i = np.meshgrid(np.arange(W), np.arange(H))
    x = G.x0 + (j - G.g + 0.5) * G.dx
    y = G.y0 + (i - G.g + 0.5) * G.dy
    phi = np.ascontiguousarray(n[0] * (x - p0[0]) + n[1] * (y - p0[1]))
    z = np.zeros_like(phi)
    oracle.lib().or_volume_fraction(G, O._p(phi), O._p(z))
    c0 = -(n[0] * p0[0] + n[1] * p0[1])
    worst, cut = 0.0, 0
    for ii in range(1, H - 1):
        for jj in range(1, W - 1):
            ex = clipped_area(x[ii, jj], y[ii, jj], G.dx, G.dy, n, c0) / (G.dx * G.dy)
            worst = max(worst, abs(z[ii, jj] - ex))
            cut += 0.0 < ex < 1.0
    assert worst <= 1e-12, worst
    assert cut >= 10
    ring = np.concatenate([z[0], z[-1], z[:, 0], z[:, -1]])
    rphi = np.concatenate([phi[0], phi[-1], phi[:, 0], phi[:, -1]])
    assert np.array_equal(ring, (rphi <= 0).astype(float))


def circle_area_error(oracle, N, R=0.3):
    G = O.Grid(N, N, 3, 0.0, 0.0, 1.0 / N, 1.0 / N, 0, 0, 0, 0)
    H, W = G.shape
    j, i = np.meshgrid(np.arange(W), np.arange(H))
    x = G.x0 + (j - G.g + 0.5) * G.dx
    y = G.y0 + (i - G.g + 0.5) * G.dy
    phi = np.ascontiguousarray(np.hypot(x - 0.5, y - 0.5) - R)
    z = np.zeros_like(phi)
    oracle.lib().or_volume_fraction(G, O._p(phi), O._p(z))
    assert z.min() >= 0.0 and z.max() <= 1.0
    return abs(z.sum() * G.dx * G.dy - math.pi * R * R) / (math.pi * R * R)


def test_volume_fraction_circle_area(oracle):
    e1, e2 = circle_area_error(oracle, 50), circle_area_error(oracle, 200)
    assert e2 <= 1e-4, (e1, e2)
    assert e1 / e2 >= 2.0 ** 2.5, (e1, e2)        # second order (ratio 16 for 4x finer)
