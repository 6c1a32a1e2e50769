"""The reference-shaped C++ drop-ins of host/gfm2d.hpp on host grids (grideuler2type, BBrange,
GFM_ITM_interface), driven through tests/shim_check and compared with the oracle on the same inputs:

- FS_MUSCL / FS_godunov::pure_fluid_update (FS_base.hpp:23-32): the full box and a narrowed realcells
  box (cells outside it unchanged, bit for bit); 1e-12 of scale, as the flux-update parity tests;
- GFM_original / real / riemann / modified::set_ghost_states (GFM_base.hpp:37-47) with the GFM_base
  flags use_extension_vfield / use_slip_BCs (GFM_base.hpp:20-21) both ways: grids, realcells and the
  stored velocity field; 1e-10 of scale, as the ghost-fill parity tests;
- extrapolate_vector (extrapolate_vector.hpp:8-15), apply_BCs_euler (euler_bc.hpp:5-103, bit for bit),
  Levelset::update_interface_coupled (the HOUC advection, 1e-12) and the pure / mixed Riemann solvers
  (pure_RS_base.hpp:13-20, mixed_RS_base.hpp:12-24; 1e-12).
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import gfm_hip as G
import oracle_lib as O
import problems as P

HERE = os.path.dirname(os.path.abspath(__file__))
SHIM = os.path.join(HERE, "shim_check")


def header(mode, info):
    return [mode, info.Nx, info.Ny, info.numGC, info.x0, info.y0, info.dx, info.dy,
            info.BC_L, info.BC_R, info.BC_B, info.BC_T]


def run_shim(tmp_path, values, fails_with=None):
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    parts = [np.asarray(v, dtype=np.float64).reshape(-1) for v in values]
    np.concatenate(parts).tofile(str(src))
    out = subprocess.run([SHIM, str(src), str(dst)], capture_output=True, text=True, timeout=120)
    if fails_with:
        assert out.returncode == 1 and fails_with in out.stderr, (out.returncode, out.stderr)
        return None
    assert out.returncode == 0, out.stderr
    return np.fromfile(str(dst), dtype=np.float64)


def ogrid(info):
    return O.Grid(info.Nx, info.Ny, info.numGC, info.x0, info.y0, info.dx, info.dy,
                  info.BC_L, info.BC_R, info.BC_B, info.BC_T)


def scaled(a, b):
    a = a.reshape(-1, a.shape[-1])
    b = b.reshape(a.shape)
    return (np.abs(a - b).max(axis=0) / np.maximum(np.abs(b).max(axis=0), 1e-300)).max()


def developed_bubble(steps=30):
    info, eos, st = P.shocked_bubble(160, 44)
    g1, g2, phi, s, _ = O.twofluid_run(info, eos, st, O.GFM_RIEMANN, steps)
    assert s == 0
    return info, eos, g1, g2, phi


@pytest.mark.parametrize("scheme", [0, 1])
@pytest.mark.parametrize("narrow", [False, True])
def test_flux_update(tmp_path, oracle, scheme, narrow):
    print("CASE shim flux scheme=%d narrow=%d" % (scheme, narrow), flush=True)
    info, eos, grid = P.circular_explosion(101)
    grid = grid.copy()
    # a developed state: a few oracle steps first
    G_ = ogrid(info)
    t, dt = np.zeros(1), np.zeros(1)
    scratch = np.zeros_like(grid)
    for s in range(6):
        assert oracle.lib().or_onefluid_step(G_, 1.4, 0.0, 0, 0.2 if s < 5 else 0.7, 0.25, O._p(t), O._p(grid),
                                             O._p(scratch), O._p(dt)) == 0
    g = info.numGC
    bb = [g + 7, g + 70, g + 12, g + 90] if narrow else [g, g + info.Ny, g, g + info.Nx]
    step = 0.6 * dt[0]
    out = run_shim(tmp_path, [header(1, info), [scheme, 1.4, 0.0, step] + bb, grid])
    H, W = grid.shape[:2]
    got = out[:H * W * 4].reshape(H, W, 4)
    fut = out[H * W * 4:].reshape(H, W, 4)
    ref, rfut = grid.copy(), grid.copy()
    fn = oracle.lib().or_fs_muscl if scheme == 0 else oracle.lib().or_fs_godunov
    assert fn(G_, 1.4, 0.0, step, O._p(np.array(bb, dtype=np.int32)), O._p(ref), O._p(rfut)) == 0
    assert scaled(got, ref) <= 1e-12
    inside = np.zeros((H, W), dtype=bool)
    inside[bb[0]:bb[1], bb[2]:bb[3]] = True
    assert np.array_equal(got[~inside], grid[~inside])
    assert np.array_equal(fut[inside], got[inside])
    assert scaled(got[inside], grid[inside]) > 1e-6            # the update did something


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
@pytest.mark.parametrize("ext,slip", [(0, 0), (1, 0), (1, 1)])
def test_set_ghost_states(tmp_path, oracle, variant, ext, slip):
    print("CASE shim ghost variant=%d ext=%d slip=%d" % (variant, ext, slip), flush=True)
    info, eos, g1, g2, phi = developed_bubble()
    args = [header(2, info), [variant, ext, slip, eos.gamma1, eos.pinf1, eos.gamma2, eos.pinf2, 0.0], phi, g1, g2]
    if variant == 0 and slip:
        # the reference asserts the flag is off for GFM_original (GFM_original.cpp:21): an error, not a
        # silent fallback
        run_shim(tmp_path, args, fails_with="GFM_original")
        return
    out = run_shim(tmp_path, args)
    H, W = phi.shape
    n = H * W * 4
    o1, o2 = out[:n].reshape(H, W, 4), out[n:2 * n].reshape(H, W, 4)
    bb = out[2 * n:2 * n + 8].astype(int)
    vf = out[2 * n + 8:].reshape(H, W, 2)
    r1, r2 = g1.copy(), g2.copy()
    bb1, bb2 = np.zeros(4, dtype=np.int32), np.zeros(4, dtype=np.int32)
    vo = np.zeros((H, W, 2))
    assert oracle.lib().or_set_ghost_states_flags(ogrid(info), O.Eos(eos.gamma1, eos.pinf1, eos.gamma2, eos.pinf2),
                                                  variant, O._p(phi), O._p(r1), O._p(r2), O._p(bb1), O._p(bb2),
                                                  O._p(vo), ext, slip) == 0
    assert max(scaled(o1, r1), scaled(o2, r2)) <= 1e-10
    assert list(bb[:4]) == list(bb1) and list(bb[4:]) == list(bb2)
    g = info.numGC
    if ext:
        assert scaled(vf, vo) <= 1e-10
    else:
        assert scaled(vf[g:-g, g:-g], vo[g:-g, g:-g]) <= 1e-10


def test_extrapolate_vector(tmp_path, oracle):
    print("CASE shim extrapolate", flush=True)
    info, eos, g1, g2, phi = developed_bubble(20)
    rng = np.random.default_rng(5)
    p1, p2 = g1.copy(), g2.copy()
    p1[..., 0] *= 1 + 0.1 * rng.random(phi.shape)
    p2[..., 3] *= 1 + 0.1 * rng.random(phi.shape)
    out = run_shim(tmp_path, [header(3, info), [5], phi, p1, p2])
    H, W = phi.shape
    n = H * W * 4
    r1, r2 = p1.copy(), p2.copy()
    oracle.lib().or_extrapolate_vector(ogrid(info), O._p(phi), 5, O._p(r1), O._p(r2), 0)
    assert max(scaled(out[:n].reshape(H, W, 4), r1), scaled(out[n:].reshape(H, W, 4), r2)) <= 1e-12


@pytest.mark.parametrize("bcs", [(0, 0, 0, 0), (1, 1, 1, 1), (0, 0, 1, 1), (2, 2, 1, 1), (0, 0, 2, 2)])
def test_apply_bcs(tmp_path, oracle, bcs):
    info = G.SimInfo(40, 24, 3, 0.0, 0.0, 0.025, 0.04, *bcs)
    rng = np.random.default_rng(sum(bcs))
    grid = rng.random((30, 46, 4)) + 0.5
    out = run_shim(tmp_path, [header(4, info), grid]).reshape(grid.shape)
    ref = grid.copy()
    oracle.lib().or_apply_bcs(ogrid(info), O._p(ref))
    assert np.array_equal(out, ref)


def test_levelset_update(tmp_path, oracle):
    print("CASE shim level set", flush=True)
    info, eos, g1, g2, phi = developed_bubble(10)
    H, W = phi.shape
    yy, xx = np.mgrid[0:H, 0:W]
    vf = np.ascontiguousarray(np.stack([300 * np.sin(xx / 7.0), -200 * np.cos(yy / 5.0)], -1).astype(np.float64))
    dt = 2e-6
    out = run_shim(tmp_path, [header(5, info), [5, 0, 0, 0.0, dt], phi, vf]).reshape(H, W)
    ref = phi.copy()
    oracle.lib().or_ls_advect(ogrid(info), 5, O._p(ref), O._p(vf), dt)
    assert scaled(out[..., None], ref[..., None]) <= 1e-12
    assert np.abs(out - phi).max() > 0


def test_riemann_solvers(tmp_path, oracle):
    rng = np.random.default_rng(11)
    n = 64
    # no vacuum: the velocity jump stays well below the sum of 2c / (gamma - 1)
    W = np.stack([rng.uniform(0.5, 5.0, n), rng.uniform(-0.5, 0.5, n), rng.uniform(-1, 1, n),
                  rng.uniform(0.5, 10, n)], 1)
    Wr = np.stack([rng.uniform(0.5, 5.0, n), rng.uniform(-0.5, 0.5, n), rng.uniform(-1, 1, n),
                   rng.uniform(0.5, 10, n)], 1)

    def cons(w, g, pinf):
        return np.stack([w[:, 0], w[:, 0] * w[:, 1], w[:, 0] * w[:, 2],
                         (w[:, 3] + g * pinf) / (g - 1) + 0.5 * w[:, 0] * (w[:, 1] ** 2 + w[:, 2] ** 2)], 1)

    UL, UR = cons(W, 1.4, 0.0), cons(Wr, 1.4, 0.0)
    info = G.SimInfo(1, 1, 1, 0.0, 0.0, 1.0, 1.0, 0, 0, 0, 0)
    for px in (1, 0):
        got = run_shim(tmp_path, [header(6, info), [px, 1.4, 0.0, n], UL, UR]).reshape(n, 4)
        ref = np.array([O.pure_rs_flux(bool(px), 1.4, 0.0, UL[k], UR[k])[0] for k in range(n)])
        assert np.isfinite(ref).all()
        assert scaled(got, ref) <= 1e-12
    UR2 = cons(Wr, 4.4, 1.0)
    got = run_shim(tmp_path, [header(7, info), [1.4, 0.0, 4.4, 1.0, n], UL, UR2]).reshape(n, 4)
    ref = np.array([O.mixed_rs(1.4, 0.0, 4.4, 1.0, UL[k], UR2[k])[0] for k in range(n)])
    assert scaled(got, ref) <= 1e-12
