"""The settings-file driver gfm2d (main.cpp / sim_base flow of the reference)
run end to end on the GPU, its output files (prims, masschange, vfield,
schlieren, GDAerr; sim_twofluid.cpp:163-331) checked against the oracle.
The output format prints 6 significant digits (as the reference's ofstream
default does), so the bar is 2e-5 relative per printed field."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GFM2D = os.path.join(os.path.dirname(HERE), "phd-2d-gfm_amd", "gfm2d")


def settings(tmp_path, **kv):
    p = tmp_path / "settingsfile.txt"
    kv.setdefault("outputpath", str(tmp_path) + "/")
    p.write_text("".join("%s %s\n" % (k, v) for k, v in kv.items()))
    return str(p)


def read_dat(path):
    rows = [l.split() for l in open(path) if l.strip()]
    return np.array(rows, dtype=np.float64)


def output_steps(tmp_path, suffix="-prims.dat"):
    """{step: file name} of the per-output files <base>-<step><suffix>"""
    return {int(p[:-len(suffix)].split("-")[-1]): p for p in os.listdir(tmp_path) if p.endswith(suffix)}


def test_gfm2d_onefluid_circular_explosion(tmp_path, oracle):
    f = settings(tmp_path, Nx=101, Ny=101, CFL=0.7, sim_type="onefluid", FS="MUSCL", test_case="circularexplosion")
    out = subprocess.run([GFM2D, f], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    files = output_steps(tmp_path)
    assert len(files) == 2 and 0 in files         # step 0 and the end (sim_onefluid.cpp:34, :60)
    nsteps = max(files)
    got = read_dat(os.path.join(tmp_path, files[nsteps]))
    G = oracle.Grid(101, 101, 2, 0.0, 0.0, 2 / 101, 2 / 101, 0, 0, 0, 0)
    H, W = 105, 105
    grid = np.zeros((H, W, 4))
    oracle.lib().or_ic_onefluid(G, 1.4, 0.0, 100, oracle._p(grid))
    first = read_dat(os.path.join(tmp_path, files[0]))
    assert np.allclose(first[:, 2], grid[..., 0].reshape(-1), rtol=2e-5, atol=0)
    s = np.zeros(1, dtype=np.int32)
    t = np.zeros(1)
    assert oracle.lib().or_onefluid_run(G, 1.4, 0.0, 0, 0.7, 0.25, 100000, oracle._p(grid), oracle._p(s),
                                        oracle._p(t)) == 0
    assert nsteps == s[0]
    rho = grid[..., 0].reshape(-1)
    assert np.allclose(got[:, 2], rho, rtol=2e-5, atol=0)
    p = (0.4 * (grid[..., 3] - 0.5 * (grid[..., 1] ** 2 + grid[..., 2] ** 2) / grid[..., 0])).reshape(-1)
    assert np.allclose(got[:, 6], p, rtol=2e-5, atol=1e-9)


def test_gfm2d_helium_bubble(tmp_path, oracle):
    """the reference's nondimensional shocked_helium_bubble case at 130x36, 40 steps"""
    f = settings(tmp_path, Nx=130, Ny=36, CFL=0.4, sim_type="twofluid", FS="MUSCL",
                 test_case="shocked_helium_bubble", ITM="LS5", GFM="riemannGFM", steps=40)
    out = subprocess.run([GFM2D, f], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0 and "status 0" in out.stdout
    files = output_steps(tmp_path)
    assert sorted(files) == [0, 40]               # T / output_freq = 140 is not reached in 40 steps
    prims = files[40]
    got = read_dat(os.path.join(tmp_path, prims))
    G = oracle.Grid(130, 36, 8, 0.0, 0.0, 325.0 / 130, 89.0 / 36, 0, 0, 1, 1)
    E = oracle.Eos(1.667, 0.0, 1.4, 0.0)
    H, W = 52, 146
    g1, g2, phi = np.zeros((H, W, 4)), np.zeros((H, W, 4)), np.zeros((H, W))
    oracle.lib().or_ic_bubble(G, E, 175.0, 44.5, 25.0, oracle._p(oracle.arr(0.138, 0, 0, 1)),
                              oracle._p(oracle.arr(1, 0, 0, 1)), oracle._p(oracle.arr(1.3764, -0.394, 0, 1.5698)),
                              225.0, oracle._p(g1), oracle._p(g2), oracle._p(phi))
    run = oracle.Run(2, 0, 5, 5, 10, 0.4, 280.0, 40)
    # or_twofluid_run's loop, step by step, to keep the last step's velocity field
    s = np.zeros(1, dtype=np.int32)
    t = np.zeros(1)
    vf, s1, s2 = np.zeros((H, W, 2)), np.zeros((H, W, 4)), np.zeros((H, W, 4))
    while t[0] < run.T and s[0] < run.max_steps:
        assert oracle.lib().or_twofluid_step(G, E, run, int(s[0]), oracle._p(t), oracle._p(g1), oracle._p(g2),
                                             oracle._p(phi), oracle._p(vf), oracle._p(s1), oracle._p(s2),
                                             None) == 0
        s[0] += 1
    rho = np.where(phi < 0, g1[..., 0], g2[..., 0])[2:H - 2, 2:W - 2].reshape(-1)
    assert np.allclose(got[:, 2], rho, rtol=2e-5, atol=0)
    # -vfield.dat: the velocity field of the last ghost fill
    vfield = read_dat(os.path.join(tmp_path, prims.replace("-prims.dat", "-vfield.dat")))
    ref = vf[2:H - 2, 2:W - 2].reshape(-1, 2)
    assert vfield.shape == (ref.shape[0], 4)
    assert np.allclose(vfield[:, :2], got[:, :2], rtol=0, atol=0)
    assert np.allclose(vfield[:, 2:], ref, rtol=2e-5, atol=1e-7 * np.abs(ref).max())
    # -schlieren.dat (sim_twofluid.cpp:226-271) from the oracle's final state, restated here
    g = 8
    rr = np.where(phi <= 0, g1[..., 0], g2[..., 0])
    gx = (rr[g:g + 36, g + 1:g + 131] - rr[g:g + 36, g - 1:g + 129]) / (2 * G.dx)
    gy = (rr[g + 1:g + 37, g:g + 130] - rr[g - 1:g + 35, g:g + 130]) / (2 * G.dy)
    gm = np.hypot(gx, gy).reshape(-1)
    sch = read_dat(os.path.join(tmp_path, prims.replace("-prims.dat", "-schlieren.dat")))
    assert sch.shape == (130 * 36, 3)
    assert np.allclose(sch[:, 2], gm / gm.max(), rtol=2e-5, atol=2e-6)
    # -GDAerr.dat: Nx and the L1 distance of fluid 1's density to the advection test's Gaussian
    xs = 0.0 + (np.arange(W) - g + 0.5) * G.dx
    ys = 0.0 + (np.arange(H) - g + 0.5) * G.dy
    X, Y = np.meshgrid(xs, ys)
    exact = 1.0 + 1000.0 * np.exp(-((X - 0.5) ** 2 + (Y - 0.5) ** 2) / (2 * 0.1 * 0.1))
    l1 = (np.abs(g1[..., 0] - exact)[g:g + 36, g:g + 130]).sum() * G.dx * G.dy
    gda = read_dat(str(tmp_path / [p for p in os.listdir(tmp_path) if p.endswith("-GDAerr.dat")][0]))
    assert gda.shape == (1, 2) and gda[0, 0] == 130
    assert abs(gda[0, 1] - l1) <= 2e-5 * l1
    mass = read_dat(os.path.join(tmp_path, [p for p in os.listdir(tmp_path) if p.endswith("masschange.dat")][0]))
    assert mass.shape == (41, 3) and mass[0, 1] == 0.0 and mass[0, 2] == 0.0
    m1 = oracle.lib().or_fluid_mass(G, -1.0, oracle._p(g1), oracle._p(phi))
    assert abs(mass[-1, 0] - t[0]) <= 1e-5 * t[0]


@pytest.mark.parametrize("gfm,variant,itm", [("VOFGFM", 4, "LS5"), ("OGFMVOF", 5, "LS5"), ("VOFGFM", 4, "CLSVOF")])
def test_gfm2d_vof_helium_bubble(tmp_path, oracle, gfm, variant, itm):
    """the settings names of the VOF-based GFMs (sim_twofluid.cpp:538-545) and of the CLSVOF tracker
    (:571-574) on the nondimensional shocked_helium_bubble at 130x36, 40 steps, against the oracle's
    GFM_VOF / OGFM_VOF on the level set or the CLSVOF tracker"""
    f = settings(tmp_path, Nx=130, Ny=36, CFL=0.4, sim_type="twofluid", FS="MUSCL",
                 test_case="shocked_helium_bubble", ITM=itm, GFM=gfm, steps=40)
    out = subprocess.run([GFM2D, f], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0 and "status 0" in out.stdout
    files = output_steps(tmp_path)
    got = read_dat(os.path.join(tmp_path, files[40]))
    G = oracle.Grid(130, 36, 8, 0.0, 0.0, 325.0 / 130, 89.0 / 36, 0, 0, 1, 1)
    E = oracle.Eos(1.667, 0.0, 1.4, 0.0)
    H, W = 52, 146
    g1, g2, phi = np.zeros((H, W, 4)), np.zeros((H, W, 4)), np.zeros((H, W))
    oracle.lib().or_ic_bubble(G, E, 175.0, 44.5, 25.0, oracle._p(oracle.arr(0.138, 0, 0, 1)),
                              oracle._p(oracle.arr(1, 0, 0, 1)), oracle._p(oracle.arr(1.3764, -0.394, 0, 1.5698)),
                              225.0, oracle._p(g1), oracle._p(g2), oracle._p(phi))
    run = oracle.Run(variant, 0, 5, 5, 10, 0.4, 280.0, 40, 0, 0,
                     oracle.ITM_CLSVOF if itm == "CLSVOF" else oracle.ITM_LS)
    s = np.zeros(1, dtype=np.int32)
    t = np.zeros(1)
    assert oracle.lib().or_twofluid_run(G, E, run, oracle._p(g1), oracle._p(g2), oracle._p(phi), oracle._p(s),
                                        oracle._p(t)) == 0
    rho = np.where(phi < 0, g1[..., 0], g2[..., 0])[2:H - 2, 2:W - 2].reshape(-1)
    assert np.allclose(got[:, 2], rho, rtol=2e-5, atol=0)


def test_gfm2d_helium_bubble_thesis_ls1(tmp_path):
    """The thesis result for the nondimensional shocked helium bubble at 325x89 (OGFM + LS1, MUSCL,
    CFL 0.4, t = 280): 3083 steps and a relative bubble-mass error of -0.208 at the end (decoded from
    the reference's figures He_test_1_masserror.eps and plot_he_test1_contours.gnu). Without periodic
    reinitialisation the run reproduces the step count exactly and the mass error within 0.01; the
    absent Levelset's reinitialisation schedule is unknown (every 5 steps gives 3087 steps, -0.099)."""
    f = settings(tmp_path, Nx=325, Ny=89, CFL=0.4, sim_type="twofluid", FS="MUSCL",
                 test_case="shocked_helium_bubble", ITM="LS1", GFM="OGFM", reinit_every=0, reinit_iters=0)
    out = subprocess.run([GFM2D, f], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0 and "status 0" in out.stdout
    steps = int(out.stdout.split("Simulation complete: ")[1].split(" steps")[0])
    mass = read_dat(os.path.join(tmp_path, [p for p in os.listdir(tmp_path) if p.endswith("masschange.dat")][0]))
    assert steps == 3083
    assert mass[-1, 0] == pytest.approx(280.0)
    assert abs(mass[-1, 1] - (-0.208)) <= 0.01


def test_gfm2d_rmi_pert_outputs(tmp_path, oracle):
    """RMI_pert (sim_twofluid.cpp:430-442, :1348-1453) at 128x32, 120 steps: the periodic output after
    every step with t - lastoutputtime > T / output_freq = 0.4 (:79-83) plus step 0 and the end, the -z
    files (fluid 1's area, 1.1, at step 0), and -pertamp / -pertampderiv (:1831-1887) against the
    oracle's run with its tracers moved in each step's velocity field."""
    n = 120
    f = settings(tmp_path, Nx=128, Ny=32, CFL=0.5, sim_type="twofluid", FS="MUSCL", test_case="RMI_pert",
                 ITM="LS5", GFM="riemannGFM", steps=n)
    out = subprocess.run([GFM2D, f], capture_output=True, text=True, timeout=120)
    print(out.stdout[-2000:], out.stderr)
    assert out.returncode == 0 and "status 0" in out.stdout
    mass = read_dat(os.path.join(tmp_path, [p for p in os.listdir(tmp_path) if p.endswith("masschange.dat")][0]))
    assert mass.shape == (n + 1, 3)
    # the output steps by the reference's rule, from the printed times (6 digits: the rule is applied
    # to the oracle's times below as well)
    G_ = oracle.Grid(128, 32, 8, 0.0, 0.0, 4.0 / 128, 1.0 / 32, 0, 0, 1, 1)
    E = oracle.Eos(1.4, 0.0, 1.093, 0.0)
    H, W = 48, 144
    g1, g2, phi = np.zeros((H, W, 4)), np.zeros((H, W, 4)), np.zeros((H, W))
    oracle.lib().or_ic_rmi(G_, E, 2.9, 0.02, 1, oracle._p(oracle.arr(1, 0, 0, 1)),
                           oracle._p(oracle.arr(1.411, -0.39, 0, 1.628)), oracle._p(oracle.arr(5.04, 0, 0, 1)),
                           3.2, oracle._p(g1), oracle._p(g2), oracle._p(phi))
    oracle.lib().or_ls_reinit(G_, oracle._p(phi), 40)
    run = oracle.Run(2, 0, 5, 5, 10, 0.5, 4.0, n)
    t = np.zeros(1)
    dt = np.zeros(1)
    vf, s1, s2 = np.zeros((H, W, 2)), np.zeros((H, W, 4)), np.zeros((H, W, 4))
    pos = np.array([[2.88, 1.0], [2.92, 0.5]])
    p0 = pos.copy()
    rows, times = [], [0.0]
    for s in range(n):
        assert oracle.lib().or_twofluid_step(G_, E, run, s, oracle._p(t), oracle._p(g1), oracle._p(g2),
                                             oracle._p(phi), oracle._p(vf), oracle._p(s1), oracle._p(s2),
                                             oracle._p(dt)) == 0
        e = np.zeros(9)
        oracle.lib().or_tracers_step(G_, oracle._p(vf), 2, oracle._p(pos), oracle._p(p0), 1, t[0], dt[0],
                                     oracle._p(e))
        rows.append(e)
        times.append(t[0])
    expect, last = {0, n}, 0.0
    for s in range(1, n + 1):
        if times[s] - last > 0.4:
            expect.add(s)
            last = times[s]
    assert len(expect) >= 4
    for suffix in ("-prims.dat", "-ls.dat", "-vfield.dat", "-z.dat", "-schlieren.dat"):
        assert set(output_steps(tmp_path, suffix)) == expect, suffix
    z0 = read_dat(os.path.join(tmp_path, output_steps(tmp_path, "-z.dat")[0]))
    inside = (z0[:, 0] > 0) & (z0[:, 0] < 4) & (z0[:, 1] > 0) & (z0[:, 1] < 1)
    assert abs(z0[inside, 2].sum() * G_.dx * G_.dy - 1.1) <= 2e-3
    amp = read_dat(os.path.join(tmp_path, [p for p in os.listdir(tmp_path) if p.endswith("-pertamp.dat")][0]))
    der = read_dat(os.path.join(tmp_path, [p for p in os.listdir(tmp_path) if p.endswith("-pertampderiv.dat")][0]))
    rows = np.array(rows)
    assert amp.shape == (n, 4) and der.shape == (n, 4)
    assert np.allclose(amp[0], [rows[0, 0], 0.02, 2.92, 2.88], rtol=2e-5, atol=0)
    ref_amp = np.stack([rows[:, 0], 0.5 * (rows[:, 5] - rows[:, 1]), rows[:, 5], rows[:, 1]], axis=1)
    assert np.allclose(amp, ref_amp, rtol=2e-5, atol=1e-7)
    ref_der = np.stack([rows[:, 0], 0.5 * (rows[:, 7] - rows[:, 3]), rows[:, 7], rows[:, 3]], axis=1)
    umax = np.abs(rows[:, [3, 7]]).max()
    assert umax > 0.1                       # the shock has reached both points
    assert np.allclose(der[:, 0], ref_der[:, 0], rtol=2e-5, atol=0)
    assert np.abs(der[:, 1:] - ref_der[:, 1:]).max() <= 2e-5 * umax
