"""The zero contour of the nondimensional shocked helium bubble (sim_twofluid.cpp:394-404, :774-859;
325x89, OGFM, MUSCL, CFL 0.4, t = 280) against the contours the reference holds from the thesis runs
(tests/golden/zerocont1.dat with LS1 after 3083 steps, zerocont5.dat with LS5 after 3090 steps; made by
plot_he_test1_contours.gnu:1-30 from the runs' -ls.dat files).

gnuplot's contour points lie where the grid lines through the cell centres cross phi = 0 (linear
interpolation between the two cell centres), and the same points are taken here from the -ls.dat file
the gfm2d driver writes at the end of the same run.  The level-set library of the thesis runs is
absent, so the shapes are compared with a tolerance rather than bit for bit:
  - mean distance from each fixture point to the nearest point of ours, and from each of ours to the
    nearest fixture point (the mean symmetric distance), in cells;
  - the symmetric Hausdorff distance (largest distance of either set to the other), in cells;
  - the extent of the contour in x and y.
Both runs use no periodic reinitialisation: that reproduces the thesis's step count with LS1 (3083) and
its bubble-mass error, while with LS5 it takes 3091 steps against 3090 and its mass error has the other
sign (+0.046 against -0.028), so the LS5 tolerance is the looser one.
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GFM2D = os.path.join(os.path.dirname(HERE), "phd-2d-gfm_amd", "gfm2d")


def fixture_points(name):
    pts = []
    for line in open(os.path.join(HERE, "golden", name)):
        if line.strip() and not line.startswith("#"):
            pts.append([float(v) for v in line.split()[:2]])
    return np.array(pts)


def read_ls(path):
    """-ls.dat: rows of 'x y phi' lines, a blank line after each grid row"""
    rows, cur = [], []
    for line in open(path):
        if line.strip():
            cur.append([float(v) for v in line.split()])
        elif cur:
            rows.append(cur)
            cur = []
    if cur:
        rows.append(cur)
    a = np.array(rows)
    return a[..., 0], a[..., 1], a[..., 2]


def zero_crossings(X, Y, P):
    pts = []
    for ax in (0, 1):
        p0 = P[:, :-1] if ax == 1 else P[:-1, :]
        p1 = P[:, 1:] if ax == 1 else P[1:, :]
        x0 = X[:, :-1] if ax == 1 else X[:-1, :]
        x1 = X[:, 1:] if ax == 1 else X[1:, :]
        y0 = Y[:, :-1] if ax == 1 else Y[:-1, :]
        y1 = Y[:, 1:] if ax == 1 else Y[1:, :]
        m = (p0 <= 0.0) != (p1 <= 0.0)
        w = p0[m] / (p0[m] - p1[m])
        pts.append(np.stack([x0[m] + w * (x1[m] - x0[m]), y0[m] + w * (y1[m] - y0[m])], -1))
    return np.concatenate(pts)


def distances(a, b):
    d = np.sqrt(((a[:, None, :] - b[None, :, :]) ** 2).sum(-1))
    return d.min(1), d.min(0)


def run_thesis(tmp_path, ls):
    f = tmp_path / "settingsfile.txt"
    f.write_text("Nx 325\nNy 89\nCFL 0.4\nsim_type twofluid\nFS MUSCL\ntest_case shocked_helium_bubble\n"
                 "ITM %s\nGFM OGFM\nreinit_every 0\nreinit_iters 0\noutputpath %s/\n" % (ls, tmp_path))
    out = subprocess.run([GFM2D, str(f)], capture_output=True, text=True, timeout=300)
    print(out.stdout[-500:], out.stderr[-500:])
    assert out.returncode == 0 and "status 0" in out.stdout
    # outputs at step 0, after the step passing t = T / output_freq = 140 (sim_twofluid.cpp:79-83) and at
    # the end; the thesis contours are of the end
    ls_files = {int(p.split("-")[-2]): p for p in os.listdir(tmp_path) if p.endswith("-ls.dat")}
    assert len(ls_files) == 3 and 0 in ls_files
    n = max(ls_files)
    return n, read_ls(os.path.join(tmp_path, ls_files[n]))


@pytest.mark.parametrize("ls,fixture,steps,mean_tol,haus_tol", [("LS1", "zerocont1.dat", 3083, 1.0, 5.0),
                                                                 ("LS5", "zerocont5.dat", 3091, 1.5, 6.0)])
def test_helium_bubble_zero_contour(tmp_path, ls, fixture, steps, mean_tol, haus_tol):
    print("CASE helium contour", ls, flush=True)
    n, (X, Y, P) = run_thesis(tmp_path, ls)
    assert n == steps
    ours = zero_crossings(X, Y, P)
    ours = ours[(ours[:, 0] >= 0) & (ours[:, 0] <= 325) & (ours[:, 1] >= 0) & (ours[:, 1] <= 89)]   # the plot's ranges
    ref = fixture_points(fixture)
    d_ref, d_ours = distances(ref, ours)
    mean_sym = 0.5 * (d_ref.mean() + d_ours.mean())
    haus = max(d_ref.max(), d_ours.max())
    ext = [ours[:, 0].min() - ref[:, 0].min(), ours[:, 0].max() - ref[:, 0].max(),
           ours[:, 1].min() - ref[:, 1].min(), ours[:, 1].max() - ref[:, 1].max()]
    print("%s: %d points (fixture %d), mean symmetric distance %.3f cells, Hausdorff %.3f cells, extent "
          "differences %s" % (ls, len(ours), len(ref), mean_sym, haus, np.round(ext, 2)), flush=True)
    assert mean_sym <= mean_tol
    assert haus <= haus_tol
    assert np.abs(ext).max() <= haus_tol
