"""Error measures used by bench.py (phd-2d-gfm_amd/accuracy.py), on
hand-made states whose errors are known in closed form."""
import numpy as np

import accuracy as A
import gfm_hip as G


def state(info, rho1, rho2, phi):
    H, W = info.shape
    g1 = np.zeros((H, W, 4)); g1[..., 0] = rho1; g1[..., 1] = 2.0; g1[..., 3] = 5.0
    g2 = np.zeros((H, W, 4)); g2[..., 0] = rho2; g2[..., 2] = 1.0; g2[..., 3] = 7.0
    return g1, g2, phi


def test_identical_states_have_zero_error():
    info = G.SimInfo(12, 10, 2, 0.0, 0.0, 0.5, 0.25, 0, 0, 1, 1)
    H, W = info.shape
    phi = np.fromfunction(lambda i, j: j - 8.0, (H, W))
    s = state(info, 1.0, 3.0, phi)
    e = A.compare(info, s, s)
    assert all(v == 0.0 for v in e.values())


def test_composite_and_contour():
    info = G.SimInfo(12, 10, 2, 0.0, 0.0, 0.5, 0.25, 0, 0, 1, 1)
    H, W = info.shape
    jj = np.fromfunction(lambda i, j: j, (H, W))
    ref = state(info, 1.0, 3.0, jj - 8.0)          # fluid 1 real for interior columns j <= 8
    got = state(info, 1.1, 3.0, jj - 9.0)          # fluid 1 denser, one more column on its side
    e = A.compare(info, got, ref)
    # composite densities: ref has 1.0 on interior columns 2..8 (7 of 12), 3.0 on 9..13;
    # got has 1.1 on columns 2..9, 3.0 on 10..13
    rr = np.where(np.arange(2, 14) <= 8, 1.0, 3.0)
    rg = np.where(np.arange(2, 14) <= 9, 1.1, 3.0)
    assert np.isclose(e["L1_rho"], np.abs(rg - rr).mean() / rr.mean())
    assert np.isclose(e["Linf_rho"], np.abs(rg - rr).max() / rr.max())
    # one column of 10 cells changed side: area 10 * dx * dy
    assert np.isclose(e["contour_L1"], 10 * 0.5 * 0.25)
    assert np.isclose(e["contour_L1_rel"], 10 / 70.0)


def test_aggregate_weights_and_max():
    r = [{"L1_rho": 1.0, "Linf_rho": 1.0}, {"L1_rho": 3.0, "Linf_rho": 0.5}]
    out = A.aggregate(r, [1, 3])
    assert np.isclose(out["L1_rho"], 2.5) and out["Linf_rho"] == 1.0
