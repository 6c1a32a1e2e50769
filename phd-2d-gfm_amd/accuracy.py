"""Error measures between two two-fluid states of the same instance (mixed
precision against the FP64 GPU path, or either against a perturbed run).

Fields are compared on the interior cells in composite form: each cell takes
the conserved state of the fluid that is real there (phi <= 0: fluid 1, the
reference's side convention; sim_twofluid.cpp:128 uses phi < 0 for the CFL
fluid, the difference only matters on phi == 0 exactly).
  L1   per conserved variable: mean |a - b| / mean |b|
  Linf per conserved variable: max |a - b| / max |b|
  with |b| of the momentum magnitude for both momentum components (one
  component can vanish identically, e.g. v in a planar flow)
  contour: area of the symmetric difference of the regions phi <= 0,
           sum |H(phi_a) - H(phi_b)| dx dy, and that area over the area of
           the phi <= 0 region of b.
"""
import numpy as np

NAMES = ("rho", "rhou", "rhov", "E")


def interior(info, a):
    g = info.numGC
    return a[g:g + info.Ny, g:g + info.Nx]


def composite(info, state):
    g1, g2, phi = state
    side = (phi <= 0.0)[..., None]
    return interior(info, np.where(side, g1, g2)), interior(info, phi)


def compare(info, a, b):
    ua, pa = composite(info, a)
    ub, pb = composite(info, b)
    out = {}
    mom = np.hypot(ub[..., 1], ub[..., 2])
    for k, n in enumerate(NAMES):
        d = np.abs(ua[..., k] - ub[..., k])
        ref = mom if k in (1, 2) else np.abs(ub[..., k])
        out["L1_" + n] = float(d.mean() / max(ref.mean(), 1e-300))
        out["Linf_" + n] = float(d.max() / max(ref.max(), 1e-300))
    ha, hb = pa <= 0.0, pb <= 0.0
    area = float(np.count_nonzero(ha != hb)) * info.dx * info.dy
    out["contour_L1"] = area
    out["contour_L1_rel"] = area / max(float(np.count_nonzero(hb)) * info.dx * info.dy, 1e-300)
    return out


def aggregate(results, weights):
    """cell-weighted mean of per-instance relative errors (max for Linf)"""
    w = np.asarray(weights, dtype=np.float64)
    w = w / w.sum()
    out = {}
    for key in results[0]:
        vals = np.array([r[key] for r in results])
        out[key] = float(vals.max()) if key.startswith("Linf") else float((vals * w).sum())
    return out
