"""The reference's circular-explosion convergence test (sim_circularexplosiontest.cpp) on the GPU path.

1. Reference solution (:9-110): one fluid, gamma 1.4, on [0,2]^2 at ref_n^2 cells (4000^2 in the
   reference), MUSCL-Hancock with exact Riemann fluxes, transmissive BCs. The initial state is
   rho = 1, p = 1 inside radius 0.4 about (1, 1) and rho = 0.125, p = 0.1 outside, with the
   primitives mixed by 10x10 sub-cell sampling in the cut cells. The run goes to t = 0.25 with the
   CFL ramp.
2. For each GFM and resolution N (:285-384): the same explosion as a two-fluid problem with identical
   gases (both gamma 1.4). The level set is the signed distance to the circle, fluid 1 holds the inside
   state everywhere and fluid 2 the outside one. The scheme is MUSCL, the level set LS9 (HOUC-9).
   The reference uses 5 ghost layers; this library needs 6 for HOUC-9 plus the normal. By t = 0.25
   the blast stays well inside the domain, so the extra layer does not change the interior.
3. Error (:420-470): L1 = dx dy sum over interior cells of |rho_real - rho_ref(x_c)|, where rho_real
   is fluid 1's density where phi <= 0, else fluid 2's. rho_ref(x_c) is the reference density
   interpolated bilinearly between cell centres at the coarse cell centre.

The thesis plots these errors (CEtest_GFM.eps); the values decoded from the figure are in THESIS. The
CFL of those runs and the absent Levelset's reinitialisation are not recorded, so both are options
here. Prints one JSON line.
"""
import argparse
import json
import os
import time

import numpy as np

import gfm_hip as G

# density L1 errors at N = 50, 100, 200, 400 (CEtest_GFM.eps:650-681; VOFGFM: CEtest_VOFGFM.eps:661-664,
# run with the reference's VOF tracker, "PYVOF")
THESIS = {"OGFM": [0.0506, 0.0266, 0.0134, 0.00732],
          "riemannGFM": [0.0902, 0.0445, 0.0240, 0.0126],
          "MGFM": [0.0413, 0.0236, 0.0119, 0.00667],
          "VOFGFM": [0.0433, 0.0256, 0.0120, 0.00642]}
GFMS = {"OGFM": G.GFM_ORIGINAL, "realGFM": G.GFM_REAL, "riemannGFM": G.GFM_RIEMANN, "MGFM": G.GFM_MODIFIED,
        "VOFGFM": G.GFM_VOFGFM, "OGFMVOF": G.GFM_OGFMVOF}
INSIDE, OUTSIDE = (1.0, 0.0, 0.0, 1.0), (0.125, 0.0, 0.0, 0.1)
T_END = 0.25
TR = G.BC_TRANSMISSIVE


def run_to_end(ctx, block=50):
    """step until the instance stops at t = 0.25 (the context deactivates it)"""
    prev, n = -1, 0
    while n != prev:
        prev = n
        ctx.step_graph(block)
        ctx.sync()
        n = ctx.get_time(0)[1]
    return n


def reference(n, cfl):
    info = G.SimInfo(n, n, 2, 0.0, 0.0, 2.0 / n, 2.0 / n, TR, TR, TR, TR)
    eos = G.Eos(1.4, 0.0, 1.4, 0.0)
    ctx = G.Context(G.methods(sim_type=G.SIM_ONEFLUID, fs=G.FS_MUSCL, precision=G.FP64))
    ctx.add(info, eos, cfl, T_END, 1 << 30)
    ctx.commit()
    ctx.ic_device([(0, G.ic_desc(G.IC_EXPLOSION))])   # 10x10 sub-cell mixing, on the device
    steps = run_to_end(ctx)
    status = ctx.status(0)
    rho = ctx.download(0)[..., 0].copy()
    ctx.close()
    return info, rho, steps, status


def bilinear(info, f, x, y):
    """f between cell centres at (x, y), indices clamped to the array"""
    H, W = f.shape
    fx = (x - info.x0) / info.dx - 0.5 + info.numGC
    fy = (y - info.y0) / info.dy - 0.5 + info.numGC
    j0 = np.clip(np.floor(fx).astype(np.int64), 0, W - 2)
    i0 = np.clip(np.floor(fy).astype(np.int64), 0, H - 2)
    tx = np.clip(fx - j0, 0.0, 1.0)
    ty = np.clip(fy - i0, 0.0, 1.0)
    return ((1.0 - ty) * ((1.0 - tx) * f[i0, j0] + tx * f[i0, j0 + 1])
            + ty * ((1.0 - tx) * f[i0 + 1, j0] + tx * f[i0 + 1, j0 + 1]))


def gfm_error(gfm, n, cfl, ref_info, ref_rho, ls_order, reinit_every, reinit_iters, itm=G.ITM_LS):
    info = G.SimInfo(n, n, 6, 0.0, 0.0, 2.0 / n, 2.0 / n, TR, TR, TR, TR)
    eos = G.Eos(1.4, 0.0, 1.4, 0.0)
    ctx = G.Context(G.methods(sim_type=G.SIM_TWOFLUID, fs=G.FS_MUSCL, gfm=GFMS[gfm], ls_order=ls_order,
                              reinit_every=reinit_every, reinit_iters=reinit_iters, precision=G.FP64, itm=itm))
    ctx.add(info, eos, cfl, T_END, 1 << 30)
    ctx.commit()
    # fluid 1 the inside state everywhere, fluid 2 the outside one, phi the distance to the circle
    ctx.ic_device([(0, G.ic_desc(G.IC_BUBBLE, INSIDE, OUTSIDE, OUTSIDE, xc=1.0, yc=1.0, R=0.4, shock_x=1e300))])
    steps = run_to_end(ctx)
    status = ctx.status(0)
    g1, g2, phi = ctx.download(0)
    ctx.close()
    g = info.numGC
    rho = np.where(phi <= 0.0, g1[..., 0], g2[..., 0])[g:g + n, g:g + n]
    xc = info.x0 + (np.arange(n) + 0.5) * info.dx
    yc = info.y0 + (np.arange(n) + 0.5) * info.dy
    X, Y = np.meshgrid(xc, yc)
    l1 = float(np.abs(rho - bilinear(ref_info, ref_rho, X, Y)).sum() * info.dx * info.dy)
    return l1, steps, status


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--cfl", type=float, default=0.4)
    ap.add_argument("--ref-n", type=int, default=4000)
    ap.add_argument("--res", type=int, nargs="+", default=[50, 100, 200, 400])
    ap.add_argument("--gfms", nargs="+", default=["OGFM", "riemannGFM", "MGFM"], choices=sorted(GFMS))
    ap.add_argument("--ls-order", type=int, default=9)
    ap.add_argument("--reinit-every", type=int, default=0)
    ap.add_argument("--reinit-iters", type=int, default=10)
    ap.add_argument("--mgfm-extrapolate", action="store_true",
                    help="MGFM with the extrapolation of GFM_modified.cpp:202 applied to the new states "
                         "(GFM_MGFM_EXTRAPOLATE=1; the thesis's form, not the reference code's)")
    a = ap.parse_args()
    if a.mgfm_extrapolate:
        os.environ["GFM_MGFM_EXTRAPOLATE"] = "1"
    t0 = time.perf_counter()
    ref_info, ref_rho, ref_steps, ref_status = reference(a.ref_n, a.cfl)
    out = {"cfl": a.cfl, "ref_n": a.ref_n, "ref_steps": ref_steps, "ref_status": ref_status,
           "ref_seconds": time.perf_counter() - t0, "ls_order": a.ls_order,
           "reinit_every": a.reinit_every, "reinit_iters": a.reinit_iters, "mgfm_extrapolate": a.mgfm_extrapolate,
           "res": a.res, "errors": {}}
    for gfm in a.gfms:
        rows = []
        for k, n in enumerate(a.res):
            l1, steps, status = gfm_error(gfm, n, a.cfl, ref_info, ref_rho, a.ls_order, a.reinit_every,
                                          a.reinit_iters)
            th = THESIS.get(gfm, [None] * 4)
            rows.append({"N": n, "L1": l1, "steps": steps, "status": status,
                         "thesis": th[[50, 100, 200, 400].index(n)] if n in (50, 100, 200, 400) else None})
        out["errors"][gfm] = rows
    print(json.dumps(out))


if __name__ == "__main__":
    main()
