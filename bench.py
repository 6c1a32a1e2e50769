"""Benchmark of the GFM per-timestep hot path on one MI355X.

Workload (one "step" = one full two-fluid time step: CFL dt, ghost-fluid
fill, MUSCL-Hancock update of both fluids, level-set advection with
reinitialisation every 5 steps, boundary conditions):
  shocked bubble, 1024 x 288 cells, plain FP64, GFM_riemann, exact-RS
  MUSCL-Hancock, HOUC-5 level set, CFL 0.6, 2000 steps by default.
Inputs are resident in HBM before the timed region; one hipGraph per step.

Prints one JSON line.  `value` = cell-updates/s (interior cells x steps /
wall time).  `roofline` is for the dominant kernel (the fused MUSCL-Hancock
flux kernel): algorithmic bytes per launch = 2 fluids x Nx*Ny cells x 8
words (4 read, 4 written) x 8 bytes, over its average duration measured
with HIP events on the context's stream.  `cpu_baseline` times the FP64
CPU oracle (a single-threaded restatement of the reference loop) on a
bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "phd-2d-gfm_amd"))

import numpy as np  # noqa: E402

import gfm_hip as G  # noqa: E402
import problems as P  # noqa: E402

HBM_PEAK_GBS = 8000.0


def cpu_baseline(Nx, Ny, steps):
    """FP64 CPU oracle, one core, on the first `steps` steps of the workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    info, eos, (g1, g2, phi) = P.shocked_bubble(Nx, Ny)
    Gr = O.Grid(info.Nx, info.Ny, info.numGC, info.x0, info.y0, info.dx, info.dy,
                info.BC_L, info.BC_R, info.BC_B, info.BC_T)
    E = O.Eos(eos.gamma1, eos.pinf1, eos.gamma2, eos.pinf2)
    run = O.Run(G.GFM_RIEMANN, 0, 5, 5, 10, 0.6, 1e300, steps)
    s = np.zeros(1, dtype=np.int32)
    t = np.zeros(1)
    t0 = time.perf_counter()
    O.lib().or_twofluid_run(Gr, E, run, O._p(g1), O._p(g2), O._p(phi), O._p(s), O._p(t))
    el = time.perf_counter() - t0
    return {"value": Nx * Ny * int(s[0]) / el, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": "FP64 oracle, %dx%d shocked bubble GFM_riemann, first %d steps, %.1f s" % (Nx, Ny, int(s[0]), el)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--ny", type=int, default=288)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=50)
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py runs on one GPU only: the per-timestep path is measured on one MI355X "
              "(multi-GPU sharding is a design, not built)", file=sys.stderr)
        sys.exit(2)

    prec = G.FP64 if a.precision == "fp64" else G.MIXED
    info, eos, st = P.shocked_bubble(a.nx, a.ny)
    m = G.methods(sim_type=G.SIM_TWOFLUID, fs=G.FS_MUSCL, gfm=G.GFM_RIEMANN, ls_order=5,
                  reinit_every=5, reinit_iters=10, precision=prec)
    ctx = G.Context(m)
    k = ctx.add(info, eos, 0.6, 1e300, 1 << 30)
    ctx.commit()

    # warmup (graph capture, code load), then restart from the initial state
    ctx.upload(k, *st)
    ctx.step_graph(a.warmup)
    ctx.sync()
    ctx.upload(k, *st)
    ctx.set_time(k, 0.0, 0)
    ctx.sync()

    t0 = time.perf_counter()
    ctx.step_graph(a.steps)
    ctx.sync()
    el = time.perf_counter() - t0
    status = ctx.status(k)
    t_sim, nstep, _ = ctx.get_time(k)
    cells = a.nx * a.ny
    value = cells * a.steps / el

    # per-phase HIP-event timing on the same workload (fresh start)
    ctx.upload(k, *st)
    ctx.set_time(k, 0.0, 0)
    ms = ctx.profile_steps(a.profile_steps)
    fs_ms = ms[2] / a.profile_steps
    es = 8 if prec == G.FP64 else 4
    fs_bytes = 2 * cells * 8 * es
    achieved = fs_bytes / (fs_ms * 1e-3) / 1e9

    out = {
        "metric": "cell-updates/s on 1 MI355X (shocked bubble two-fluid step)",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": el * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == G.FP64 else "f32+f64",
        "data": "synthetic",
        "config": {"workload": "shocked bubble %dx%d GFM_riemann MUSCL exact-RS HOUC-5 %s, %d steps"
                               % (a.nx, a.ny, a.precision, a.steps),
                   "Nx": a.nx, "Ny": a.ny, "gfm": "riemannGFM", "fs": "MUSCL", "ls": "LS5",
                   "reinit_every": 5, "cfl": 0.6, "precision": a.precision},
        "roofline": {"bound": "hbm", "kernel": "k_fs (MUSCL-Hancock, both fluids)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None},
        "phase_ms_per_step": {n: v / a.profile_steps for n, v in zip(("dt", "ghost", "fs", "ls", "bc"), ms)},
        "sim_time": t_sim,
        "status": status,
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(a.nx, a.ny, a.cpu_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
