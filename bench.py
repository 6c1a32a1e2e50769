"""Benchmark of the GFM per-timestep hot path on one MI355X.

A "step" is one full two-fluid time step of every instance: CFL dt, ghost-
fluid fill (GFM_riemann), MUSCL-Hancock exact-RS update of both fluids,
HOUC-5 level-set advection with reinitialisation every 5 steps, boundary
conditions.  Inputs are resident in HBM before the timed region; each step
is one hipGraph launch.

Workloads (--workload):
  batch   (default) 64 seeded shocked bubbles, grids s*(256x72), s~U[1,8]
          (2.9e7 cells), per-instance Mach/radius/density/gamma, mixed
          precision (FP32 state, FP64 Newton, FP64 interfacial states),
          per-instance dt, 1500 steps;
  bubble  one 1024x288 shocked bubble, plain FP64, 2000 steps;
  large   one 4096x1152 shocked bubble, mixed precision, 2000 steps.
`value` = interior cells x steps / wall time, summed over instances.
`roofline` is for the dominant kernel, the fused MUSCL-Hancock flux kernel:
algorithmic bytes per launch = 2 fluids x cells x 8 words (4 read, 4
written) x word size, over its mean duration from HIP events on the
context's stream during a profiled pass of the same workload.
`fp64_plain` runs the same batch in FP64 for the mixed-vs-plain ratio.
`cpu_baseline` times the single-threaded FP64 CPU oracle on a bounded
sample (first steps of the smallest instances).
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


def build(cases, precision, steps):
    m = G.methods(sim_type=G.SIM_TWOFLUID, fs=G.FS_MUSCL, gfm=G.GFM_RIEMANN, ls_order=5,
                  reinit_every=5, reinit_iters=10, precision=precision)
    ctx = G.Context(m)
    states = []
    cells = 0
    for c in cases:
        info, eos, st = P.shocked_bubble(**c)
        ctx.add(info, eos, 0.6, 1e300, steps)
        states.append(st)
        cells += info.Nx * info.Ny
    ctx.commit()
    return ctx, states, cells


def load_all(ctx, states):
    for k, st in enumerate(states):
        ctx.upload(k, *st)
        ctx.set_time(k, 0.0, 0)


def timed_run(ctx, states, steps, warmup, eager=False):
    step = ctx.step if eager else ctx.step_graph
    load_all(ctx, states)
    step(warmup)
    ctx.sync()
    load_all(ctx, states)
    ctx.sync()
    t0 = time.perf_counter()
    step(steps)
    ctx.sync()
    return time.perf_counter() - t0


def cpu_baseline(cases, steps):
    """FP64 CPU oracle, one core, first `steps` steps of the given instances"""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    tot_cells, el = 0, 0.0
    for c in cases:
        info, eos, (g1, g2, phi) = P.shocked_bubble(**c)
        Gr = O.Grid(info.Nx, info.Ny, info.numGC, info.x0, info.y0, info.dx, info.dy,
                    info.BC_L, info.BC_R, info.BC_B, info.BC_T)
        E = O.Eos(eos.gamma1, eos.pinf1, eos.gamma2, eos.pinf2)
        run = O.Run(G.GFM_RIEMANN, 0, 5, 5, 10, 0.6, 1e300, steps)
        s = np.zeros(1, dtype=np.int32)
        t = np.zeros(1)
        t0 = time.perf_counter()
        O.lib().or_twofluid_run(Gr, E, run, O._p(g1), O._p(g2), O._p(phi), O._p(s), O._p(t))
        el += time.perf_counter() - t0
        tot_cells += info.Nx * info.Ny * int(s[0])
    desc = ", ".join("%dx%d" % (c["Nx"], c["Ny"]) for c in cases)
    return {"value": tot_cells / el, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": "FP64 oracle (single thread), GFM_riemann shocked bubble(s) %s, first %d steps, %.1f s"
                      % (desc, steps, el)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="batch", choices=["batch", "bubble", "large"])
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=None,
                    help="steps of the per-phase HIP-event pass (default: the timed step count)")
    ap.add_argument("--eager", action="store_true", help="launch kernels directly instead of one hipGraph per step")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one MI355X: the per-timestep path runs on one GPU here "
              "(instances would be bin-packed across GPUs; not built)", file=sys.stderr)
        sys.exit(2)

    if a.workload == "batch":
        cases = P.bubble_batch(a.instances)
        prec = G.MIXED
        steps = a.steps or 1500
        name = "%d seeded shocked bubbles s*(256x72), s~U[1,8], GFM_riemann, mixed FP32/FP64" % len(cases)
    elif a.workload == "bubble":
        cases = [dict(Nx=1024, Ny=288)]
        prec = G.FP64
        steps = a.steps or 2000
        name = "shocked bubble 1024x288, GFM_riemann, plain FP64"
    else:
        cases = [dict(Nx=4096, Ny=1152)]
        prec = G.MIXED
        steps = a.steps or 2000
        name = "shocked bubble 4096x1152, GFM_riemann, mixed FP32/FP64"

    ctx, states, cells = build(cases, prec, steps + a.warmup)
    el = timed_run(ctx, states, steps, a.warmup, a.eager)
    status = [ctx.status(k) for k in range(len(cases))]
    value = cells * steps / el

    # per-phase HIP-event timing: the same workload replayed from the initial
    # state with events around each phase (eager launches)
    nprof = a.profile_steps or steps
    load_all(ctx, states)
    ms = ctx.profile_steps(nprof)
    fs_ms = ms[2] / nprof
    es = 8 if prec == G.FP64 else 4
    fs_bytes = 2 * cells * 8 * es
    achieved = fs_bytes / (fs_ms * 1e-3) / 1e9
    ctx.close()

    out = {
        "metric": "cell-updates/s on 1 MI355X (two-fluid GFM step; mixed+batched vs FP64 plain)",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": a.warmup,
        "ms_per_step": el * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == G.FP64 else "f32+f64",
        "data": "synthetic",
        "config": {"workload": name, "instances": len(cases), "cells": cells, "steps": steps,
                   "gfm": "riemannGFM", "fs": "MUSCL", "ls": "LS5", "reinit_every": 5, "cfl": 0.6},
        "roofline": {"bound": "hbm", "kernel": "k_fs (MUSCL-Hancock, both fluids)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None},
        "phase_ms_per_step": {n: v / nprof for n, v in zip(("dt", "ghost", "fs", "ls", "bc"), ms)},
        "phase_steps": nprof,
        "status": sorted(set(status)),
    }
    if prec == G.MIXED and not a.no_fp64:
        nfp = steps
        ctx64, st64, cells64 = build(cases, G.FP64, nfp + a.warmup)
        el64 = timed_run(ctx64, st64, nfp, a.warmup, a.eager)
        out["fp64_plain"] = {"value": cells64 * nfp / el64, "steps": nfp,
                             "ratio_mixed_over_fp64": value / (cells64 * nfp / el64)}
        ctx64.close()
    if not a.no_cpu:
        small = sorted(cases, key=lambda c: c["Nx"] * c["Ny"])[:2]
        out["cpu_baseline"] = cpu_baseline(small, a.cpu_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
