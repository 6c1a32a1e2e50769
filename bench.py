"""Benchmark of the GFM per-timestep hot path on one MI355X.

A "step" is one full two-fluid time step of every instance: CFL dt, ghost-
fluid fill, MUSCL-Hancock exact-RS update of both fluids, HOUC-5 level-set
advection with reinitialisation every 5 steps, boundary conditions.  Inputs
are resident in HBM before the timed region; a single instance runs one
hipGraph per step, a batch runs its instance groups (two streams in mixed
precision, four in FP64) from one graph per 100-step block (--eager: direct
launches).

Workloads (--workload):
  batch   (default) 64 seeded shocked bubbles, grids s*(256x72), s~U[1,8]
          (2.55e7 interior cells), per-instance Mach/radius/density/gamma,
          GFM_riemann, mixed precision (FP32 state and sweeps, FP64 Newton,
          FP64 interfacial states, FP64 sums), per-instance dt, 1500 steps;
  bubble  one 1024x288 shocked bubble, plain FP64, GFM_riemann, 2000 steps;
  large   one 4096x1152 shocked bubble (bubble gamma and density drawn from
          seed 20261203), mixed precision, GFM_riemann, 2000 steps; with
          --accuracy also GFM_original and GFM_real, each against FP64;
  rmi     256 seeded single-mode Richtmyer-Meshkov instances, grids
          s*(512x128), s~U[1,4], periodic in y, GFM_modified, mixed
          precision; 300 steps unless --steps says otherwise.
`value` = interior cells x steps / wall time, summed over instances.
`roofline` is for the dominant kernel, the fused MUSCL-Hancock flux kernel:
algorithmic bytes per launch = 2 fluids x cells x 8 words (4 read, 4
written) x word size, over its mean duration from HIP events on the
context's stream during a replay of the same workload (same steps, from the
same initial state, with events around each phase).  `step_roofline` is
the whole step against the compulsory state traffic 2 (4 F + L) s bytes
per cell-update (F = 2 fluids, L = 1 level set, s = word size).
For mixed-precision workloads the FP64 path runs the same job (`fp64_plain`)
and the final states are compared (`accuracy`: relative L1 / Linf per
conserved variable of the composite real-fluid state, interface-contour
L1); --perturb adds the same comparison between FP64 and FP64 from an
initial state perturbed by 1e-7 relative, which measures how much the flow
itself amplifies a last-digit change.
`cpu_baseline` times the single-threaded FP64 CPU oracle on a bounded
sample (first steps of the smallest instances).
Before timing, each precision's schedule (instance groups, blocked or
one-sweep sweeps) is autotuned on short runs of the same job (autotune.py;
`schedule` holds the trials; --no-autotune keeps the library default).
--host-io K adds `host_io_inclusive`: K steps with every instance's grids
moved host -> device -> host around each step (the per-phase drop-in
boundary, PCIe included), reported beside `value`, never as it.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "phd-2d-gfm_amd"))

import numpy as np  # noqa: E402

import accuracy as A  # noqa: E402
import autotune as T  # noqa: E402
import gfm_hip as G  # noqa: E402
import problems as P  # noqa: E402

HBM_PEAK_GBS = 8000.0
SEED = 20261203
# HBM bytes per interior cell (both fluids) of one flux-kernel launch, from
# rocprof FETCH_SIZE / WRITE_SIZE scaled by a device copy of known size
# (profiles/k_fs_batch64_mixed_traffic.md); measured for the default workload
K_FS_TRAFFIC_PER_CELL = {"batch": 66.1}
GFM_LABEL = {G.GFM_ORIGINAL: "OGFM", G.GFM_REAL: "realGFM", G.GFM_RIEMANN: "riemannGFM", G.GFM_MODIFIED: "MGFM"}


def make_cases(workload, n_inst):
    """(problem builder, per-instance kwargs, variant, precision, steps, name, init reinit iterations)"""
    if workload == "batch":
        cases = P.bubble_batch(n_inst)
        return (P.shocked_bubble, cases, G.GFM_RIEMANN, G.MIXED, 1500,
                "%d seeded shocked bubbles s*(256x72), s~U[1,8], GFM_riemann, mixed FP32/FP64" % len(cases), 0)
    if workload == "bubble":
        return (P.shocked_bubble, [dict(Nx=1024, Ny=288)], G.GFM_RIEMANN, G.FP64, 2000,
                "shocked bubble 1024x288, GFM_riemann, plain FP64", 0)
    if workload == "large":
        rng = np.random.default_rng(SEED)
        c = dict(Nx=4096, Ny=1152, bubble_gamma=float(rng.uniform(1.09, 1.67)),
                 bubble_rho=float(rng.uniform(0.18, 5.0)))
        return (P.shocked_bubble, [c], G.GFM_RIEMANN, G.MIXED, 2000,
                "shocked bubble 4096x1152 (gamma %.4f, rho %.4f, seed %d), GFM_riemann, mixed FP32/FP64"
                % (c["bubble_gamma"], c["bubble_rho"], SEED), 0)
    cases = P.rmi_batch(n_inst)
    return (P.rmi, cases, G.GFM_MODIFIED, G.MIXED, 300,
            "%d seeded single-mode RMI s*(512x128), s~U[1,4], periodic y, GFM_modified, mixed FP32/FP64"
            % len(cases), 40)


def make_setup(builder, cases, perturb=0.0):
    """per instance (sim info, EOS) and the initial host grids"""
    setup, states = [], []
    rng = np.random.default_rng(7)
    for c in cases:
        info, eos, st = builder(**c)
        if perturb:
            st = tuple(np.ascontiguousarray(a * (1.0 + perturb * rng.uniform(-1.0, 1.0, a.shape)))
                       if k < 2 else a for k, a in enumerate(st))
        setup.append((info, eos))
        states.append(st)
    return setup, states


def new_ctx(setup, variant, precision, steps):
    m = G.methods(sim_type=G.SIM_TWOFLUID, fs=G.FS_MUSCL, gfm=variant, ls_order=5,
                  reinit_every=5, reinit_iters=10, precision=precision)
    ctx = G.Context(m)
    for info, eos in setup:
        ctx.add(info, eos, 0.6, 1e300, steps)
    ctx.commit()
    return ctx


def build(builder, cases, variant, precision, steps, perturb=0.0):
    setup, states = make_setup(builder, cases, perturb)
    return new_ctx(setup, variant, precision, steps), states, [s[0] for s in setup]


def tuned_ctx(a, setup, states, variant, precision, steps, init_reinit):
    """a context on the autotuned schedule (autotune.py), unless --no-autotune, --eager, or the
    schedule is set in the environment; returns (ctx, schedule record)"""
    if a.no_autotune or a.eager or any(k in os.environ for k in T.KEYS):
        rec = {"tuned": False, "env": {k: os.environ[k] for k in T.KEYS if k in os.environ}}
        return new_ctx(setup, variant, precision, steps), rec
    ts = min(a.tune_steps, steps)
    choice = T.tune(lambda: new_ctx(setup, variant, precision, ts + a.warmup),
                    lambda c: load_all(c, states, init_reinit), len(setup), steps=ts, warmup=a.warmup)
    with T.applied(choice):
        ctx = new_ctx(setup, variant, precision, steps)
    return ctx, dict(choice, tuned=True, tune_steps=ts)


def load_all(ctx, states, init_reinit):
    for k, st in enumerate(states):
        ctx.upload(k, *st)
        ctx.set_time(k, 0.0, 0)
    if init_reinit:
        ctx.ls_reinit(init_reinit)   # ls->initialisation() (sim_twofluid.cpp:47)


def timed_run(ctx, states, steps, warmup, init_reinit, eager=False):
    step = ctx.step if eager else ctx.step_graph
    load_all(ctx, states, init_reinit)
    step(warmup)
    ctx.sync()
    load_all(ctx, states, init_reinit)
    ctx.sync()
    t0 = time.perf_counter()
    step(steps)
    ctx.sync()
    return time.perf_counter() - t0


def cpu_baseline(builder, cases, variant, steps):
    """FP64 CPU oracle, one core, first `steps` steps of the given instances"""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    tot_cells, el = 0, 0.0
    for c in cases:
        info, eos, (g1, g2, phi) = builder(**c)
        Gr = O.Grid(info.Nx, info.Ny, info.numGC, info.x0, info.y0, info.dx, info.dy,
                    info.BC_L, info.BC_R, info.BC_B, info.BC_T)
        E = O.Eos(eos.gamma1, eos.pinf1, eos.gamma2, eos.pinf2)
        run = O.Run(variant, 0, 5, 5, 10, 0.6, 1e300, steps)
        s = np.zeros(1, dtype=np.int32)
        t = np.zeros(1)
        t0 = time.perf_counter()
        O.lib().or_twofluid_run(Gr, E, run, O._p(g1), O._p(g2), O._p(phi), O._p(s), O._p(t))
        el += time.perf_counter() - t0
        tot_cells += info.Nx * info.Ny * int(s[0])
    desc = ", ".join("%dx%d" % (c["Nx"], c["Ny"]) for c in cases)
    return {"value": tot_cells / el, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": "FP64 oracle (single thread), first %d steps of %s, %.1f s" % (steps, desc, el)}


def states_of(ctx, n):
    return [ctx.download(k) for k in range(n)]


def host_io_run(ctx, states, steps, init_reinit):
    """the rate through the per-phase drop-in boundary: every step uploads
    each instance's host grids (FP64 cells, both fluids and phi), runs one
    step (eager launches) and downloads all grids again"""
    load_all(ctx, states, init_reinit)
    ctx.sync()
    cur = list(states)
    t0 = time.perf_counter()
    for _ in range(steps):
        for k, st in enumerate(cur):
            ctx.upload(k, *st)
        ctx.step(1)
        cur = states_of(ctx, len(cur))
    ctx.sync()
    return time.perf_counter() - t0


def accuracy(infos, got, ref):
    res = [A.compare(info, a, b) for info, a, b in zip(infos, got, ref)]
    return A.aggregate(res, [i.Nx * i.Ny for i in infos])


def run_job(builder, cases, variant, precision, steps, init_reinit, perturb=0.0):
    ctx, states, _ = build(builder, cases, variant, precision, steps, perturb)
    load_all(ctx, states, init_reinit)
    ctx.step_graph(steps)
    ctx.sync()
    out = states_of(ctx, len(cases))
    ctx.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="batch", choices=["batch", "bubble", "large", "rmi"])
    ap.add_argument("--instances", type=int, default=None)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison run of mixed workloads")
    ap.add_argument("--accuracy", action="store_true",
                    help="large: also run GFM_original and GFM_real, mixed against FP64")
    ap.add_argument("--perturb", action="store_true", help="also compare FP64 against FP64 from a perturbed start")
    ap.add_argument("--profile-steps", type=int, default=None,
                    help="steps of the per-phase HIP-event pass (default: the timed step count)")
    ap.add_argument("--eager", action="store_true", help="launch kernels directly instead of one hipGraph per step")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the library's default schedule instead of autotuning it (autotune.py)")
    ap.add_argument("--tune-steps", type=int, default=100, help="steps per autotuning trial")
    ap.add_argument("--host-io", type=int, default=0, metavar="K",
                    help="also time K steps that move every instance's grids host -> device -> host around each "
                         "step (the per-phase drop-in's rate, PCIe included; reported beside value, never as it)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run device-to-device copies of known size (the workload's conserved grids, "
                         "through gfm_set_ghost_states) for scaling rocprof FETCH_SIZE / WRITE_SIZE; prints "
                         "the bytes copied")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one MI355X: the per-timestep path runs on one GPU here "
              "(instances would be bin-packed across GPUs; not built)", file=sys.stderr)
        sys.exit(2)

    n_inst = a.instances or (64 if a.workload == "batch" else 256)
    builder, cases, variant, prec, dsteps, name, init_reinit = make_cases(a.workload, n_inst)
    steps = a.steps or dsteps

    setup, states = make_setup(builder, cases)
    infos = [s[0] for s in setup]
    cells = sum(i.Nx * i.Ny for i in infos)
    if a.copy_calibration:
        ctx = new_ctx(setup, variant, prec, steps + a.warmup)
        # the phase API ends a ghost fill with Q[.][1] -> Q[.][0], one
        # hipMemcpyAsync per instance and fluid of 4 padded planes
        es = 8 if prec == G.FP64 else 4
        nbytes = 0
        for i in infos:
            H, W = i.shape
            pitch = (W + 63) // 64 * 64
            plane = (pitch * H + 63) // 64 * 64
            nbytes += 2 * 4 * plane * es
        load_all(ctx, states, init_reinit)
        ctx.set_ghost_states()
        ctx.sync()
        print(json.dumps({"copy_calibration_bytes": nbytes, "copies": 2 * len(infos)}))
        return
    ctx, schedule = tuned_ctx(a, setup, states, variant, prec, steps + a.warmup, init_reinit)
    el = timed_run(ctx, states, steps, a.warmup, init_reinit, a.eager)
    status = [ctx.status(k) for k in range(len(cases))]
    value = cells * steps / el

    # per-phase HIP-event timing: the same workload replayed from the initial
    # state with events around each phase (eager launches)
    nprof = a.profile_steps or steps
    load_all(ctx, states, init_reinit)
    ms = ctx.profile_steps(nprof)
    fs_ms = ms[2] / nprof
    es = 8 if prec == G.FP64 else 4
    fs_bytes = 2 * cells * 8 * es
    achieved = fs_bytes / (fs_ms * 1e-3) / 1e9
    step_bytes = 2 * (4 * 2 + 1) * es
    step_gbs = step_bytes * value / 1e9
    final = states_of(ctx, len(cases)) if nprof == steps else None
    host_io = None
    if a.host_io > 0:
        el_io = host_io_run(ctx, states, a.host_io, init_reinit)
        host_io = {"value": cells * a.host_io / el_io, "unit": "cell-updates/s", "steps": a.host_io,
                   "ms_per_step": el_io * 1e3 / a.host_io,
                   "what": "per step: upload of every instance's FP64 host grids (both fluids, phi), one eager "
                           "step, download of all grids, from Python through the C ABI"}
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
                   "gfm": GFM_LABEL[variant], "fs": "MUSCL", "ls": "LS5", "reinit_every": 5, "cfl": 0.6},
        "roofline": {"bound": "hbm", "kernel": "k_fs (MUSCL-Hancock + exact RS, both fluids)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (K_FS_TRAFFIC_PER_CELL[a.workload] * cells
                                 if a.workload in K_FS_TRAFFIC_PER_CELL and prec == G.MIXED else None),
                     "traffic_source": "rocprof FETCH_SIZE+WRITE_SIZE, copy-scaled (profiles/k_fs_batch64_mixed_traffic.md)"},
        "step_roofline": {"bytes_per_cell_update": step_bytes, "achieved": step_gbs, "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": step_gbs / HBM_PEAK_GBS},
        "phase_ms_per_step": {n: v / nprof for n, v in zip(("dt", "ghost", "fs", "ls", "bc"), ms)},
        "phase_steps": nprof,
        "status": sorted(set(status)),
        "schedule": schedule,
    }
    if host_io is not None:
        out["host_io_inclusive"] = host_io
    if prec == G.MIXED and not a.no_fp64:
        ctx64, sched64 = tuned_ctx(a, setup, states, variant, G.FP64, steps + a.warmup, init_reinit)
        el64 = timed_run(ctx64, states, steps, a.warmup, init_reinit, a.eager)
        out["fp64_plain"] = {"value": cells * steps / el64, "steps": steps,
                             "ratio_mixed_over_fp64": value / (cells * steps / el64), "schedule": sched64}
        ref = states_of(ctx64, len(cases))
        ctx64.close()
        if final is not None:
            acc = {"vs": "FP64 GPU path, same job, after %d steps" % steps,
                   GFM_LABEL[variant]: accuracy(infos, final, ref)}
            if a.perturb:
                pert = run_job(builder, cases, variant, G.FP64, steps, init_reinit, perturb=1e-7)
                acc["fp64_perturbed_1e-7"] = accuracy(infos, pert, ref)
            if a.accuracy and a.workload == "large":
                for v in (G.GFM_ORIGINAL, G.GFM_REAL):
                    mixed = run_job(builder, cases, v, G.MIXED, steps, init_reinit)
                    fp64 = run_job(builder, cases, v, G.FP64, steps, init_reinit)
                    acc[GFM_LABEL[v]] = accuracy(infos, mixed, fp64)
            out["accuracy"] = acc
    if not a.no_cpu:
        small = sorted(cases, key=lambda c: c["Nx"] * c["Ny"])[:2]
        out["cpu_baseline"] = cpu_baseline(builder, small, variant, a.cpu_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
