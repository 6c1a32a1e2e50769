"""Accuracy of the mixed-precision path on the benchmark workloads, at full size and full length.

For one workload and GFM variant, these jobs run side by side on the GPU from the same initial condition:
  - the FP64 path (the reference restatement): the yardstick;
  - the mixed-precision path (FP32 state and sweeps, FP64 Newton iterations, FP64 interfacial states,
    FP64 sums);
  - the FP64 path from the initial state perturbed by `--perturb` relative (uniform noise on the
    conserved variables of both fluids), once per seed: how far the flow itself carries a change in the
    last digits.  Shocked interfaces are unstable, so after some steps any perturbation grows; an error
    of the mixed path within the spread of these runs is error the flow amplifies, not error the FP32
    stages add;
  - the FP64 path with its state (both fluids' conserved grids and phi) rounded to FP32 after every
    step (GFM_ROUND_STATE_F32=1, `fp64_f32storage`): FP32 storage alone, which perturbs the state by
    up to half an FP32 ulp on every step rather than once.  The mixed path stores its state in FP32
    too, so the part of its error above this job's is what its FP32 arithmetic adds.
At each mark every job's state is compared with the FP64 job's (accuracy.py: relative L1 / Linf per
conserved variable of the composite real-fluid state, and the interface-contour L1; cell-weighted mean
over instances, max for Linf).

Workloads: `large` (one 4096x1152 shocked bubble, 2000 steps), `batch` (64 seeded shocked bubbles, 1500
steps), `rmi` (256 seeded periodic-y RMI instances, GFM_modified, 3000 steps).

usage: python accuracy_study.py --workload large --variant riemannGFM [--steps N] [--marks K] [--seeds S]
prints one JSON object and a markdown table
"""
import argparse
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import accuracy as A  # noqa: E402
import gfm_hip as G  # noqa: E402
import problems as P  # noqa: E402

SEED = 20261203
VARIANTS = {"OGFM": G.GFM_ORIGINAL, "realGFM": G.GFM_REAL, "riemannGFM": G.GFM_RIEMANN, "MGFM": G.GFM_MODIFIED}


def workload(name, n_inst=None):
    """(builder, cases, default variant, job steps, initial reinitialisation iterations)"""
    if name == "large":
        rng = np.random.default_rng(SEED)
        c = dict(Nx=4096, Ny=1152, bubble_gamma=float(rng.uniform(1.09, 1.67)),
                 bubble_rho=float(rng.uniform(0.18, 5.0)))
        return P.shocked_bubble, [c], "riemannGFM", 2000, 0
    if name == "batch":
        return P.shocked_bubble, P.bubble_batch(n_inst or 64), "riemannGFM", 1500, 0
    if name == "rmi":
        return P.rmi, P.rmi_batch(n_inst or 256), "MGFM", 3000, 40
    raise ValueError(name)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def perturbed(states, rel, seed):
    rng = np.random.default_rng(seed)
    out = []
    for g1, g2, phi in states:
        out.append((np.ascontiguousarray(g1 * (1.0 + rel * rng.uniform(-1.0, 1.0, g1.shape))),
                    np.ascontiguousarray(g2 * (1.0 + rel * rng.uniform(-1.0, 1.0, g2.shape))), phi))
    return out


def job(setup, states, variant, precision, steps, init_reinit, env=None):
    ctx = G.Context(G.methods(sim_type=G.SIM_TWOFLUID, fs=G.FS_MUSCL, gfm=variant, ls_order=5,
                              reinit_every=5, reinit_iters=10, precision=precision))
    for info, eos in setup:
        ctx.add(info, eos, 0.6, 1e300, steps + 1)
    env = env or {}
    saved = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        ctx.commit()   # the library reads its switches here
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    for k, st in enumerate(states):
        ctx.upload(k, *st)
        ctx.set_time(k, 0.0, 0)
    if init_reinit:
        ctx.ls_reinit(init_reinit)
    return ctx


def study(name, variant_name=None, steps=None, marks=10, seeds=3, rel=1e-7, n_inst=None):
    builder, cases, vdef, job_steps, init_reinit = workload(name, n_inst)
    variant = VARIANTS[variant_name or vdef]
    steps = steps or job_steps
    setup, states = [], []
    for c in cases:
        info, eos, st = builder(**c)
        setup.append((info, eos))
        states.append(st)
    infos = [s[0] for s in setup]
    w = [i.Nx * i.Ny for i in infos]
    log("initial conditions of %d instances built" % len(setup))
    runs = {"fp64": job(setup, states, variant, G.FP64, steps, init_reinit),
            "mixed": job(setup, states, variant, G.MIXED, steps, init_reinit),
            "fp64_f32storage": job(setup, states, variant, G.FP64, steps, init_reinit,
                                   {"GFM_ROUND_STATE_F32": "1"})}
    for s in range(seeds):
        runs["fp64_perturbed_%d" % (s + 1)] = job(setup, perturbed(states, rel, 1000 + s), variant, G.FP64, steps,
                                                  init_reinit)
    log("%d jobs loaded" % len(runs))
    at = sorted({int(round(steps * (k + 1) / marks)) for k in range(marks)})
    done = 0
    curves = {k: [] for k in runs if k != "fp64"}
    for m in at:
        for key, ctx in runs.items():
            ctx.step_graph(m - done)
            ctx.sync()
            log("%s at step %d" % (key, m))
        done = m
        ref = [runs["fp64"].download(k) for k in range(len(setup))]
        for key, ctx in runs.items():
            if key == "fp64":
                continue
            got = [ctx.download(k) for k in range(len(setup))]
            agg = A.aggregate([A.compare(i, a, b) for i, a, b in zip(infos, got, ref)], w)
            agg["step"] = m
            curves[key].append(agg)
            log("step %d %s L1_rho %.2e" % (m, key, agg["L1_rho"]))
            del got
        del ref
    status = {k: sorted({ctx.status(i) for i in range(len(setup))}) for k, ctx in runs.items()}
    for ctx in runs.values():
        ctx.close()
    return {"workload": name, "variant": variant_name or vdef, "steps": steps, "instances": len(setup),
            "cells": int(sum(w)), "perturbation": rel, "seeds": seeds, "status": status, "curves": curves}


def table(res, key="L1_rho"):
    cols = list(res["curves"])
    lines = ["| step | " + " | ".join(cols) + " |", "|---:|" + "---:|" * len(cols)]
    for r in range(len(res["curves"][cols[0]])):
        lines.append("| %d | " % res["curves"][cols[0]][r]["step"]
                     + " | ".join("%.2e" % res["curves"][c][r][key] for c in cols) + " |")
    return "\n".join(lines)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="large", choices=["large", "batch", "rmi"])
    ap.add_argument("--variant", default=None, choices=list(VARIANTS))
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--marks", type=int, default=10)
    ap.add_argument("--seeds", type=int, default=3)
    ap.add_argument("--perturb", type=float, default=1e-7)
    ap.add_argument("--instances", type=int, default=None)
    a = ap.parse_args(argv)
    res = study(a.workload, a.variant, a.steps, a.marks, a.seeds, a.perturb, a.instances)
    print(json.dumps(res))
    for key in ("L1_rho", "Linf_rho", "L1_E", "contour_L1_rel"):
        print("\n%s\n" % key)
        print(table(res, key))
    return 0


if __name__ == "__main__":
    sys.exit(main())
