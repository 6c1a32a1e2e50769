"""Benchmark of the GFM per-timestep hot path on one MI355X.

A "step" is one full two-fluid time step of every instance: CFL dt, ghost-
fluid fill, MUSCL-Hancock exact-RS update of both fluids, HOUC-5 level-set
advection with reinitialisation every 5 steps, boundary conditions.

Timed window.  Each precision's job starts from the initial condition and
is advanced untimed to mid-job (--advance, default half the job length), so
that the timed steps do the per-step work of a developed flow rather than of
the quiescent start.  The state there is checkpointed (downloaded with its
time and step count); the schedule autotuner times its candidates from that
checkpoint; then --warmup untimed steps, and the --steps timed steps.
Inputs are resident in HBM before the timed region; a single instance runs
one hipGraph per step, a batch runs its instance groups (two streams in
mixed precision, four in FP64) from one graph per 100-step block (--eager:
direct launches).

Workloads (--workload):
  batch   (default) 64 seeded shocked bubbles, grids s*(256x72), s~U[1,8]
          (2.55e7 interior cells), per-instance Mach/radius/density/gamma,
          GFM_riemann, mixed precision (FP32 state and sweeps, FP64 Newton,
          FP64 interfacial states, FP64 sums), per-instance dt; 1500-step job;
  bubble  one 1024x288 shocked bubble, plain FP64, GFM_riemann, 2000 steps;
  large   one 4096x1152 shocked bubble (bubble gamma and density drawn from
          seed 20261203), mixed precision, GFM_riemann, 2000 steps; with
          --accuracy also GFM_original and GFM_real, each against FP64;
  rmi     256 seeded single-mode Richtmyer-Meshkov instances, grids
          s*(512x128), s~U[1,4], periodic in y, GFM_modified, mixed
          precision; 300-step job unless --job-steps says otherwise.
`value` = interior cells x timed steps / wall time, summed over instances.

`roofline` is for the dominant kernel, the fused MUSCL-Hancock flux kernel
k_fs: algorithmic bytes per launch = 2 fluids x cells x 8 words (4 read, 4
written) x word size, over its mean duration from HIP events on the
context's stream.  Those events come from a replay of exactly the timed
steps: the checkpoint taken at the start of the timed window is restored
and the same steps are launched eagerly with events around each phase
(`replay_bitwise` says whether the replay reproduced the timed run's final
state bit for bit).  `traffic` is HBM bytes per k_fs launch from rocprof
FETCH_SIZE / WRITE_SIZE per launch (--traffic-kib, with --copy-factors: what those counters read for a cold
device copy of known size, --copy-calibration); for the default workload
the values measured over its default window are built in, with the read
bytes taken from the L2-to-fabric read requests by size.  `step_roofline` is the
whole step against the compulsory state traffic 2 (4 F + L) s bytes per
cell-update (F = 2 fluids, L = 1 level set, s = word size).

For mixed-precision workloads the FP64 path runs the same job the same way
(`fp64_plain`) and the states after the timed window are compared
(`accuracy`: relative L1 / Linf per conserved variable of the composite
real-fluid state, interface-contour L1), beside the same comparison for the
FP64 job with its state rounded to FP32 after every step
(`fp64_f32storage`: what FP32 storage alone does to the job; the mixed path
stores its state in FP32 too); --perturb adds the same comparison
between FP64 and FP64 from an initial state perturbed by 1e-7 relative,
which measures how much the flow itself amplifies a last-digit change.
`cpu_baseline` times the FP64 CPU oracle on a bounded sample of the same
window (the checkpointed mid-job states of the smallest instances, a few
steps each): serially, as the reference runs, and with one host thread per
instance on up to 16 cores (`cpu_baseline.all_cores`).
--host-io K adds `host_io_inclusive`: K steps with every instance's grids
moved host -> device -> host around each step (the per-phase drop-in
boundary, PCIe included), reported beside `value`, never as it.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "phd-2d-gfm_amd"))

import numpy as np  # noqa: E402

import accuracy as A  # noqa: E402
import autotune as T  # noqa: E402
import gfm_hip as G  # noqa: E402
import problems as P  # noqa: E402

HBM_PEAK_GBS = 8000.0
SEED = 20261203
GFM_LABEL = {v: k for k, v in G.GFM_NAMES.items()}   # settings-file names: OGFM, realGFM, ..., OGFMVOF
GFM_CLASS = {G.GFM_ORIGINAL: "GFM_original", G.GFM_REAL: "GFM_real", G.GFM_RIEMANN: "GFM_riemann",
             G.GFM_MODIFIED: "GFM_modified", G.GFM_VOFGFM: "GFM_VOF", G.GFM_OGFMVOF: "OGFM_VOF"}
HOST_THREADS_MAX = 16   # the host cores one GPU's share of a box offers
# rocprof counters of the flux kernel per launch over the default workload's window (steps 760-860),
# profiles/batch64_window_counters.md: read bytes from the L2-to-fabric read requests by size, raw
# FETCH_SIZE and WRITE_SIZE in KiB (FETCH_SIZE counts gfx950's 128-byte read requests as 64 bytes),
# VALU busy, LDS bank conflicts per active LDS instruction
K_FS_COUNTERS = {"batch": {"read_bytes": 1.3057e9, "fetch_kib_raw": 6.4096e5, "write_kib": 4.3430e5,
                           "valu_busy": 0.785, "lds_bank_conflict_per_lds_inst": 0.68,
                           "measured_window": [760, 860], "source": "profiles/batch64_window_counters.md"}}


def log(msg):
    """progress on stderr, one line per stage, so that a long run shows it is alive"""
    print("[bench %.1fs] %s" % (time.perf_counter() - T0, msg), file=sys.stderr, flush=True)


T0 = time.perf_counter()


def make_cases(workload, n_inst):
    """(problem builder, per-instance kwargs, variant, precision, job steps, name, init reinit iterations)"""
    if workload == "batch":
        cases = P.bubble_batch(n_inst)
        return (P.shocked_bubble, cases, G.GFM_RIEMANN, G.MIXED, 1500,
                "%d seeded shocked bubbles s*(256x72), s~U[1,8], {gfm}, mixed FP32/FP64" % len(cases), 0)
    if workload == "bubble":
        return (P.shocked_bubble, [dict(Nx=1024, Ny=288)], G.GFM_RIEMANN, G.FP64, 2000,
                "shocked bubble 1024x288, {gfm}, plain FP64", 0)
    if workload == "large":
        rng = np.random.default_rng(SEED)
        c = dict(Nx=4096, Ny=1152, bubble_gamma=float(rng.uniform(1.09, 1.67)),
                 bubble_rho=float(rng.uniform(0.18, 5.0)))
        return (P.shocked_bubble, [c], G.GFM_RIEMANN, G.MIXED, 2000,
                "shocked bubble 4096x1152 (gamma %.4f, rho %.4f, seed %d), {gfm}, mixed FP32/FP64"
                % (c["bubble_gamma"], c["bubble_rho"], SEED), 0)
    cases = P.rmi_batch(n_inst)
    return (P.rmi, cases, G.GFM_MODIFIED, G.MIXED, 300,
            "%d seeded single-mode RMI s*(512x128), s~U[1,4], periodic y, {gfm}, mixed FP32/FP64"
            % len(cases), 40)


def make_setup(builder, cases, perturb=0.0):
    """per instance (sim info, EOS) and the initial host grids"""
    setup, states = [], []
    rng = np.random.default_rng(7)
    for k, c in enumerate(cases):
        if k and k % 32 == 0:
            log("  %d of %d built" % (k, len(cases)))
        info, eos, st = builder(**c)
        if perturb:
            st = tuple(np.ascontiguousarray(a * (1.0 + perturb * rng.uniform(-1.0, 1.0, a.shape)))
                       if k < 2 else a for k, a in enumerate(st))
        setup.append((info, eos))
        states.append(st)
    return setup, states


ITM = G.ITM_LS   # the interface tracker of every context (--itm)


def new_ctx(setup, variant, precision, max_steps):
    m = G.methods(sim_type=G.SIM_TWOFLUID, fs=G.FS_MUSCL, gfm=variant, ls_order=5,
                  reinit_every=5, reinit_iters=10, precision=precision, itm=ITM)
    ctx = G.Context(m)
    for info, eos in setup:
        ctx.add(info, eos, 0.6, 1e300, max_steps)
    ctx.commit()
    return ctx


def load_ic(ctx, states, init_reinit):
    for k, st in enumerate(states):
        ctx.upload(k, *st)
        ctx.set_time(k, 0.0, 0)
    if init_reinit:
        ctx.ls_reinit(init_reinit)   # ls->initialisation() (sim_twofluid.cpp:47)


def checkpoint(ctx, n):
    """every instance's grids, time and step count, and the CLSVOF tracker's z"""
    ctx.sync()
    return [(ctx.download(k), ctx.get_time(k)[:2], ctx.download_z(k) if ITM == G.ITM_CLSVOF else None)
            for k in range(n)]


def restore(ctx, ck):
    for k, (st, (t, nstep), z) in enumerate(ck):
        ctx.upload(k, *st)
        if z is not None:
            ctx.upload_z(k, z)
        ctx.set_time(k, t, nstep)
    ctx.sync()


def advanced(setup, states, variant, precision, advance, cap, init_reinit, eager=False):
    """the job advanced `advance` steps from its initial condition (default schedule), checkpointed"""
    ctx = new_ctx(setup, variant, precision, cap)
    load_ic(ctx, states, init_reinit)
    (ctx.step if eager else ctx.step_graph)(advance)
    ck = checkpoint(ctx, len(setup))
    status = sorted({ctx.status(k) for k in range(len(setup))})
    ctx.close()
    return ck, status


def tuned_ctx(a, setup, ck, variant, precision, cap):
    """a context on the autotuned schedule (autotune.py), trials timed from the checkpoint `ck`, unless
    --no-autotune, --eager, or the schedule is set in the environment; returns (ctx, schedule record)"""
    if a.no_autotune or a.eager or any(k in os.environ for k in T.KEYS):
        rec = {"tuned": False, "env": {k: os.environ[k] for k in T.KEYS if k in os.environ}}
        return new_ctx(setup, variant, precision, cap), rec
    ts = a.tune_steps if a.tune_steps else a.steps
    choice = T.tune(lambda: new_ctx(setup, variant, precision, cap), lambda c: restore(c, ck), len(setup),
                    steps=ts, warmup=a.warmup)
    with T.applied(choice):
        ctx = new_ctx(setup, variant, precision, cap)
    return ctx, dict(choice, tuned=True, tune_steps=ts, from_step=ck[0][1][1])


def timed_window(ctx, ck, steps, warmup, eager=False):
    """restore `ck`, `warmup` untimed steps, checkpoint, `steps` timed steps; returns
    (seconds, checkpoint at the window start, final states)"""
    step = ctx.step if eager else ctx.step_graph
    if not eager:
        # capture every graph the timed call replays (blocks of steps are
        # captured at their first use), then start again from `ck`
        restore(ctx, ck)
        step(steps)
    restore(ctx, ck)
    step(warmup)
    win = checkpoint(ctx, len(ck))
    t0 = time.perf_counter()
    step(steps)
    ctx.sync()
    el = time.perf_counter() - t0
    return el, win, [ctx.download(k) for k in range(len(ck))]


def same_states(a, b):
    return all(all(np.array_equal(x, y) for x, y in zip(sa, sb)) for sa, sb in zip(a, b))


def oracle_job(builder_args):
    """one instance's FP64 oracle run from a checkpointed state (called from host threads; the oracle's
    scratch is thread-local and ctypes releases the GIL)"""
    import oracle_lib as O
    info, eos, st, variant, steps = builder_args
    g1, g2, phi = (np.ascontiguousarray(x, dtype=np.float64).copy() for x in st)
    Gr = O.Grid(info.Nx, info.Ny, info.numGC, info.x0, info.y0, info.dx, info.dy,
                info.BC_L, info.BC_R, info.BC_B, info.BC_T)
    E = O.Eos(eos.gamma1, eos.pinf1, eos.gamma2, eos.pinf2)
    run = O.Run(variant, 0, 5, 5, 10, 0.6, 1e300, steps, 0, 0,
                O.ITM_CLSVOF if ITM == G.ITM_CLSVOF else O.ITM_LS)
    s = np.zeros(1, dtype=np.int32)
    t = np.zeros(1)
    O.lib().or_twofluid_run(Gr, E, run, O._p(g1), O._p(g2), O._p(phi), O._p(s), O._p(t))
    return info.Nx * info.Ny * int(s[0])


CPU_BUDGET = 1.5e7   # cell-steps per CPU leg: about 10 s of one core at the oracle's ~1.5e6 cell-updates/s


def cpu_baseline(setup, ck, variant, max_steps, n_serial, n_threads):
    """FP64 CPU oracle on the smallest instances' checkpointed mid-job states: serially (one core, as the
    reference runs) and with one host thread per instance; each leg runs about CPU_BUDGET cell-steps
    (at most max_steps steps per instance)"""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    O.lib()
    order = sorted(range(len(setup)), key=lambda k: setup[k][0].Nx * setup[k][0].Ny)
    size = lambda k: setup[k][0].Nx * setup[k][0].Ny  # noqa: E731
    ser = order[:n_serial]
    steps_s = int(max(1, min(max_steps, round(CPU_BUDGET / sum(size(k) for k in ser)))))
    jobs = [(setup[k][0], setup[k][1], ck[k][0], variant, steps_s) for k in ser]
    t0 = time.perf_counter()
    cells_s = sum(oracle_job(j) for j in jobs)
    el_s = time.perf_counter() - t0
    desc = lambda js: ", ".join("%dx%d" % (j[0].Nx, j[0].Ny) for j in js)  # noqa: E731
    start = ck[0][1][1]
    out = {"value": cells_s / el_s, "unit": "cell-updates/s", "cores": 1, "kind": "port",
           "sample": "FP64 oracle, one thread, %d steps from the step-%d checkpoint of %s: %.1f s"
                     % (steps_s, start, desc(jobs), el_s)}
    threads = max(1, min(n_threads, HOST_THREADS_MAX, os.cpu_count() or 1, len(setup)))
    if threads < 2:
        out["all_cores"] = {"value": None, "cores": 1, "nproc": os.cpu_count(),
                            "sample": "one instance: the oracle is serial, so there is no instance-level "
                                      "parallelism to time"}
        return out
    par = order[:threads]
    steps_p = int(max(1, min(max_steps, round(CPU_BUDGET / max(size(k) for k in par)))))
    pjobs = [(setup[k][0], setup[k][1], ck[k][0], variant, steps_p) for k in par]
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        cells_p = sum(ex.map(oracle_job, pjobs))
    el_p = time.perf_counter() - t0
    out["all_cores"] = {"value": cells_p / el_p, "unit": "cell-updates/s", "cores": threads,
                        "nproc": os.cpu_count(),
                        "sample": "FP64 oracle, one host thread per instance, %d steps from the step-%d "
                                  "checkpoint of %d instances (%s): %.1f s"
                                  % (steps_p, start, len(pjobs), desc(pjobs), el_p)}
    return out


def host_io_run(ctx, ck, steps):
    """the rate through the per-phase drop-in boundary: every step uploads
    each instance's host grids (FP64 cells, both fluids and phi), runs one
    step (eager launches) and downloads all grids again"""
    restore(ctx, ck)
    cur = [c[0] for c in ck]
    t0 = time.perf_counter()
    for _ in range(steps):
        for k, st in enumerate(cur):
            ctx.upload(k, *st)
        ctx.step(1)
        cur = [ctx.download(k) for k in range(len(cur))]
    ctx.sync()
    return time.perf_counter() - t0


def accuracy(infos, got, ref):
    res = [A.compare(info, a, b) for info, a, b in zip(infos, got, ref)]
    return A.aggregate(res, [i.Nx * i.Ny for i in infos])


def run_job(setup, states, variant, precision, steps, init_reinit, env=None):
    with T.environment(**(env or {})):   # the library reads its switches at commit
        ctx = new_ctx(setup, variant, precision, steps + 1)
    load_ic(ctx, states, init_reinit)
    ctx.step_graph(steps)
    ctx.sync()
    out = [ctx.download(k) for k in range(len(setup))]
    ctx.close()
    return out


def copy_calibration(nbytes):
    """one cold device-to-device copy of nbytes (hipMemcpy, the runtime's copy kernel): the source is
    written first, then a buffer of the same size, so that at most the last-written bytes sit in the
    caches when the copy reads the source.  rocprof's FETCH_SIZE / WRITE_SIZE of that dispatch over
    nbytes are the counters' scale factors."""
    hip = C.CDLL("libamdhip64.so")
    vp = C.c_void_p
    bufs = [vp(), vp(), vp()]
    for b in bufs:
        if hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc")
    src, dst, flush = bufs
    hip.hipMemset(src, 1, C.c_size_t(nbytes))
    hip.hipMemset(flush, 2, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3)   # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    for b in bufs:
        hip.hipFree(b)
    return {"copy_calibration_bytes": nbytes, "dispatch": "the last __amd_rocclr_copyBuffer of the run"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed steps")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--advance", type=int, default=None,
                    help="untimed steps from the initial condition before the window (default: half the job)")
    ap.add_argument("--job-steps", type=int, default=None, help="job length whose middle is timed")
    ap.add_argument("--workload", default="batch", choices=["batch", "bubble", "large", "rmi"])
    ap.add_argument("--instances", type=int, default=None)
    ap.add_argument("--gfm", default=None, choices=sorted(G.GFM_NAMES),
                    help="the workload with another ghost fluid method (settings-file name, e.g. VOFGFM)")
    ap.add_argument("--itm", default="LS5", choices=["LS5", "CLSVOF"],
                    help="interface tracker: the level set, or the CLSVOF tracker beside it")
    ap.add_argument("--cpu-steps", type=int, default=200, help="most oracle steps per instance in the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison run of mixed workloads")
    ap.add_argument("--no-replay", action="store_true", help="skip the per-phase HIP-event replay of the window")
    ap.add_argument("--accuracy", action="store_true",
                    help="large: also run GFM_original and GFM_real, mixed against FP64")
    ap.add_argument("--perturb", action="store_true", help="also compare FP64 against FP64 from a perturbed start")
    ap.add_argument("--eager", action="store_true", help="launch kernels directly instead of hipGraphs")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the library's default schedule instead of autotuning it (autotune.py)")
    ap.add_argument("--tune-steps", type=int, default=None, help="steps per autotuning trial (default: --steps)")
    ap.add_argument("--traffic-kib", type=float, nargs=2, default=None, metavar=("FETCH", "WRITE"),
                    help="rocprof FETCH_SIZE and WRITE_SIZE of k_fs per launch over the replay (KiB), with the "
                         "copy scale factors from --copy-factors")
    ap.add_argument("--copy-factors", type=float, nargs=2, default=(1.0, 1.0), metavar=("FETCH", "WRITE"))
    ap.add_argument("--host-io", type=int, default=0, metavar="K",
                    help="also time K steps that move every instance's grids host -> device -> host around each "
                         "step (the per-phase drop-in's rate, PCIe included; reported beside value, never as it)")
    ap.add_argument("--copy-calibration", type=float, default=None, metavar="GIB",
                    help="only run one cold device-to-device copy of GIB GiB, for scaling FETCH_SIZE / WRITE_SIZE")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one MI355X: the per-timestep path runs on one GPU here "
              "(instances would be bin-packed across GPUs; not built)", file=sys.stderr)
        sys.exit(2)
    if a.copy_calibration:
        print(json.dumps(copy_calibration(int(a.copy_calibration * (1 << 30)))))
        return

    global ITM
    ITM = G.ITM_CLSVOF if a.itm == "CLSVOF" else G.ITM_LS
    n_inst = a.instances or (64 if a.workload == "batch" else 256)
    builder, cases, variant, prec, job, name, init_reinit = make_cases(a.workload, n_inst)
    if a.gfm:
        variant = G.GFM_NAMES[a.gfm]
    name = name.format(gfm=GFM_CLASS[variant] + (" on the CLSVOF tracker" if ITM == G.ITM_CLSVOF else ""))
    job = a.job_steps or job
    advance = job // 2 if a.advance is None else a.advance
    cap = advance + a.warmup + max(a.steps, a.tune_steps or 0) + 8

    log("building the initial conditions of %d instances" % len(cases))
    setup, states = make_setup(builder, cases)
    infos = [s[0] for s in setup]
    cells = sum(i.Nx * i.Ny for i in infos)
    n = len(setup)

    log("advancing %d steps" % advance)
    ck, st_adv = advanced(setup, states, variant, prec, advance, cap, init_reinit, a.eager)
    log("autotuning")
    ctx, schedule = tuned_ctx(a, setup, ck, variant, prec, cap)
    log("timing %d steps" % a.steps)
    el, win, final = timed_window(ctx, ck, a.steps, a.warmup, a.eager)
    log("timed: %.3f ms per step" % (el * 1e3 / a.steps))
    status = sorted({ctx.status(k) for k in range(n)} | set(st_adv))
    value = cells * a.steps / el

    es = 8 if prec == G.FP64 else 4
    out = {
        "metric": "cell-updates/s on 1 MI355X (two-fluid GFM step; mixed+batched vs FP64 plain)",
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
        "config": {"workload": name, "instances": n, "cells": cells, "job_steps": job,
                   "window": [advance + a.warmup, advance + a.warmup + a.steps],
                   "gfm": GFM_LABEL[variant], "fs": "MUSCL", "ls": a.itm, "reinit_every": 5, "cfl": 0.6},
    }
    if not a.no_replay:
        # the same steps again from the window's checkpoint, eager, events around each phase
        restore(ctx, win)
        ms = ctx.profile_steps(a.steps)
        replay = [ctx.download(k) for k in range(n)]
        fs_ms = ms[2] / a.steps
        fs_bytes = 2 * cells * 8 * es
        achieved = fs_bytes / (fs_ms * 1e-3) / 1e9
        traffic, counters = None, None
        if a.traffic_kib:
            counters = {"fetch_kib": a.traffic_kib[0], "write_kib": a.traffic_kib[1],
                        "factors": tuple(a.copy_factors), "source": "--traffic-kib"}
        elif a.workload in K_FS_COUNTERS and prec == G.MIXED and n == 64:
            counters = K_FS_COUNTERS[a.workload]
        if counters:
            if "read_bytes" in counters:
                traffic = counters["read_bytes"] + 1024.0 * counters["write_kib"]
            else:
                traffic = 1024.0 * (counters["fetch_kib"] * counters["factors"][0]
                                    + counters["write_kib"] * counters["factors"][1])
        out["roofline"] = {"bound": "hbm", "kernel": "k_fs (MUSCL-Hancock + exact RS, both fluids)",
                           "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                           "bytes_per_launch": fs_bytes, "ms_per_launch": fs_ms,
                           "launches": a.steps, "window": "the timed steps, replayed from their checkpoint"}
        if counters:
            out["roofline"]["traffic_counters"] = counters
        out["phase_ms_per_step"] = {k: v / a.steps for k, v in zip(("dt", "ghost", "fs", "ls", "bc"), ms)}
        out["replay_bitwise"] = same_states(final, replay)
    step_bytes = 2 * (4 * 2 + 1) * es
    step_gbs = step_bytes * value / 1e9
    out["step_roofline"] = {"bytes_per_cell_update": step_bytes, "achieved": step_gbs, "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": step_gbs / HBM_PEAK_GBS}
    out["status"] = status
    out["schedule"] = schedule
    if a.host_io > 0:
        el_io = host_io_run(ctx, win, a.host_io)
        out["host_io_inclusive"] = {
            "value": cells * a.host_io / el_io, "unit": "cell-updates/s", "steps": a.host_io,
            "ms_per_step": el_io * 1e3 / a.host_io,
            "what": "per step: upload of every instance's FP64 host grids (both fluids, phi), one eager "
                    "step, download of all grids, from Python through the C ABI"}
    ctx.close()

    if prec == G.MIXED and not a.no_fp64:
        log("FP64 job")
        ck64, st64 = advanced(setup, states, variant, G.FP64, advance, cap, init_reinit, a.eager)
        ctx64, sched64 = tuned_ctx(a, setup, ck64, variant, G.FP64, cap)
        el64, _, ref = timed_window(ctx64, ck64, a.steps, a.warmup, a.eager)
        ctx64.close()
        out["fp64_plain"] = {"value": cells * a.steps / el64, "steps": a.steps,
                             "ratio_mixed_over_fp64": value / (cells * a.steps / el64), "schedule": sched64,
                             "status": st64}
        total = advance + a.warmup + a.steps
        acc = {"vs": "FP64 GPU path, same job, after %d steps from the initial condition" % total,
               GFM_LABEL[variant]: accuracy(infos, final, ref)}
        # the same FP64 job with its state rounded to FP32 after every step:
        # what FP32 storage alone does, beside which the mixed error is judged
        log("FP64 job with FP32 storage")
        st32 = run_job(setup, states, variant, G.FP64, total, init_reinit, {"GFM_ROUND_STATE_F32": "1"})
        acc["fp64_f32storage"] = accuracy(infos, st32, ref)
        acc["mixed_over_f32storage_L1_rho"] = (acc[GFM_LABEL[variant]]["L1_rho"] /
                                               max(acc["fp64_f32storage"]["L1_rho"], 1e-300))
        del st32
        if a.perturb:
            log("FP64 job from a perturbed start")
            _, pstates = make_setup(builder, cases, perturb=1e-7)
            pert = run_job(setup, pstates, variant, G.FP64, total, init_reinit)
            acc["fp64_perturbed_1e-7"] = accuracy(infos, pert, ref)
        if a.accuracy and a.workload == "large":
            for v in (G.GFM_ORIGINAL, G.GFM_REAL):
                log("variant %s" % GFM_LABEL[v])
                mixed = run_job(setup, states, v, G.MIXED, total, init_reinit)
                fp64 = run_job(setup, states, v, G.FP64, total, init_reinit)
                acc[GFM_LABEL[v]] = accuracy(infos, mixed, fp64)
        out["accuracy"] = acc
    if not a.no_cpu:
        log("CPU baseline")
        out["cpu_baseline"] = cpu_baseline(setup, ck, variant, a.cpu_steps, 4, HOST_THREADS_MAX)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
