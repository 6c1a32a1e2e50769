"""Schedule autotuner for a batched two-fluid context.

It picks the step schedule of a workload by timing short runs of the workload itself:
  - the number of instance groups, each stepped on its own stream (GFM_STEP_GROUPS, 1-4);
  - the sweep schedule: the temporally blocked kernels (one launch per chain of sweeps) or one launch
    per sweep (GFM_NO_TEMPORAL_BLOCKING=1). It is tried only at the best group count, since the
    blocked kernels have won on every measured workload.

The library reads both settings when a context is committed, so every candidate gets a fresh context,
built by the caller's `build_ctx` under the candidate's environment. The tile shapes of the kernels
are compile-time constants and are not tuned here. The steps per captured graph (GFM_GRAPH_STEPS)
matter only for runs longer than a tuning run; 100 and 300 measured alike and 1500 slower.
"""
import contextlib
import os
import time

GROUPS = (1, 2, 3, 4)
KEYS = ("GFM_STEP_GROUPS", "GFM_NO_TEMPORAL_BLOCKING")


@contextlib.contextmanager
def _env(**kv):
    old = {k: os.environ.get(k) for k in kv}
    try:
        for k, v in kv.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
        yield
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def _settings(groups, blocking):
    return {"GFM_STEP_GROUPS": str(groups), "GFM_NO_TEMPORAL_BLOCKING": None if blocking else "1"}


def _time(build_ctx, load, settings, steps, warmup):
    with _env(**settings):
        ctx = build_ctx()
    try:
        # the warmup has the timed call's shape, so that every graph the
        # timed call replays is captured before the clock starts
        load(ctx)
        ctx.step_graph(max(warmup, steps))
        ctx.sync()
        load(ctx)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.step_graph(steps)
        ctx.sync()
        return (time.perf_counter() - t0) * 1e3 / steps
    finally:
        ctx.close()


def tune(build_ctx, load, n_instances, steps=100, warmup=10, groups=GROUPS):
    """Return the fastest schedule for the workload.

    build_ctx() -> a committed Context; load(ctx) -> None puts the initial state on the device.
    The result is {"groups": g, "blocking": b, "ms_per_step": t, "trials": [...]}, with one trial per
    candidate as {"groups", "blocking", "ms_per_step"}. A single instance takes one group, so only
    the sweep schedule is tried for it. The caller's environment is left as it was.
    """
    trials = []
    cand = [g for g in groups if g <= max(1, n_instances)] or [1]
    for g in cand:
        trials.append({"groups": g, "blocking": True,
                       "ms_per_step": _time(build_ctx, load, _settings(g, True), steps, warmup)})
    best = min(trials, key=lambda t: t["ms_per_step"])
    trials.append({"groups": best["groups"], "blocking": False,
                   "ms_per_step": _time(build_ctx, load, _settings(best["groups"], False), steps, warmup)})
    best = min(trials, key=lambda t: t["ms_per_step"])
    return dict(best, trials=trials)


def settings_of(choice):
    """the environment that makes a context use a tuned schedule"""
    return _settings(choice["groups"], choice["blocking"])


@contextlib.contextmanager
def applied(choice):
    """apply a tuned schedule to the contexts committed inside the block"""
    with _env(**settings_of(choice)):
        yield
