"""Schedule and tile-shape autotuner for a two-fluid context.

It picks the step schedule and the kernels' tile shapes of a workload by timing short runs of the
workload itself, in this order (each stage keeps the best of the stages before it):
  1. the number of instance groups, each stepped on its own stream (GFM_STEP_GROUPS, 1-4);
  2. the sweep schedule: the temporally blocked kernels (one launch per chain of sweeps) or one launch
     per sweep (GFM_NO_TEMPORAL_BLOCKING=1);
  3. the flux kernel's tile (GFM_FS_TILE: 32x16, 64x8, 16x32 or 32x8 cells, one thread per cell);
  4. with the blocked schedule, the blocked extrapolation's tile (GFM_TB_TILE: 64x16 cells on 512
     threads, 32x16 or 64x8 on 256);
  5. the flux sweep's form: one workgroup per tile, or the persistent kernel whose grid of as many
     workgroups as the GPU holds drains the tiles through per-XCD queues (GFM_FS_PERSISTENT=1).
The library reads these settings when a context is committed (an unknown tile shape is an error), so
every candidate gets a fresh context, built by the caller's `build_ctx` under the candidate's
environment. The steps per captured graph (GFM_GRAPH_STEPS) matter only for runs longer than a tuning
run; 100 and 300 measured alike and 1500 slower.
"""
import contextlib
import os
import time

GROUPS = (1, 2, 3, 4)
FS_TILES = ("32x16", "64x8", "16x32", "32x8")   # the library's compiled flux tiles, default first
TB_TILES = ("64x16", "32x16", "64x8")           # the blocked extrapolation's, default first
KEYS = ("GFM_STEP_GROUPS", "GFM_NO_TEMPORAL_BLOCKING", "GFM_FS_TILE", "GFM_TB_TILE", "GFM_FS_PERSISTENT")


@contextlib.contextmanager
def environment(**kv):
    """set (value) or unset (None) environment variables inside the block"""
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


def _settings(groups, blocking, fs_tile=FS_TILES[0], tb_tile=TB_TILES[0], persistent=False):
    return {"GFM_STEP_GROUPS": str(groups), "GFM_NO_TEMPORAL_BLOCKING": None if blocking else "1",
            "GFM_FS_TILE": fs_tile, "GFM_TB_TILE": tb_tile, "GFM_FS_PERSISTENT": "1" if persistent else None}


def _trial(build_ctx, load, steps, warmup, groups, blocking, fs_tile=FS_TILES[0], tb_tile=TB_TILES[0],
           persistent=False):
    ms = _time(build_ctx, load, _settings(groups, blocking, fs_tile, tb_tile, persistent), steps, warmup)
    return {"groups": groups, "blocking": blocking, "fs_tile": fs_tile, "tb_tile": tb_tile,
            "persistent": persistent, "ms_per_step": ms}


def _time(build_ctx, load, settings, steps, warmup):
    with environment(**settings):
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


def tune(build_ctx, load, n_instances, steps=100, warmup=10, groups=GROUPS, fs_tiles=FS_TILES,
         tb_tiles=TB_TILES):
    """Return the fastest schedule and tile shapes for the workload.

    build_ctx() -> a committed Context; load(ctx) -> None puts the initial state on the device.
    The result is {"groups", "blocking", "fs_tile", "tb_tile", "persistent", "ms_per_step", "trials": [...]}, with
    one trial per candidate in the same form. A single instance takes one group, so only the other
    stages are tried for it. The caller's environment is left as it was.
    """
    trials = []

    def best():
        return min(trials, key=lambda t: t["ms_per_step"])

    cand = [g for g in groups if g <= max(1, n_instances)] or [1]
    for g in cand:
        trials.append(_trial(build_ctx, load, steps, warmup, g, True))
    trials.append(_trial(build_ctx, load, steps, warmup, best()["groups"], False))
    b = best()
    for ft in fs_tiles:
        if ft != b["fs_tile"]:
            trials.append(_trial(build_ctx, load, steps, warmup, b["groups"], b["blocking"], ft))
    b = best()
    if b["blocking"]:
        for tt in tb_tiles:
            if tt != b["tb_tile"]:
                trials.append(_trial(build_ctx, load, steps, warmup, b["groups"], True, b["fs_tile"], tt))
    b = best()
    trials.append(_trial(build_ctx, load, steps, warmup, b["groups"], b["blocking"], b["fs_tile"], b["tb_tile"],
                         True))
    return dict(best(), trials=trials)


def settings_of(choice):
    """the environment that makes a context use a tuned schedule and tile shapes"""
    return _settings(choice["groups"], choice["blocking"], choice.get("fs_tile", FS_TILES[0]),
                     choice.get("tb_tile", TB_TILES[0]), choice.get("persistent", False))


@contextlib.contextmanager
def applied(choice):
    """apply a tuned schedule to the contexts committed inside the block"""
    with environment(**settings_of(choice)):
        yield
