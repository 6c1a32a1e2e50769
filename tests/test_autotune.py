"""Schedule and tile autotuner (phd-2d-gfm_amd/autotune.py) on a fake context and a fake clock: it
must try each group count with the blocked sweeps, then the one-sweep schedule at the best count, then
the other flux tiles, then (blocked schedule only) the other extrapolation tiles, then the persistent
flux kernel, pick the fastest, and leave the caller's environment as it was."""
import os
import sys

import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "phd-2d-gfm_amd"))
import autotune as T  # noqa: E402

MS = {1: 4.51, 2: 3.80, 3: 3.92, 4: 3.94}   # ms per step by group count, blocked sweeps
FS = {"32x16": 1.0, "64x8": 1.05, "16x32": 0.97, "32x8": 1.1}   # factor by flux tile
TB = {"64x16": 1.0, "32x16": 1.02, "64x8": 0.99}                # factor by extrapolation tile


class Clock:
    t = 0.0

    def __call__(self):
        return self.t


class FakeCtx:
    def __init__(self, clock, built):
        self.clock = clock
        self.groups = int(os.environ["GFM_STEP_GROUPS"])
        self.blocking = os.environ.get("GFM_NO_TEMPORAL_BLOCKING") != "1"
        self.fs = os.environ.get("GFM_FS_TILE", "32x16")
        self.tb = os.environ.get("GFM_TB_TILE", "64x16")
        self.persistent = os.environ.get("GFM_FS_PERSISTENT") == "1"
        self.closed = False
        built.append((self.groups, self.blocking, self.fs, self.tb))

    def factor(self):
        return FS[self.fs] * TB[self.tb] * (1.01 if self.persistent else 1.0)

    def step_graph(self, n):
        self.clock.t += n * MS[self.groups] * (1.0 if self.blocking else 2.0) * self.factor() * 1e-3

    def sync(self):
        pass

    def close(self):
        self.closed = True


@pytest.fixture
def clock(monkeypatch):
    c = Clock()
    monkeypatch.setattr(T.time, "perf_counter", c)
    return c


def test_picks_fastest_schedule(clock, monkeypatch):
    monkeypatch.setenv("GFM_STEP_GROUPS", "3")
    monkeypatch.delenv("GFM_NO_TEMPORAL_BLOCKING", raising=False)
    built, ctxs, loads = [], [], []

    def build():
        ctxs.append(FakeCtx(clock, built))
        return ctxs[-1]

    choice = T.tune(build, loads.append, n_instances=64, steps=50, warmup=5)
    assert (choice["groups"], choice["blocking"], choice["fs_tile"], choice["tb_tile"]) == (2, True, "16x32", "64x8")
    assert choice["ms_per_step"] == pytest.approx(3.80 * 0.97 * 0.99)
    d = ("32x16", "64x16")
    assert built == [(1, True) + d, (2, True) + d, (3, True) + d, (4, True) + d, (2, False) + d,
                     (2, True, "64x8", "64x16"), (2, True, "16x32", "64x16"), (2, True, "32x8", "64x16"),
                     (2, True, "16x32", "32x16"), (2, True, "16x32", "64x8"), (2, True, "16x32", "64x8")]
    assert len(choice["trials"]) == 11 and all(c.closed for c in ctxs)
    assert [t["persistent"] for t in choice["trials"]] == [False] * 10 + [True]
    assert choice["persistent"] is False                 # 1% slower here
    assert len(loads) == 2 * len(ctxs)                  # warmup from the initial state, then the timed run
    assert os.environ["GFM_STEP_GROUPS"] == "3"          # caller's environment restored
    assert "GFM_NO_TEMPORAL_BLOCKING" not in os.environ
    with T.applied(choice):
        assert os.environ["GFM_STEP_GROUPS"] == "2"
        assert "GFM_NO_TEMPORAL_BLOCKING" not in os.environ
        assert (os.environ["GFM_FS_TILE"], os.environ["GFM_TB_TILE"]) == ("16x32", "64x8")
    assert os.environ["GFM_STEP_GROUPS"] == "3"
    assert "GFM_FS_TILE" not in os.environ and "GFM_TB_TILE" not in os.environ
    assert "GFM_FS_PERSISTENT" not in os.environ


def test_single_instance_tries_one_group(clock, monkeypatch):
    monkeypatch.delenv("GFM_STEP_GROUPS", raising=False)
    built = []
    choice = T.tune(lambda: FakeCtx(clock, built), lambda c: None, n_instances=1, steps=10, warmup=2)
    assert [b[:2] for b in built] == [(1, True), (1, False)] + [(1, True)] * 6
    assert (choice["groups"], choice["blocking"]) == (1, True)
    assert "GFM_STEP_GROUPS" not in os.environ


def test_blocking_off_wins_when_faster(clock, monkeypatch):
    monkeypatch.setitem(MS, 2, 3.80)

    class Inverted(FakeCtx):
        def step_graph(self, n):
            self.clock.t += n * MS[self.groups] * (1.0 if self.blocking else 0.5) * 1e-3

    built = []
    choice = T.tune(lambda: Inverted(clock, built), lambda c: None, n_instances=8, steps=10, warmup=2)
    assert (choice["groups"], choice["blocking"]) == (2, False)
    # the one-sweep schedule has no blocked extrapolation: its tile is not tried
    assert len(built) == 5 + 3 + 1 and all(b[3] == "64x16" for b in built)
    with T.applied(choice):
        assert os.environ["GFM_NO_TEMPORAL_BLOCKING"] == "1"
