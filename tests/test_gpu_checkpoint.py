"""Checkpoint and resume: a batch stepped n1 steps, downloaded with each
instance's time and step count (gfm_download_state, gfm_get_time), uploaded
into a fresh context (gfm_upload_state, gfm_set_time) and stepped n2 more
steps, ends bit for bit where the same batch stepped n1 + n2 steps without
interruption ends.  Everything a step reads is either in the downloaded
state or rebuilt by the step itself (velocity field, primitives,
interfacial list, dt).  The resumed context uses another instance-group
schedule, so the test also covers the bench's window replay."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import gfm_hip as G
import problems as P


def ctx_of(cases, variant, precision, groups):
    old = os.environ.get("GFM_STEP_GROUPS")
    os.environ["GFM_STEP_GROUPS"] = str(groups)
    try:
        ctx = G.Context(G.methods(sim_type=G.SIM_TWOFLUID, fs=G.FS_MUSCL, gfm=variant, ls_order=5,
                                  reinit_every=5, reinit_iters=10, precision=precision))
        for info, eos, _ in cases:
            ctx.add(info, eos, 0.6, 1e300, 1000)
        ctx.commit()
    finally:
        if old is None:
            del os.environ["GFM_STEP_GROUPS"]
        else:
            os.environ["GFM_STEP_GROUPS"] = old
    return ctx


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED])
@pytest.mark.parametrize("variant", [G.GFM_RIEMANN, G.GFM_MODIFIED, G.GFM_ORIGINAL])
def test_resume_matches_uninterrupted(variant, precision):
    print("case checkpoint variant=%d precision=%d" % (variant, precision), flush=True)
    cases = [P.shocked_bubble(**c) for c in P.bubble_batch(3, smin=0.6, smax=1.0)]
    n1, n2 = 23, 17       # n1 ends mid reinitialisation period
    a = ctx_of(cases, variant, precision, groups=1)
    for k, (_, _, st) in enumerate(cases):
        a.upload(k, *st)
    a.step_graph(n1)
    ck = [(a.download(k), a.get_time(k)[:2]) for k in range(len(cases))]
    a.step_graph(n2)
    a.sync()
    ref = [a.download(k) for k in range(len(cases))]
    assert all(a.status(k) == 0 for k in range(len(cases)))
    a.close()
    b = ctx_of(cases, variant, precision, groups=3)
    for k, (st, (t, nstep)) in enumerate(ck):
        assert nstep == n1
        b.upload(k, *st)
        b.set_time(k, t, nstep)
    b.step_graph(n2)
    b.sync()
    for k in range(len(cases)):
        assert b.get_time(k)[1] == n1 + n2
        for x, y in zip(b.download(k), ref[k]):
            assert np.array_equal(x, y)
    b.close()
