"""The GPU initial-condition library (gfm_ic_device) against the host one (gfm_ic_onefluid, gfm_ic_bubble,
gfm_ic_rmi, gfm_ic_planar): the same FP64 arithmetic per cell, so the same state bit for bit in FP64 and
in FP32 storage, except the RMI interface's cos / sin, whose device and host libraries may differ in the
last bits (level set to 4e-16 relative).  Several instances of different kinds and sizes in one call;
time and step count reset; instances not listed keep their state."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import gfm_hip as G
import problems as P

T, R = G.BC_TRANSMISSIVE, G.BC_REFLECTIVE


def twofluid_cases():
    out = []
    for c in P.bubble_batch(3, smin=0.5, smax=1.0):
        info, eos, desc = P.shocked_bubble_ic(**c)
        out.append((info, eos, desc, P.shocked_bubble(**c)[2], False))
    for c in P.rmi_batch(2, smin=0.3, smax=0.5):
        info, eos, desc = P.rmi_ic(**c)
        out.append((info, eos, desc, P.rmi(**c)[2], True))
    info = G.SimInfo(64, 40, 8, -2.0, -2.0, 4.0 / 64, 4.0 / 40, T, T, T, T)
    eos = G.Eos(4.4, 6e8, 1.4, 0.0)
    w1, w2 = [1000.0, 0.0, 0.0, 1e9], [50.0, 0.0, 0.0, 1e5]
    out.append((info, eos, G.ic_desc(G.IC_PLANAR, w1, w2, axis=1, pos=0.7), G.ic_planar(info, eos, 1, 0.7, w1, w2),
                False))
    return out


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED])
def test_device_ic_twofluid_matches_host(precision):
    print("CASE device ic twofluid precision=%d" % precision, flush=True)
    cases = twofluid_cases()
    ctx = G.Context(G.methods(sim_type=G.SIM_TWOFLUID, gfm=G.GFM_RIEMANN, precision=precision))
    for info, eos, _, _, _ in cases:
        ctx.add(info, eos, 0.6)
    ctx.commit()
    for k, (info, _, _, st, _) in enumerate(cases):
        ctx.upload(k, *[np.zeros_like(a) for a in st])
        ctx.set_time(k, 1.0, 7)
    keep = len(cases) - 1   # not listed: keeps its (zero) state and time
    ctx.ic_device([(k, c[2]) for k, c in enumerate(cases) if k != keep])
    ctx.sync()
    for k, (info, _, _, host, trig) in enumerate(cases):
        got = ctx.download(k)
        if k == keep:
            assert all(not a.any() for a in got) and ctx.get_time(k)[:2] == (1.0, 7)
            continue
        assert ctx.get_time(k)[:2] == (0.0, 0)
        if precision == G.MIXED:
            host = [a.astype(np.float32).astype(np.float64) for a in host]
        for n, (a, b) in enumerate(zip(got, host)):
            if trig and n == 2:
                assert np.abs(a - b).max() <= 4e-16 * np.abs(b).max() + (6e-8 * np.abs(b).max() if precision else 0)
            else:
                assert np.array_equal(a, b), (k, n)


@pytest.mark.parametrize("tc", [1, 3, 5, 12, 14, 100])
def test_device_ic_onefluid_matches_host(tc):
    info = G.SimInfo(80, 60, 2, 0.0, 0.0, 1.0 / 80 * (2 if tc == 100 else 1), 1.0 / 60 * (2 if tc == 100 else 1),
                     T, T, T, T)
    eos = G.Eos(1.4, 0.0, 1.4, 0.0)
    ctx = G.Context(G.methods(sim_type=G.SIM_ONEFLUID, fs=G.FS_MUSCL))
    ctx.add(info, eos, 0.8)
    ctx.commit()
    kind = G.IC_EXPLOSION if tc == 100 else G.IC_TTC
    ctx.ic_device([(0, G.ic_desc(kind, test_case=tc))])
    assert np.array_equal(ctx.download(0), G.ic_onefluid(info, eos, tc))


def test_device_ic_refuses_bad_requests():
    info, eos, desc = P.shocked_bubble_ic(64, 20)
    ctx = G.Context(G.methods(sim_type=G.SIM_TWOFLUID))
    ctx.add(info, eos, 0.6)
    ctx.commit()
    with pytest.raises(G.GfmError, match="kind"):
        ctx.ic_device([(0, G.ic_desc(G.IC_TTC, test_case=1))])
    with pytest.raises(G.GfmError, match="twice"):
        ctx.ic_device([(0, desc), (0, desc)])
    with pytest.raises(G.GfmError, match="out of range"):
        ctx.ic_device([(3, desc)])
