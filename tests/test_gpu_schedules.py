"""The fast schedules against the simple ones, bit for bit.

1. Temporally blocked ghost extrapolation (all Jacobi iterations of a ghost
   fill in one launch, tile and halo in LDS) against one launch per
   iteration, ghost frame included.  Covered: all four GFM variants, FP64 and
   mixed precision, each boundary combination the blocked path takes
   (reflective and transmissive on every side; the right side reflects with
   the left BC as in the reference), partial tiles, several instances per
   context, and full steps.
2. The interfacial solve ordered by dependency level against plain rounds
   (solve whatever is ready, repeat), for the variants with dependencies.
3. Steps of a batch split into instance groups on separate streams (with
   multi-step graphs and staggered group starts) against one stream.
4. The blocked extrapolation and level-set kernels on periodic-in-y
   instances (wrapped region rows) against the per-iteration ones.
The mixed path's ghost band is off in this module (GFM_GHOST_BAND=0): the
band leaves ghost states far from the interface alone, and which of them
depends on the schedule and the tiles (the blocked extrapolation passes whole
tiles through, the one-sweep kernels extrapolate every cell).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import gfm_hip as G
import problems as P

VARIANTS = [G.GFM_ORIGINAL, G.GFM_REAL, G.GFM_RIEMANN, G.GFM_MODIFIED]


@pytest.fixture(autouse=True)
def no_ghost_band(monkeypatch):
    monkeypatch.setenv("GFM_GHOST_BAND", "0")
T, R = G.BC_TRANSMISSIVE, G.BC_REFLECTIVE
BCS = [(T, T, R, R), (R, R, T, T), (R, T, R, T), (T, R, T, R)]


def make_ctx(cases, variant, precision, blocked, plain_rounds=False, groups=4, fs_tile="32x16", tb_tile="64x16",
             persistent=False, iface_lds=None):
    env = {"GFM_NO_TEMPORAL_BLOCKING": "0" if blocked else "1", "GFM_IFACE_PLAIN_ROUNDS": "1" if plain_rounds else "0",
           "GFM_STEP_GROUPS": str(groups), "GFM_FS_TILE": fs_tile, "GFM_TB_TILE": tb_tile,
           "GFM_FS_PERSISTENT": "1" if persistent else "0",
           "GFM_IFACE_LDS_ENTRIES": "" if iface_lds is None else str(iface_lds)}
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        ctx = G.Context(G.methods(sim_type=G.SIM_TWOFLUID, fs=G.FS_MUSCL, gfm=variant, ls_order=5,
                                  reinit_every=5, reinit_iters=10, precision=precision))
        for info, eos, _ in cases:
            ctx.add(info, eos, 0.6, 1e300, 1 << 30)
        ctx.commit()
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v
    for k, (_, _, st) in enumerate(cases):
        ctx.upload(k, *st)
    return ctx


def perturbed_bubble(Nx, Ny, bcs, seed, numGC=6, **kw):
    """shocked bubble with every cell made distinct (conserved variables
    scaled by 1 + 1e-3 U[0,1)) so that no stencil reads equal neighbours"""
    info, eos, (g1, g2, phi) = P.shocked_bubble(Nx, Ny, numGC=numGC, **kw)
    info.BC_L, info.BC_R, info.BC_B, info.BC_T = bcs
    rng = np.random.default_rng(seed)
    g1 = g1 * (1.0 + 1e-3 * rng.random(g1.shape))
    g2 = g2 * (1.0 + 1e-3 * rng.random(g2.shape))
    return info, eos, (np.ascontiguousarray(g1), np.ascontiguousarray(g2), phi)


def assert_same(a, b, what):
    assert a.shape == b.shape
    bad = ~((a == b) | (np.isnan(a) & np.isnan(b)))
    assert not bad.any(), "%s: %d cells differ, max |d| %g" % (what, bad.sum(), np.nanmax(np.abs(a - b)))


def compare(ca, cb, n):
    for k in range(n):
        a, b = ca.download(k), cb.download(k)
        for x, y, name in zip(a, b, ("fluid1", "fluid2", "phi")):
            assert_same(x, y, "instance %d %s" % (k, name))
        assert_same(ca.download_vfield(k), cb.download_vfield(k), "instance %d vfield" % k)
        assert ca.bbox(k) == cb.bbox(k)
        assert ca.status(k) == cb.status(k)


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED])
@pytest.mark.parametrize("bcs", BCS)
@pytest.mark.parametrize("variant", VARIANTS)
def test_ghost_fill_blocked_matches_per_iteration(variant, bcs, precision):
    print("CASE ghost_fill variant=%d bcs=%s precision=%d" % (variant, bcs, precision), flush=True)
    cases = [perturbed_bubble(150, 37, bcs, 1), perturbed_bubble(70, 20, bcs, 2, numGC=7)]
    ca = make_ctx(cases, variant, precision, True)
    cb = make_ctx(cases, variant, precision, False)
    ca.set_ghost_states()
    cb.set_ghost_states()
    ca.sync()
    cb.sync()
    compare(ca, cb, len(cases))


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED])
@pytest.mark.parametrize("variant", VARIANTS)
def test_steps_blocked_matches_per_iteration(variant, precision):
    print("CASE steps variant=%d precision=%d" % (variant, precision), flush=True)
    cases = [perturbed_bubble(200, 56, BCS[0], 3), perturbed_bubble(96, 40, BCS[2], 4, numGC=8)]
    ca = make_ctx(cases, variant, precision, True)
    cb = make_ctx(cases, variant, precision, False)
    ca.step_graph(12)
    cb.step_graph(12)
    ca.sync()
    cb.sync()
    compare(ca, cb, len(cases))
    for k in range(len(cases)):
        assert ca.get_time(k) == cb.get_time(k)


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED])
@pytest.mark.parametrize("variant", [G.GFM_REAL, G.GFM_RIEMANN])
def test_iface_levels_match_plain_rounds(variant, precision):
    print("CASE iface levels variant=%d precision=%d" % (variant, precision), flush=True)
    cases = [perturbed_bubble(300, 84, BCS[0], 5, R=0.03), perturbed_bubble(128, 40, BCS[1], 6)]
    ca = make_ctx(cases, variant, precision, True)
    cb = make_ctx(cases, variant, precision, True, plain_rounds=True)
    ca.step_graph(8)
    cb.step_graph(8)
    ca.sync()
    cb.sync()
    compare(ca, cb, len(cases))
    for k in range(len(cases)):
        assert ca.iface_info(k)[0] > 0


@pytest.mark.parametrize("plain", [False, True])
@pytest.mark.parametrize("precision", [G.FP64, G.MIXED])
@pytest.mark.parametrize("variant", [G.GFM_REAL, G.GFM_RIEMANN, G.GFM_VOFGFM])
def test_iface_flags_in_global_memory_match_lds(variant, precision, plain):
    """an interfacial list longer than the ordered solve's LDS flag array
    keeps its flags in global memory (here every list: the array is cut to
    16 entries); both schedules give the bits of the LDS form"""
    print("CASE iface global flags variant=%d precision=%d plain=%d" % (variant, precision, plain), flush=True)
    cases = [perturbed_bubble(300, 84, BCS[0], 5, R=0.03), perturbed_bubble(128, 40, BCS[1], 6)]
    ca = make_ctx(cases, variant, precision, True)
    cb = make_ctx(cases, variant, precision, True, plain_rounds=plain, iface_lds=16)
    ca.step_graph(8)
    cb.step_graph(8)
    ca.sync()
    cb.sync()
    compare(ca, cb, len(cases))
    for k in range(len(cases)):
        assert ca.iface_info(k)[0] > 16


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED])
@pytest.mark.parametrize("iters", [1, 7, 23])
def test_reinit_blocked_matches_per_iteration(iters, precision):
    """explicit reinitialisation of a distorted level set (1, 7 and 23
    iterations: one partial block, one block, three blocks)"""
    print("CASE reinit iters=%d precision=%d" % (iters, precision), flush=True)
    cases = []
    for k, bcs in enumerate(BCS[:2]):
        info, eos, (g1, g2, phi) = perturbed_bubble(150 + 17 * k, 44 + 5 * k, bcs, 7 + k)
        H, W = phi.shape
        yy, xx = np.mgrid[0:H, 0:W]
        phi = np.ascontiguousarray(phi * (1.0 + 0.3 * np.sin(xx / 9.0) * np.cos(yy / 7.0)))
        cases.append((info, eos, (g1, g2, phi)))
    ca = make_ctx(cases, G.GFM_RIEMANN, precision, True)
    cb = make_ctx(cases, G.GFM_RIEMANN, precision, False)
    for c in (ca, cb):
        c.ls_reinit(iters)
        c.sync()
    for k in range(len(cases)):
        assert_same(ca.download(k)[2], cb.download(k)[2], "instance %d phi" % k)


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED])
@pytest.mark.parametrize("variant", VARIANTS)
def test_periodic_y_ghost_fill_blocked_matches_per_iteration(variant, precision):
    """blocked ghost extrapolation of periodic-in-y instances (wrapped region
    rows) straight after the upload: the perturbed ghost rows are not periodic
    images, so the first iteration must read them as stored"""
    print("CASE periodic-y ghost_fill variant=%d precision=%d" % (variant, precision), flush=True)
    P_ = G.BC_PERIODIC
    cases = [perturbed_bubble(150, 37, (T, T, P_, P_), 1), perturbed_bubble(70, 12, (R, T, P_, P_), 2, numGC=7),
             perturbed_bubble(96, 40, BCS[1], 3)]
    ca = make_ctx(cases, variant, precision, True)
    cb = make_ctx(cases, variant, precision, False)
    ca.set_ghost_states()
    cb.set_ghost_states()
    ca.sync()
    cb.sync()
    compare(ca, cb, len(cases))


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED])
@pytest.mark.parametrize("variant", VARIANTS)
def test_periodic_y_steps_blocked_matches_per_iteration(variant, precision):
    """periodic-in-y instances (the blocked extrapolation and level-set
    kernels wrap their region rows) beside a mirror-bounded one: 12 steps
    with reinitialisation every 5, then an explicit 23-iteration
    reinitialisation.  Ny = 12 makes one region wrap several times."""
    print("CASE periodic-y level set variant=%d precision=%d" % (variant, precision), flush=True)
    P_ = G.BC_PERIODIC
    cases = [perturbed_bubble(150, 37, (T, T, P_, P_), 31), perturbed_bubble(96, 40, BCS[0], 32, numGC=8),
             perturbed_bubble(80, 12, (T, R, P_, P_), 33)]
    ca = make_ctx(cases, variant, precision, True)
    cb = make_ctx(cases, variant, precision, False)
    for c in (ca, cb):
        c.step_graph(12)
        c.sync()
    compare(ca, cb, len(cases))
    for c in (ca, cb):
        c.ls_reinit(23)
        c.sync()
    for k in range(len(cases)):
        assert_same(ca.download(k)[2], cb.download(k)[2], "instance %d phi after reinit" % k)
        assert ca.get_time(k) == cb.get_time(k)


@pytest.mark.parametrize("variant", VARIANTS)
def test_instance_groups_match_one_stream(variant):
    """seven instances of different sizes in four groups: 105 steps (a
    staggered first step, one 100-step graph and four one-step graphs per
    group), then 7 eager steps"""
    print("CASE groups variant=%d" % variant, flush=True)
    sizes = [(96, 40), (64, 24), (130, 33), (72, 72), (40, 20), (110, 28), (64, 40)]
    cases = [perturbed_bubble(nx, ny, BCS[k % len(BCS)], 20 + k) for k, (nx, ny) in enumerate(sizes)]
    ca = make_ctx(cases, variant, G.MIXED, True, groups=4)
    cb = make_ctx(cases, variant, G.MIXED, True, groups=1)
    for c in (ca, cb):
        c.step_graph(105)
        c.step(7)
        c.sync()
    compare(ca, cb, len(cases))
    for k in range(len(cases)):
        assert ca.get_time(k) == cb.get_time(k)
        assert ca.get_time(k)[1] == 112


def test_autotune_small_batch():
    """the schedule and tile autotuner (autotune.py) on a real batch: eleven trials (four group counts
    with the blocked sweeps, the one-sweep schedule at the best count, then the three other flux tiles,
    with the blocked schedule the two other extrapolation tiles, and the persistent flux kernel), and a
    context on the chosen settings steps the batch to the same bits as one stream"""
    print("CASE autotune", flush=True)
    import autotune as AT
    sizes = [(96, 40), (64, 24), (130, 33), (72, 72), (40, 20)]
    cases = [perturbed_bubble(nx, ny, BCS[k % len(BCS)], 40 + k) for k, (nx, ny) in enumerate(sizes)]

    def build():
        ctx = G.Context(G.methods(sim_type=G.SIM_TWOFLUID, fs=G.FS_MUSCL, gfm=G.GFM_RIEMANN, ls_order=5,
                                  reinit_every=5, reinit_iters=10, precision=G.MIXED))
        for info, eos, _ in cases:
            ctx.add(info, eos, 0.6, 1e300, 1 << 30)
        ctx.commit()
        return ctx

    def load(ctx):
        for k, (_, _, st) in enumerate(cases):
            ctx.upload(k, *st)
            ctx.set_time(k, 0.0, 0)

    before = {k: os.environ.get(k) for k in AT.KEYS}
    choice = AT.tune(build, load, len(cases), steps=6, warmup=2)
    assert {k: os.environ.get(k) for k in AT.KEYS} == before
    assert len(choice["trials"]) == (11 if choice["blocking"] else 9)
    assert {k: choice[k] for k in ("groups", "blocking", "fs_tile", "tb_tile", "persistent", "ms_per_step")} \
        in choice["trials"]
    with AT.applied(choice):
        ca = build()
    load(ca)
    cb = make_ctx(cases, G.GFM_RIEMANN, G.MIXED, True, groups=1)
    for c in (ca, cb):
        c.step_graph(12)
        c.sync()
    compare(ca, cb, len(cases))


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED], ids=["fp64", "mixed"])
def test_tile_shapes_match_default(precision):
    """every compiled flux tile (32x16, 64x8, 16x32, 32x8) and blocked-extrapolation tile (64x16,
    32x16, 64x8) steps a batch of uneven instances (partial tiles, every boundary combination of the
    blocked path) to the same bits as the default tiles: the per-cell and per-face arithmetic and the
    update's accumulation order do not depend on the tile"""
    import autotune as AT
    sizes = [(96, 40), (70, 24), (130, 33), (72, 72)]
    cases = [perturbed_bubble(nx, ny, BCS[k % len(BCS)], 70 + k) for k, (nx, ny) in enumerate(sizes)]
    ref = make_ctx(cases, G.GFM_RIEMANN, precision, True, groups=1)
    ref.step_graph(8)
    ref.sync()
    for fs in AT.FS_TILES[1:]:
        print("CASE tile fs=%s precision=%d" % (fs, precision), flush=True)
        c = make_ctx(cases, G.GFM_RIEMANN, precision, True, groups=1, fs_tile=fs)
        c.step_graph(8)
        c.sync()
        compare(c, ref, len(cases))
        c.close()
    for tb in AT.TB_TILES[1:]:
        print("CASE tile tb=%s precision=%d" % (tb, precision), flush=True)
        c = make_ctx(cases, G.GFM_RIEMANN, precision, True, groups=1, tb_tile=tb)
        c.step_graph(8)
        c.sync()
        compare(c, ref, len(cases))
        c.close()
    ref.close()


def test_unknown_tile_shape_is_an_error():
    """a tile shape the library was not compiled for is refused at commit, naming the shapes it has"""
    print("CASE tile error", flush=True)
    cases = [perturbed_bubble(40, 20, BCS[0], 3)]
    with pytest.raises(G.GfmError, match="32x16, 64x8, 16x32, 32x8"):
        make_ctx(cases, G.GFM_RIEMANN, G.MIXED, True, fs_tile="48x16")
    with pytest.raises(G.GfmError, match="64x16, 32x16, 64x8"):
        make_ctx(cases, G.GFM_RIEMANN, G.MIXED, True, tb_tile="7x7")


def test_round_state_f32_switch():
    """GFM_ROUND_STATE_F32=1 (FP64 path, analysis): after every step both fluids' conserved grids and
    phi hold FP32 values; the first step matches the plain FP64 step rounded to FP32, and later steps
    differ from the plain FP64 run by no more than FP32 rounding carried a few steps"""
    print("CASE round f32", flush=True)
    cases = [perturbed_bubble(96, 40, BCS[0], 11), perturbed_bubble(70, 24, BCS[1], 12)]
    old = os.environ.get("GFM_ROUND_STATE_F32")
    os.environ["GFM_ROUND_STATE_F32"] = "1"
    try:
        cr = make_ctx(cases, G.GFM_RIEMANN, G.FP64, True, groups=1)
    finally:
        if old is None:
            del os.environ["GFM_ROUND_STATE_F32"]
        else:
            os.environ["GFM_ROUND_STATE_F32"] = old
    cp = make_ctx(cases, G.GFM_RIEMANN, G.FP64, True, groups=1)
    for c in (cr, cp):
        c.step(1)
        c.sync()
    for k in range(len(cases)):
        for a, b in zip(cr.download(k), cp.download(k)):
            assert np.array_equal(a, b.astype(np.float32).astype(np.float64))
    for c in (cr, cp):
        c.step_graph(9)
        c.sync()
    for k in range(len(cases)):
        for a, b in zip(cr.download(k), cp.download(k)):
            assert np.array_equal(a, a.astype(np.float32).astype(np.float64))
            scale = np.abs(b).max()
            assert np.abs(a - b).max() <= 1e-5 * scale
        assert cr.status(k) == 0
    cr.close()
    cp.close()


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("groups", [1, 2])
def test_persistent_flux_matches_per_tile(precision, groups, monkeypatch):
    """the persistent flux kernel (GFM_FS_PERSISTENT=1: as many workgroups as the GPU holds, draining
    the tiles through eight per-XCD queues) against one workgroup per tile, bit for bit, on a batch of
    uneven instances in one or two instance groups (each group its own queue counters), with the
    default tile and the 16x32 tile, and in mixed precision with the ghost band on"""
    sizes = [(96, 40), (70, 24), (130, 33), (72, 72), (200, 64)]
    cases = [perturbed_bubble(nx, ny, BCS[k % len(BCS)], 90 + k) for k, (nx, ny) in enumerate(sizes)]
    if precision == G.MIXED:
        monkeypatch.setenv("GFM_GHOST_BAND", "4")
    for fs in ("32x16", "16x32"):
        print("CASE persistent flux precision=%d groups=%d tile=%s" % (precision, groups, fs), flush=True)
        ref = make_ctx(cases, G.GFM_RIEMANN, precision, True, groups=groups, fs_tile=fs)
        per = make_ctx(cases, G.GFM_RIEMANN, precision, True, groups=groups, fs_tile=fs, persistent=True)
        for c in (ref, per):
            c.step_graph(10)
            c.sync()
        compare(per, ref, len(cases))
        ref.close()
        per.close()
