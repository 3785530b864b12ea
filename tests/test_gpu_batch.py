"""Batched persistent advect + marching cubes (vr_batch) vs the per-grid path
and vs the oracle, in all three modes (loop, eager, graph replay)."""
import numpy as np
import pytest

from util import assert_nano_equal, white_noise_block

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[gpu] {msg}", flush=True)


def scenes(vr, rng, count, vs):
    out = []
    for _ in range(count):
        ns = int(rng.integers(1, 5))
        prims = []
        for _ in range(ns):
            c = rng.uniform(0.2, 0.8, 3)
            r = float(np.exp(rng.uniform(np.log(0.03), np.log(0.15))))
            prims.append((vr.PRIM_SPHERE, vr.CSG_UNION, (float(c[0]), float(c[1]), float(c[2]), r)))
        vel = tuple(float(x) for x in rng.uniform(-1, 1, 3))
        out.append((prims, vel))
    return out


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_batch_matches_oracle(orc, ctx, vr, mode):
    announce(f"batch mode={mode}")
    rng = np.random.default_rng(21)
    vs, hw = 1.0 / 128, 3.0
    spec = scenes(vr, rng, 5, vs)
    grids, vels, dts = [], [], []
    for prims, vel in spec:
        grids.append(ctx.from_primitives(prims, vs, hw))
        vels.append(ctx.velocity_analytic(vr.VEL_UNIFORM, vel))
        dts.append(0.5 * vs / max(abs(v) for v in vel))
    dt = min(dts)
    b = vr.Batch(ctx, grids, vels)
    if mode == vr.BATCH_GRAPH:
        b.autotune(dt, vr.RK3)
        b.step(dt, 0.0, vr.RK3, 0.0, vr.BATCH_GRAPH)  # first call sizes and captures
    b.step(dt, 0.0, vr.RK3, 0.0, mode)
    for i, (prims, vel) in enumerate(spec):
        og = orc.from_nano(grids[i].download())
        want = orc.advect_analytic(og, vr.VEL_UNIFORM, vel, dt, 3)
        assert_nano_equal(b.output_nano(i), want.to_nano(), f"grid {i}")
        vw, tw, _ = orc.mesh(want, 0.0)
        vg, tg = b.mesh(i, mode)
        assert tg.shape == tw.shape and np.array_equal(tg, tw), f"grid {i} triangles"
        assert np.array_equal(vg.view(np.uint64), vw.view(np.uint64)), f"grid {i} vertices"


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_batch_with_empty_member(orc, ctx, vr, mode):
    """A grid without leaves first, between and after non-empty ones: its
    output and mesh are empty, and the others still equal the oracle."""
    announce(f"batch empty member mode={mode}")
    rng = np.random.default_rng(8)
    vs, hw = 1.0 / 128, 3.0
    spec = scenes(vr, rng, 2, vs)
    empty = orc.new_level_set(vs, hw).to_nano()
    order = [None, spec[0], None, spec[1], None]
    grids, vels = [], []
    for item in order:
        grids.append(ctx.upload(empty) if item is None else ctx.from_primitives(item[0], vs, hw))
        vels.append(ctx.velocity_analytic(vr.VEL_UNIFORM, (0.5, -0.25, 0.75) if item is None else item[1]))
    dt = 0.3 * vs
    b = vr.Batch(ctx, grids, vels)
    if mode == vr.BATCH_GRAPH:
        b.step(dt, 0.0, vr.RK3, 0.0, vr.BATCH_GRAPH)  # first call sizes and captures
    b.step(dt, 0.0, vr.RK3, 0.0, mode)
    for i, item in enumerate(order):
        vg, tg = b.mesh(i, mode)
        if item is None:
            assert_nano_equal(b.output_nano(i), empty, f"empty grid {i}")
            assert len(vg) == 0 and len(tg) == 0, f"empty grid {i} mesh"
            continue
        og = orc.from_nano(grids[i].download())
        want = orc.advect_analytic(og, vr.VEL_UNIFORM, item[1], dt, 3)
        assert_nano_equal(b.output_nano(i), want.to_nano(), f"grid {i}")
        vw, tw, _ = orc.mesh(want, 0.0)
        assert tg.shape == tw.shape and np.array_equal(tg, tw), f"grid {i} triangles"
        assert np.array_equal(vg.view(np.uint64), vw.view(np.uint64)), f"grid {i} vertices"


def test_batch_graph_replays_are_stable(orc, ctx, vr):
    announce("batch graph replay")
    rng = np.random.default_rng(5)
    vs = 1.0 / 128
    spec = scenes(vr, rng, 3, vs)
    grids = [ctx.from_primitives(p, vs, 3.0) for p, _ in spec]
    vels = [ctx.velocity_analytic(vr.VEL_UNIFORM, v) for _, v in spec]
    b = vr.Batch(ctx, grids, vels)
    dt = 0.3 * vs
    b.step(dt, 0.0, vr.RK3, 0.0, vr.BATCH_EAGER)
    ref = [(b.output_nano(i), b.mesh(i, vr.BATCH_EAGER)) for i in range(3)]
    for rep in range(3):
        b.step(dt, 0.0, vr.RK3, 0.0, vr.BATCH_GRAPH)
        ctx.synchronize()
        for i in range(3):
            assert b.output_nano(i) == ref[i][0], f"replay {rep} grid {i}"
            v, t = b.mesh(i, vr.BATCH_GRAPH)
            assert np.array_equal(t, ref[i][1][1]) and np.array_equal(v, ref[i][1][0]), f"replay {rep} mesh {i}"


def test_batch_graph_new_parameters_resize(orc, ctx, vr):
    """Graph steps whose new dt or iso gives a larger mesh than the first
    capture: the batch sizes its mesh again for each parameter set."""
    announce("batch graph new parameters")
    rng = np.random.default_rng(13)
    vs = 1.0 / 128
    spec = scenes(vr, rng, 3, vs)
    grids = [ctx.from_primitives(p, vs, 3.0) for p, _ in spec]
    vels = [ctx.velocity_analytic(vr.VEL_UNIFORM, v) for _, v in spec]
    b = vr.Batch(ctx, grids, vels)
    for dt, iso in [(0.3 * vs, 0.0), (0.3 * vs, 2.0 * vs), (0.6 * vs, 2.0 * vs), (0.3 * vs, 0.0)]:
        b.step(dt, 0.0, vr.RK3, iso, vr.BATCH_GRAPH)
        ctx.synchronize()
        for i, (_, vel) in enumerate(spec):
            og = orc.from_nano(grids[i].download())
            want = orc.advect_analytic(og, vr.VEL_UNIFORM, vel, dt, 3)
            vw, tw, _ = orc.mesh(want, iso)
            vg, tg = b.mesh(i, vr.BATCH_GRAPH)
            assert tg.shape == tw.shape and np.array_equal(tg, tw), f"dt {dt} iso {iso} grid {i} triangles"
            assert np.array_equal(vg.view(np.uint64), vw.view(np.uint64)), f"dt {dt} iso {iso} grid {i} vertices"


@pytest.mark.parametrize("kinds,tuning", [
    ("mixed", (1, 0)),      # per-leaf velocity switch, one block per chunk
    ("mixed", (1, 8)),      # persistent, one-leaf chunks
    ("mixed", (3, 8)),      # persistent, chunk size not dividing the grids
    ("vortex", (16, 4)),    # compacted 4-wave leaves, persistent
    ("rigid", (2, 0)),      # slab leaves, one block per chunk
])
def test_batch_velocity_kinds_and_tunings(orc, ctx, vr, kinds, tuning):
    """Every launch form of the batched advect (specialised or mixed kinds,
    persistent or one block per chunk) against the oracle, with per-grid dt."""
    announce(f"batch kinds={kinds} tuning={tuning}")
    rng = np.random.default_rng(8)
    vs, hw = 1.0 / 128, 3.0
    spec = scenes(vr, rng, 4, vs)
    fields = []
    for i, (_, vel) in enumerate(spec):
        k = {"mixed": (vr.VEL_UNIFORM, vr.VEL_RIGID, vr.VEL_VORTEX)[i % 3], "vortex": vr.VEL_VORTEX,
             "rigid": vr.VEL_RIGID}[kinds]
        params = {vr.VEL_UNIFORM: vel, vr.VEL_RIGID: (0.0, 0.0, 2.0, 0.5, 0.5, 0.5), vr.VEL_VORTEX: ()}[k]
        fields.append((k, params))
    grids = [ctx.from_primitives(p, vs, hw) for p, _ in spec]
    vels = [ctx.velocity_analytic(k, p) for k, p in fields]
    dts = [(0.3 + 0.1 * i) * vs for i in range(len(spec))]
    b = vr.Batch(ctx, grids, vels)
    b.set_dt(dts)
    b.set_tuning(*tuning)
    for mode in (vr.BATCH_EAGER, vr.BATCH_GRAPH):
        b.step(0.0, 0.0, vr.RK3, 0.0, mode)
        for i, (k, p) in enumerate(fields):
            og = orc.from_nano(grids[i].download())
            want = orc.advect_analytic(og, k, p, dts[i], 3)
            assert_nano_equal(b.output_nano(i), want.to_nano(), f"mode {mode} grid {i}")
            vw, tw, _ = orc.mesh(want, 0.0)
            vg, tg = b.mesh(i, mode)
            assert tg.shape == tw.shape and np.array_equal(tg, tw), f"mode {mode} grid {i} triangles"
            assert np.array_equal(vg.view(np.uint64), vw.view(np.uint64)), f"mode {mode} grid {i} vertices"


@pytest.mark.parametrize("lo,hi", [(0.3, 0.7), (-0.2, 0.2)])
def test_batch_mesh_large_tile_list(orc, ctx, vr, lo, hi):
    """A batch whose tile list (at least its leaf count) takes the
    tile-per-lane neighbour kernel (32768 tiles and up), while each grid
    alone (at most 7 x its leaves) takes the per-slot one: every grid's mesh
    from the batch equals its own mesh, and grid 0's equals the oracle's.
    Centres in [lo, hi]^3: the second set puts the spheres across the origin
    (negative coordinates, roots on both sides of 0)."""
    announce(f"batch mesh large tile list {lo} {hi}")
    rng = np.random.default_rng(31)
    vs, hw = 1.0 / 256, 3.0
    spec = []
    for _ in range(40):
        c = rng.uniform(lo, hi, 3)
        r = float(rng.uniform(0.15, 0.3))
        spec.append([(vr.PRIM_SPHERE, vr.CSG_UNION, (float(c[0]), float(c[1]), float(c[2]), r))])
    grids = [ctx.from_primitives(p, vs, hw) for p in spec]
    leaves = sum(g.counts()[0] for g in grids)
    assert leaves >= 32768 and all(7 * g.counts()[0] < 32768 for g in grids), leaves
    vels = [ctx.velocity_analytic(vr.VEL_UNIFORM, (0.0, 0.0, 0.0)) for _ in grids]
    b = vr.Batch(ctx, grids, vels)
    b.step(0.25 * vs, 0.0, vr.RK1, 0.0, vr.BATCH_EAGER)
    for i in range(len(grids)):
        vg, tg = b.mesh(i, vr.BATCH_EAGER)
        vs1, ts1 = ctx.upload(b.output_nano(i)).mesh(0.0)
        assert np.array_equal(tg, ts1), f"grid {i} triangles"
        assert np.array_equal(vg.view(np.uint64), vs1.view(np.uint64)), f"grid {i} vertices"
        if i == 0:
            vw, tw, _ = orc.mesh(orc.from_nano(b.output_nano(0)), 0.0)
            assert np.array_equal(tg, tw) and np.array_equal(vg.view(np.uint64), vw.view(np.uint64)), "oracle"


def check_against_oracle(orc, b, inputs, spec, dts, iso, mode, what):
    """Each output grid and mesh of the batch's last `mode` step equals the
    oracle's RK3 advect of the input bytes with that grid's dt, then MC at iso."""
    for i, (_, vel) in enumerate(spec):
        want = orc.advect_analytic(orc.from_nano(inputs[i]), 2, vel, dts[i], 3)
        assert_nano_equal(b.output_nano(i), want.to_nano(), f"{what}: grid {i}")
        vw, tw, _ = orc.mesh(want, iso)
        vg, tg = b.mesh(i, mode)
        assert tg.shape == tw.shape and np.array_equal(tg, tw), f"{what}: grid {i} triangles"
        assert np.array_equal(vg.view(np.uint64), vw.view(np.uint64)), f"{what}: grid {i} vertices"


def small_batch(ctx, vr, seed, count=4, vs=1.0 / 128):
    spec = scenes(vr, np.random.default_rng(seed), count, vs)
    grids = [ctx.from_primitives(p, vs, 3.0) for p, _ in spec]
    vels = [ctx.velocity_analytic(vr.VEL_UNIFORM, v) for _, v in spec]
    return spec, grids, vels, [g.download() for g in grids]


def test_batch_graph_after_autotune(orc, ctx, vr):
    """GRAPH at the default tuning, then the autotuner (whose one-leaf pass
    grows the chunk table), then GRAPH again: the step never replays a graph
    that reads the old chunk table."""
    announce("batch graph after autotune")
    vs = 1.0 / 128
    spec, grids, vels, inputs = small_batch(ctx, vr, 31)
    b = vr.Batch(ctx, grids, vels)
    dt = 0.3 * vs
    b.step(dt, 0.0, vr.RK3, 0.0, vr.BATCH_GRAPH)
    b.autotune(dt, vr.RK3)
    for tuning in [None, (4, 8), (1, 0), (4, 8)]:
        if tuning:
            b.set_tuning(*tuning)
        b.step(dt, 0.0, vr.RK3, 0.0, vr.BATCH_GRAPH)
        check_against_oracle(orc, b, inputs, spec, [dt] * len(spec), 0.0, vr.BATCH_GRAPH, f"tuning {tuning}")


def test_batch_graph_eager_graph_with_larger_mesh(orc, ctx, vr):
    """GRAPH(a), EAGER(b) with a larger mesh (reallocated buffers), GRAPH(a):
    the last step captures again instead of writing the freed buffers, and
    its mesh is whole."""
    announce("batch graph/eager/graph")
    vs = 1.0 / 128
    spec, grids, vels, inputs = small_batch(ctx, vr, 37)
    b = vr.Batch(ctx, grids, vels)
    a = (0.3 * vs, 0.0)
    for dt, iso, mode in [(*a, vr.BATCH_GRAPH), (0.6 * vs, 2.0 * vs, vr.BATCH_EAGER), (*a, vr.BATCH_GRAPH),
                          (0.6 * vs, 2.0 * vs, vr.BATCH_GRAPH), (*a, vr.BATCH_EAGER), (*a, vr.BATCH_GRAPH)]:
        b.step(dt, 0.0, vr.RK3, iso, mode)
        check_against_oracle(orc, b, inputs, spec, [dt] * len(spec), iso, mode, f"dt={dt} iso={iso} mode={mode}")


def test_batch_per_grid_dts_then_uniform(orc, ctx, vr):
    """Per-grid dts through a graph, then set_dt(None): the next graph step
    uses the step's dt for every grid, with a mesh sized for it."""
    announce("batch per-grid dts then uniform")
    vs = 1.0 / 128
    spec, grids, vels, inputs = small_batch(ctx, vr, 41)
    b = vr.Batch(ctx, grids, vels)
    dts = [0.5 * vs / max(abs(c) for c in v) for _, v in spec]
    b.set_dt(dts)
    b.step(0.0, 0.0, vr.RK3, 0.0, vr.BATCH_GRAPH)
    check_against_oracle(orc, b, inputs, spec, dts, 0.0, vr.BATCH_GRAPH, "per-grid dts")
    b.set_dt(None)
    b.step(0.2 * vs, 0.0, vr.RK3, 0.0, vr.BATCH_GRAPH)
    check_against_oracle(orc, b, inputs, spec, [0.2 * vs] * len(spec), 0.0, vr.BATCH_GRAPH, "uniform dt")


def test_batch_inputs_changed_after_create(orc, ctx, vr):
    """The batch keeps its own copy of the inputs: rebuilding, trimming,
    rolling out or destroying an input grid afterwards does not change the
    batch's results."""
    announce("batch inputs changed after create")
    vs = 1.0 / 128
    spec, grids, vels, inputs = small_batch(ctx, vr, 43)
    b = vr.Batch(ctx, grids, vels)
    dt = 0.3 * vs
    grids[0].rebuild(1.5)
    grids[1].trim(1.0)
    grids[2].rollout(vels[2], 0.0, dt, vr.RK1, 3, 2, 3.0, 0, 0.0, None)
    grids[3].close()
    for mode in [vr.BATCH_EAGER, vr.BATCH_GRAPH, vr.BATCH_LOOP]:
        b.step(dt, 0.0, vr.RK3, 0.0, mode)
        check_against_oracle(orc, b, inputs, spec, [dt] * len(spec), 0.0, mode, f"mode {mode}")


@pytest.mark.parametrize("mode", ["eager", "graph"])
def test_batch_white_noise_meshes(orc, ctx, vr, mode):
    """The batch's sign-mask marching cubes on random values (zero velocity:
    the advected grid keeps them): dense rounds and tiles whose earlier-tile
    edges overflow the side table, bit-exact against the oracle."""
    announce(f"batch white noise {mode}")
    og = white_noise_block(orc, 11)
    inputs = [og.to_nano()]
    spec = [(None, (0.0, 0.0, 0.0))]
    g = ctx.upload(inputs[0])
    b = vr.Batch(ctx, [g], [ctx.velocity_analytic(vr.VEL_UNIFORM, (0.0, 0.0, 0.0))])
    m = vr.BATCH_EAGER if mode == "eager" else vr.BATCH_GRAPH
    for _ in range(2):
        b.step(0.25, 0.0, vr.RK3, 0.0, m)
    check_against_oracle(orc, b, inputs, spec, [0.25], 0.0, m, f"white noise {mode}")
