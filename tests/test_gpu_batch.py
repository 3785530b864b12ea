"""Batched persistent advect + marching cubes (vr_batch) vs the per-grid path
and vs the oracle, in all three modes (loop, eager, graph replay)."""
import numpy as np
import pytest

from util import assert_nano_equal

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
