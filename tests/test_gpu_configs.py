"""The benchmark configurations under test, on the GPU against the oracle:

- the Zalesak rollout at full size (1/512, 100 RK3 steps, rebuild every 5,
  marching cubes every 10), final grid bytes and meshes;
- grids of the batched 64-sphere-union scene at its real resolution (1/1024)
  through a graph-replayed batch, and the whole 64-grid batch against the
  per-grid path;
- the differentiable rollout at full size (1/256 sphere, 64^3 field, 20 RK1
  steps): the gradient against the FP64 adjoint and against FP64 central
  differences (h = 1e-3) in 32 entries;
- the noisy-sphere RK3 adjoint with a checkpoint every 2 steps (the large
  batched differentiable workload's schedule) at reduced size.
"""
import importlib.util
import multiprocessing as mp
import os

import numpy as np
import pytest

from conftest import ROOT
from util import active_voxels, assert_nano_equal, config4_fd64, config4_problem, mesh_equal, slotted_sphere

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[gpu] {msg}", flush=True)


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


# ---------------------------------------------------------------- Zalesak rollout, full size
def oracle_rollout_bounded(orc, g, period, dt, steps, rebuild_every, hw, mesh_every):
    """advect (RK3, single vortex) -> rebuild every k (the bounded-radius
    rebuild, equal to the reference's brute force bit for bit) -> marching
    cubes every m; returns the final grid, the voxel updates and every mesh."""
    updates, meshes = 0, []
    for s in range(steps):
        tf = orc.stage_factors(period, s * dt, dt, 3)
        updates += g.active_count()
        g = orc.advect_analytic(g, 4, (), dt, 3, tf)
        if rebuild_every and (s + 1) % rebuild_every == 0:
            orc.rebuild_bounded(g, hw)
        if mesh_every and (s + 1) % mesh_every == 0:
            meshes.append(orc.mesh(g, 0.0))
    return g, updates, meshes


@pytest.mark.parametrize("steps", [10, 100])
def test_zalesak_rollout_full_size(orc, ctx, vr, steps):
    """Slotted sphere at 1/512, single vortex with cos(pi t / 8), RK3 at CFL
    0.5 (dt = 0.5 vs), rebuild every 5, marching cubes every 10: the GPU
    rollout's final NanoGrid bytes, counts and last mesh equal the oracle's
    composition of the reference tools (level_set_track.rs:80-114,
    volume_to_mesh.rs:46-201)."""
    announce(f"zalesak 1/512 rollout steps={steps}")
    vs, hw, period = 1.0 / 512, 3.0, 8.0
    dt = 0.5 * vs
    og = slotted_sphere(orc, vs)
    g = ctx.upload(og.to_nano())
    want, updates, meshes = oracle_rollout_bounded(orc, og, period, dt, steps, 5, hw, 10)
    vel = ctx.velocity_analytic(vr.VEL_VORTEX, (), period)
    mesh = vr.Mesh(ctx)
    st = g.rollout(vel, 0.0, dt, vr.RK3, steps, 5, hw, 10, 0.0, mesh)
    assert_nano_equal(g.download(), want.to_nano(), f"zalesak {steps} steps")
    assert st.voxel_updates == updates
    assert (st.final_leaf_count, st.final_active_count) == (want.leaf_count(), want.active_count())
    assert st.meshes == len(meshes) == steps // 10
    vw, tw, _ = meshes[-1]
    vg, tg = mesh.arrays()
    mesh_equal(vg, tg, vw, tw, vs, f"zalesak last mesh ({steps} steps)")
    assert np.array_equal(vg.view(np.uint64), vw.view(np.uint64))
    if steps == 10:
        assert len(tw) > 1000  # the mesh at step 10 is not empty


# ---------------------------------------------------------------- batched scene at 1/1024
@pytest.fixture(scope="module")
def sphere_union_scene(ctx, vr):
    bench = load_bench()
    spec = bench.batched_sphere_scenes(vr)
    vs = 1.0 / 1024
    grids = [ctx.from_primitives(p, vs, 3.0) for p, _ in spec]
    vels = [ctx.velocity_analytic(vr.VEL_UNIFORM, v) for _, v in spec]
    dts = [0.5 * vs / max(abs(c) for c in v) for _, v in spec]
    return spec, grids, vels, dts


def test_sphere_union_grids_full_resolution(orc, ctx, vr, sphere_union_scene):
    """The four smallest grids of the 64-grid scene at 1/1024 (about 4e5
    active voxels each) in one batch with their own dts, graph-replayed:
    outputs and meshes equal the oracle's RK3 advect + marching cubes."""
    announce("sphere-union grids at 1/1024")
    spec, grids, vels, dts = sphere_union_scene
    sizes = [g.counts()[1] for g in grids]
    pick = sorted(range(len(grids)), key=lambda i: sizes[i])[:4]
    b = vr.Batch(ctx, [grids[i] for i in pick], [vels[i] for i in pick])
    b.set_dt([dts[i] for i in pick])
    b.autotune(dts[pick[0]], vr.RK3)
    for _ in range(2):
        b.step(0.0, 0.0, vr.RK3, 0.0, vr.BATCH_GRAPH)
    for k, i in enumerate(pick):
        og = orc.from_nano(grids[i].download())
        want = orc.advect_analytic(og, vr.VEL_UNIFORM, spec[i][1], dts[i], 3)
        assert_nano_equal(b.output_nano(k), want.to_nano(), f"scene grid {i}")
        vw, tw, _ = orc.mesh(want, 0.0)
        vg, tg = b.mesh(k, vr.BATCH_GRAPH)
        assert len(tw) > 0 and np.array_equal(tg, tw), f"scene grid {i} triangles"
        assert np.array_equal(vg.view(np.uint64), vw.view(np.uint64)), f"scene grid {i} vertices"
    b.close()


def test_sphere_union_batch_all_64_against_loop(ctx, vr, sphere_union_scene):
    """All 64 grids (6e8 active voxels): the graph-replayed batch gives every
    grid the same output bytes and mesh as the per-grid path (one advect and
    one marching-cubes pipeline per grid)."""
    announce("sphere-union batch of 64 vs per-grid loop")
    spec, grids, vels, dts = sphere_union_scene
    b = vr.Batch(ctx, grids, vels)
    b.set_dt(dts)
    b.step(0.0, 0.0, vr.RK3, 0.0, vr.BATCH_GRAPH)
    b.step(0.0, 0.0, vr.RK3, 0.0, vr.BATCH_GRAPH)
    graph = [(b.output_nano(i), b.mesh_counts(i, vr.BATCH_GRAPH)) for i in range(len(grids))]
    meshes = {i: b.mesh(i, vr.BATCH_GRAPH) for i in (0, 17, 63)}
    b.step(0.0, 0.0, vr.RK3, 0.0, vr.BATCH_LOOP)
    for i in range(len(grids)):
        assert b.output_nano(i) == graph[i][0], f"grid {i} output"
        assert b.mesh_counts(i, vr.BATCH_LOOP) == graph[i][1], f"grid {i} mesh counts"
    for i, (v, t) in meshes.items():
        vl, tl = b.mesh(i, vr.BATCH_LOOP)
        assert np.array_equal(t, tl) and np.array_equal(v.view(np.uint64), vl.view(np.uint64)), f"grid {i} mesh"
    b.close()


# ---------------------------------------------------------------- differentiable rollout, full size
def test_config4_gradient_full_size(orc, ctx, vr):
    """Sphere r = 0.3 at 1/256 (444 524 active voxels), dense 64^3 field of
    N(0, 0.5^2), 20 RK1 steps with a checkpoint at every step, L2 loss against
    the sphere moved by (0.1, 0, 0): the GPU gradient against the FP64 CPU
    adjoint (all 786 432 entries) and against FP64 central differences
    (h = 1e-3) in 32 seeded entries, 24 of them where the gradient is large."""
    announce("config-4 gradient, full size")
    u, g0, tg, dt, n, steps = config4_problem(orc)
    g = ctx.upload(g0.to_nano())
    target = ctx.upload(tg.to_nano())
    vel = ctx.velocity_dense(u, n, 1.0 / n)
    for ckpt in (1, 4):
        st, grad = g.adjoint_rollout(vel, n, dt, vr.RK1, steps, ckpt, target)
        if ckpt == 1:
            grad1 = grad
        else:
            assert np.abs(grad - grad1).max() <= 1e-3 * np.abs(grad1).max()
    loss64, g64 = orc.rollout_adjoint_f64(g0, u.astype(np.float64), n, 1.0 / n, dt, 1, steps, tg)
    g64 = g64.reshape(3, n, n, n)
    scale = np.abs(g64).max()
    err = np.abs(grad1 - g64).max()
    rel = np.linalg.norm(grad1 - g64) / np.linalg.norm(g64)
    print(f"[gpu] config-4 gradient vs FP64 adjoint: max err {err / scale:.2e} of max|g|, rel L2 {rel:.2e}", flush=True)
    assert err <= 2e-3 * scale and rel < 1e-3
    assert abs(st.loss - loss64) <= 1e-4 * loss64
    flat = g64.reshape(-1)
    rng = np.random.default_rng(2024)
    big = np.nonzero(np.abs(flat) > 0.05 * scale)[0]
    idx = np.concatenate([rng.choice(big, 24, replace=False), rng.choice(flat.size, 8, replace=False)])
    with mp.get_context("spawn").Pool(8) as pool:
        fd = np.array(pool.map(config4_fd64, [(int(i), 1e-3) for i in idx]))
    got = grad1.reshape(-1)[idx]
    fd_err = np.abs(fd - got).max()
    print(f"[gpu] config-4 central differences (32 entries): max |fd - grad| {fd_err / scale:.2e} of max|g|, "
          f"max |fd - fp64 adjoint| {np.abs(fd - flat[idx]).max() / scale:.2e}", flush=True)
    assert fd_err <= 3e-3 * scale
    assert np.abs(fd - flat[idx]).max() <= 1e-4 * scale


def test_noisy_sphere_rk3_adjoint_checkpoint_2(orc, ctx, vr):
    """The large differentiable workload's schedule at reduced size: a sphere
    r = 0.35 displaced by 4-octave value noise (amplitude 0.02, base frequency
    8, seed 1000) at 1/128, a dense 32^3 N(0, 0.5^2) field, 10 RK3 steps with a
    checkpoint every 2 steps, loss against the undisplaced sphere: the GPU
    gradient against the FP64 CPU adjoint."""
    announce("noisy sphere rk3 adjoint ckpt 2")
    vs, n, steps = 1.0 / 128, 32, 10
    g0 = orc.scene([(2, (0.5, 0.5, 0.5, 0.35, 0.02, 8.0, 1000))], [0], vs, 3.0)
    tg = orc.sphere(0.35, (0.5, 0.5, 0.5), vs, 3.0)
    gdev = ctx.from_primitives([(vr.PRIM_NOISY_SPHERE, 0, (0.5, 0.5, 0.5, 0.35, 0.02, 8.0, 1000))], vs, 3.0)
    assert_nano_equal(gdev.download(), g0.to_nano(), "noisy sphere fixture")
    u = np.random.default_rng(100).normal(0.0, 0.5, size=(3, n, n, n)).astype(np.float32)
    dt = 0.5 * vs / float(np.abs(u).max())
    st, grad = gdev.adjoint_rollout(ctx.velocity_dense(u, n, 1.0 / n), n, dt, vr.RK3, steps, 2,
                                    ctx.upload(tg.to_nano()))
    fin = orc.rollout_dense(g0, u, n, 1.0 / n, dt, 3, steps)
    c, v = active_voxels(fin.to_nano())
    t = np.array([tg.get(*xyz) for xyz in c.tolist()], dtype=np.float64)
    loss_f32 = 0.5 * float(((v.astype(np.float64) - t) ** 2).sum())
    assert abs(st.loss - loss_f32) <= 1e-12 * loss_f32
    loss64, g64 = orc.rollout_adjoint_f64(g0, u.astype(np.float64), n, 1.0 / n, dt, 3, steps, tg)
    g64 = g64.reshape(3, n, n, n)
    scale = np.abs(g64).max()
    err = np.abs(grad - g64).max()
    rel = np.linalg.norm(grad - g64) / np.linalg.norm(g64)
    print(f"[gpu] noisy-sphere RK3 gradient vs FP64 adjoint: max err {err / scale:.2e} of max|g|, rel L2 {rel:.2e}",
          flush=True)
    assert err <= 2e-3 * scale and rel < 1e-3
    assert abs(st.loss - loss64) <= 1e-4 * loss64
