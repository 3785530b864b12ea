"""Differentiable rollout on the GPU (vr_adjoint_rollout) vs the FP64 CPU
adjoint (itself pinned to central differences in test_oracle_math.py), and
against central differences of the GPU's own forward loss."""
import numpy as np
import pytest

from util import active_voxels

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[gpu] {msg}", flush=True)


def setup(orc, ctx, vr, n=16, vs=1.0 / 64, seed=4):
    rng = np.random.default_rng(seed)
    u = rng.normal(0.0, 0.5, size=(3, n, n, n)).astype(np.float32)
    g0 = orc.sphere(0.3, (0.5, 0.5, 0.5), vs, 3.0)
    tg = orc.sphere(0.3, (0.6, 0.5, 0.5), vs, 3.0)
    dt = 0.5 * vs / float(np.abs(u).max())
    return u, g0, tg, dt


@pytest.mark.parametrize("scheme,ckpt", [(1, 1), (1, 3), (3, 2)])
def test_adjoint_matches_cpu_adjoint(orc, ctx, vr, scheme, ckpt):
    announce(f"adjoint scheme={scheme} checkpoint_every={ckpt}")
    n, steps = 16, 6
    u, g0, tg, dt = setup(orc, ctx, vr, n)
    vel = ctx.velocity_dense(u, n, 1.0 / n)
    g = ctx.upload(g0.to_nano())
    target = ctx.upload(tg.to_nano())
    st, grad = g.adjoint_rollout(vel, n, dt, scheme, steps, ckpt, target)
    # forward loss: the f32 rollout is the oracle's f32 rollout bit for bit
    fin = orc.rollout_dense(g0, u, n, 1.0 / n, dt, scheme, steps)
    from util import active_voxels
    c, v = active_voxels(fin.to_nano())
    t = np.array([tg.get(*xyz) for xyz in c.tolist()], dtype=np.float64)
    loss_f32 = 0.5 * float(((v.astype(np.float64) - t) ** 2).sum())
    assert abs(st.loss - loss_f32) <= 1e-12 * loss_f32
    loss64, g64 = orc.rollout_adjoint_f64(g0, u.astype(np.float64), n, 1.0 / n, dt, scheme, steps, tg)
    g64 = g64.reshape(3, n, n, n)
    scale = np.abs(g64).max()
    err = np.abs(grad - g64).max()
    assert err <= 2e-3 * scale, (err, scale)
    rel = np.linalg.norm(grad - g64) / np.linalg.norm(g64)
    assert rel < 1e-3, rel
    assert abs(st.loss - loss64) <= 1e-4 * loss64
    assert st.forward_ms > 0 and st.adjoint_ms > 0


@pytest.mark.parametrize("scheme", [1, 3])
def test_adjoint_leaf_count_not_multiple_of_8(orc, ctx, vr, scheme):
    """137 leaves: the one-leaf-per-block kernels (forward, terminal, steps)
    see XCD ranges of unequal length; two calls, the second on reused memory."""
    announce(f"adjoint 137 leaves scheme={scheme}")
    n, steps = 16, 3
    u, _, tg, dt = setup(orc, ctx, vr, n)
    g0 = orc.sphere(0.27, (0.5, 0.5, 0.5), 1.0 / 64, 3.0)
    nano = g0.to_nano()
    assert ((len(nano) - 320) // 2128) % 8 != 0
    g = ctx.upload(nano)
    target = ctx.upload(tg.to_nano())
    vel = ctx.velocity_dense(u, n, 1.0 / n)
    fin = orc.rollout_dense(g0, u, n, 1.0 / n, dt, scheme, steps)
    from util import active_voxels
    c, v = active_voxels(fin.to_nano())
    t = np.array([tg.get(*xyz) for xyz in c.tolist()], dtype=np.float64)
    loss_f32 = 0.5 * float(((v.astype(np.float64) - t) ** 2).sum())
    loss64, g64 = orc.rollout_adjoint_f64(g0, u.astype(np.float64), n, 1.0 / n, dt, scheme, steps, tg)
    g64 = g64.reshape(3, n, n, n)
    for _ in range(2):
        st, grad = g.adjoint_rollout(vel, n, dt, scheme, steps, 1, target)
        assert abs(st.loss - loss_f32) <= 1e-12 * loss_f32
        rel = np.linalg.norm(grad - g64) / np.linalg.norm(g64)
        assert rel < 1e-3, rel


def test_adjoint_matches_central_differences_of_gpu_forward(orc, ctx, vr):
    announce("adjoint vs FD of the GPU forward")
    n, steps, scheme = 16, 6, 1
    u, g0, tg, dt = setup(orc, ctx, vr, n)
    g = ctx.upload(g0.to_nano())
    target = ctx.upload(tg.to_nano())
    st, grad = g.adjoint_rollout(ctx.velocity_dense(u, n, 1.0 / n), n, dt, scheme, steps, 1, target)
    flat = grad.reshape(-1)
    rng = np.random.default_rng(9)
    big = np.nonzero(np.abs(flat) > 0.05 * np.abs(flat).max())[0]
    h = 1e-3
    for i in rng.choice(big, 8, replace=False):
        up, dn = u.reshape(-1).copy(), u.reshape(-1).copy()
        up[i] += h
        dn[i] -= h
        lu = g.adjoint_rollout(ctx.velocity_dense(up, n, 1.0 / n), n, dt, scheme, steps, 1, target, False)[0].loss
        ld = g.adjoint_rollout(ctx.velocity_dense(dn, n, 1.0 / n), n, dt, scheme, steps, 1, target, False)[0].loss
        fd = (lu - ld) / (2 * h)
        assert abs(fd - flat[i]) <= 2e-2 * abs(flat).max(), (i, fd, flat[i])


@pytest.mark.parametrize("scheme", [1, 3])
def test_adjoint_fine_field_uses_atomic_scatter(orc, ctx, vr, scheme):
    """Velocity nodes as fine as the voxels (h = vs): a leaf spans more nodes
    than the separable window holds, so the scatter takes the atomic path."""
    announce(f"adjoint fine field scheme={scheme}")
    n, steps, vs = 64, 4, 1.0 / 64
    u, g0, tg, dt = setup(orc, ctx, vr, n, vs=vs, seed=11)
    vel = ctx.velocity_dense(u, n, 1.0 / n)
    st, grad = ctx.upload(g0.to_nano()).adjoint_rollout(vel, n, dt, scheme, steps, 2, ctx.upload(tg.to_nano()))
    loss64, g64 = orc.rollout_adjoint_f64(g0, u.astype(np.float64), n, 1.0 / n, dt, scheme, steps, tg)
    g64 = g64.reshape(3, n, n, n)
    assert np.abs(grad - g64).max() <= 2e-3 * np.abs(g64).max()
    assert np.linalg.norm(grad - g64) / np.linalg.norm(g64) < 1e-3
    assert abs(st.loss - loss64) <= 1e-4 * loss64


@pytest.mark.parametrize("scheme,ckpt,steps", [(3, 2, 6), (1, 1, 5)])
def test_adjoint_batch_matches_per_grid(orc, ctx, vr, scheme, ckpt, steps):
    """Several level sets in one launch per step (vr_adjoint_rollout_batch):
    noisy spheres and a sphere of 137 leaves, each with its own dense field
    (of its own size), time step and target; each grid's loss equals its own
    f32 rollout's, and its gradient the FP64 CPU adjoint's and the per-grid
    GPU call's."""
    announce(f"adjoint batch scheme={scheme} ckpt={ckpt}")
    vs = 1.0 / 64
    grids = [orc.scene([(2, (0.5, 0.5, 0.5, 0.3, 0.02, 8.0, 1000 + i))], [0], vs, 3.0) for i in range(2)]
    grids.append(orc.sphere(0.27, (0.5, 0.5, 0.5), vs, 3.0))
    targets = [orc.sphere(0.3, (0.55, 0.5, 0.5), vs, 3.0) for _ in grids]
    ns = [16, 12, 16]
    us = [np.random.default_rng(200 + i).normal(0.0, 0.5, size=(3, n, n, n)).astype(np.float32) for i, n in enumerate(ns)]
    dts = [0.5 * vs / float(np.abs(u).max()) for u in us]
    dg = [ctx.upload(g.to_nano()) for g in grids]
    dt_ = [ctx.upload(t.to_nano()) for t in targets]
    vels = [ctx.velocity_dense(u, n, 1.0 / n) for u, n in zip(us, ns)]
    stats, grads = vr.adjoint_rollout_batch(ctx, dg, vels, ns, dts, scheme, steps, ckpt, dt_)
    for i in range(len(grids)):
        fin = orc.rollout_dense(grids[i], us[i], ns[i], 1.0 / ns[i], dts[i], scheme, steps)
        c, v = active_voxels(fin.to_nano())
        t = np.array([targets[i].get(*xyz) for xyz in c.tolist()], dtype=np.float64)
        loss_f32 = 0.5 * float(((v.astype(np.float64) - t) ** 2).sum())
        assert abs(stats[i].loss - loss_f32) <= 1e-12 * loss_f32, i
        st1, g1 = dg[i].adjoint_rollout(vels[i], ns[i], dts[i], scheme, steps, ckpt, dt_[i])
        assert st1.loss == stats[i].loss
        scale = np.abs(g1).max()
        assert np.abs(grads[i] - g1).max() <= 1e-5 * scale, i
        loss64, g64 = orc.rollout_adjoint_f64(grids[i], us[i].astype(np.float64), ns[i], 1.0 / ns[i], dts[i], scheme,
                                              steps, targets[i])
        g64 = g64.reshape(3, ns[i], ns[i], ns[i])
        assert np.abs(grads[i] - g64).max() <= 2e-3 * np.abs(g64).max(), i
        assert np.linalg.norm(grads[i] - g64) / np.linalg.norm(g64) < 1e-3, i
    assert stats[0].forward_ms > 0 and stats[0].adjoint_ms > 0
