"""Helpers shared by the tests: NanoGrid parsing, fixtures, comparisons."""
import math
import struct

import numpy as np

NANO_HEADER = 320
NANO_LEAF = 2128

LEAF_DTYPE = np.dtype([("origin", "<i4", 3), ("pad", "<i4"), ("mask", "<u8", 8), ("values", "<f4", 512)])
assert LEAF_DTYPE.itemsize == NANO_LEAF


def leaf_array(b):
    n = struct.unpack_from("<I", b, 28)[0]
    return np.frombuffer(b, dtype=LEAF_DTYPE, count=n, offset=NANO_HEADER)


def header(b):
    return {
        "magic": b[0:4],
        "version": struct.unpack_from("<I", b, 4)[0],
        "class": struct.unpack_from("<I", b, 8)[0],
        "vs": struct.unpack_from("<d", b, 16)[0],
        "bg": struct.unpack_from("<f", b, 24)[0],
        "leaf_count": struct.unpack_from("<I", b, 28)[0],
        "bbox_min": struct.unpack_from("<3i", b, 32),
        "bbox_max": struct.unpack_from("<3i", b, 44),
    }


def active_voxels(b):
    """(coords int32 [A,3], values float32 [A]) in leaf order, offset order."""
    L = leaf_array(b)
    if len(L) == 0:
        return np.zeros((0, 3), np.int32), np.zeros(0, np.float32)
    bits = np.unpackbits(L["mask"].view(np.uint8).reshape(len(L), 64), axis=1, bitorder="little").astype(bool)
    li, off = np.nonzero(bits)
    o = L["origin"][li]
    c = np.stack([o[:, 0] + (off >> 6), o[:, 1] + ((off >> 3) & 7), o[:, 2] + (off & 7)], axis=1).astype(np.int32)
    return c, L["values"][li, off]


def assert_nano_equal(got, want, what=""):
    """Byte-exact NanoGrid comparison with a readable first difference."""
    if got == want:
        return
    hg, hw = header(got), header(want)
    assert hg == hw, f"{what}: header differs: {hg} vs {hw}"
    assert got[:NANO_HEADER] == want[:NANO_HEADER], f"{what}: header bytes differ"
    Lg, Lw = leaf_array(got), leaf_array(want)
    for i in range(len(Lw)):
        if tuple(Lg["origin"][i]) != tuple(Lw["origin"][i]):
            raise AssertionError(f"{what}: leaf {i} origin {Lg['origin'][i]} vs {Lw['origin'][i]}")
        if not np.array_equal(Lg["mask"][i], Lw["mask"][i]):
            raise AssertionError(f"{what}: leaf {i} at {Lw['origin'][i]} mask differs")
        vg, vw = Lg["values"][i].view(np.uint32), Lw["values"][i].view(np.uint32)
        if not np.array_equal(vg, vw):
            k = int(np.nonzero(vg != vw)[0][0])
            o = Lw["origin"][i]
            raise AssertionError(
                f"{what}: leaf {i} voxel ({o[0] + (k >> 6)},{o[1] + ((k >> 3) & 7)},{o[2] + (k & 7)}) "
                f"value {Lg['values'][i][k]!r} vs {Lw['values'][i][k]!r}; "
                f"{int((vg != vw).sum())} differing values in this leaf")
    raise AssertionError(f"{what}: bytes differ outside header/leaves")


def rigid_velocity_grids(orc, sdf_nano, vs, omega=(0.0, 0.0, 2 * math.pi), center=(0.5, 0.5, 0.5)):
    """vx, vy, vz oracle grids holding the rigid-rotation field (host f64 -> f32)
    at every active voxel of the level set, background 0."""
    c, _ = active_voxels(sdf_nano)
    w = c.astype(np.float64) * vs
    r = w - np.array(center)
    u = np.stack([omega[1] * r[:, 2] - omega[2] * r[:, 1],
                  omega[2] * r[:, 0] - omega[0] * r[:, 2],
                  omega[0] * r[:, 1] - omega[1] * r[:, 0]], axis=1).astype(np.float32)
    grids = [orc.new_grid(0.0, vs) for _ in range(3)]
    for k in range(3):
        g = grids[k]
        for (x, y, z), val in zip(c.tolist(), u[:, k].tolist()):
            g.set(x, y, z, val)
    return grids


def sphere128(orc):
    return orc.sphere(0.3, (0.5, 0.5, 0.5), 1.0 / 128, 3.0)


def sphere128_rigid_dt():
    return 0.5 * (1.0 / 128) / math.pi


def slotted_sphere(orc, vs, hw=3.0):
    """Zalesak sphere: r = 0.15 at (0.5, 0.75, 0.5) minus a slot 0.05 wide in x,
    from the sphere's bottom (y = 0.6) up to y = 0.725, through all z."""
    prims = [(0, (0.5, 0.75, 0.5, 0.15)), (1, (0.475, 0.55, 0.3, 0.525, 0.725, 0.7))]
    return orc.scene(prims, [0, 2], vs, hw)


def mesh_equal(vg, tg, vw, tw, vs, what=""):
    assert tg.shape == tw.shape, f"{what}: {len(tg)} vs {len(tw)} triangles"
    assert vg.shape == vw.shape, f"{what}: {len(vg)} vs {len(vw)} vertices"
    if len(tw):
        bad = np.nonzero((tg != tw).any(axis=1))[0]
        assert len(bad) == 0, f"{what}: triangle {bad[0]} {tg[bad[0]]} vs {tw[bad[0]]}"
    if len(vw):
        err = np.abs(vg - vw).max()
        assert err <= 1e-5 * vs, f"{what}: vertex error {err}"
    return float(np.abs(vg - vw).max()) if len(vw) else 0.0


def config4_problem(orc, seed=0):
    """The differentiable-rollout workload: sphere r = 0.3 at 1/256, target the
    sphere moved by (0.1, 0, 0), dense 64^3 field of N(0, 0.5^2) (numpy
    default_rng(seed)), dt = 0.5 vs / max|u|, 20 RK1 steps."""
    vs, n = 1.0 / 256, 64
    u = np.random.default_rng(seed).normal(0.0, 0.5, size=(3, n, n, n)).astype(np.float32)
    g0 = orc.sphere(0.3, (0.5, 0.5, 0.5), vs, 3.0)
    tg = orc.sphere(0.3, (0.6, 0.5, 0.5), vs, 3.0)
    dt = 0.5 * vs / float(np.abs(u).max())
    return u, g0, tg, dt, n, 20


def config4_fd64(args):
    """Central difference (FP64 oracle loss) of the config-4 loss in one
    velocity entry: (L(u + h e_i) - L(u - h e_i)) / 2h.  Runs in a worker
    process (CPU only)."""
    index, h = args
    import sys
    import os
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
    import oracle as orc
    u, g0, tg, dt, n, steps = config4_problem(orc)
    u64 = u.astype(np.float64).reshape(-1)
    up, dn = u64.copy(), u64.copy()
    up[index] += h
    dn[index] -= h
    lu = orc.rollout_loss_f64(g0, up, n, 1.0 / n, dt, 1, steps, tg)
    ld = orc.rollout_loss_f64(g0, dn, n, 1.0 / n, dt, 1, steps, tg)
    return (lu - ld) / (2 * h)


def white_noise_block(orc, seed, n=24, origin=(-8, -8, -8), bg=3.0):
    """Seeded uniform [-1, 1) values at every voxel of an n^3 block (all
    active): nearly every lattice edge crosses 0, so marching-cubes tiles hold
    far more vertices, triangles and edges owned by earlier tiles than a
    level set does."""
    vals = np.random.default_rng(seed).uniform(-1.0, 1.0, size=(n, n, n)).astype(np.float32)
    g = orc.new_grid(bg, 1.0)
    ox, oy, oz = origin
    for x in range(n):
        for y in range(n):
            for z in range(n):
                g.set(ox + x, oy + y, oz + z, float(vals[x, y, z]))
    return g
