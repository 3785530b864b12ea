"""The .vol boundary (read_vol / write_vol, src/io/vol_format.rs:53-147) and
extend_velocity (src/tools/level_set_track.rs:127-169) on the GPU against the
oracle, byte for byte."""
import struct

import numpy as np
import pytest

from util import active_voxels, assert_nano_equal, sphere128

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[gpu] {msg}", flush=True)


def vol_bytes(cls, vs, bg, leaves):
    b = b"VOLR" + struct.pack("<IIdfI", 1, cls, vs, bg, len(leaves)) + bytes(4)
    for o, mask, vals in leaves:
        b += struct.pack("<3i", *o) + np.asarray(mask, "<u8").tobytes() + np.asarray(vals, "<f4").tobytes()
    return b


# ---------------------------------------------------------------- .vol
def vol_cases(orc):
    g1 = orc.new_grid(-1.0, 0.25)
    orc.lib().or_set_class(g1.h, 2)
    for p, v in [((0, 0, 0), 1.0), ((7, 7, 7), 2.0), ((100, 200, 50), 3.14), ((-300, 5, -9), -4.5)]:
        g1.set(*p, v)
    g2 = orc.new_grid(0.0, 1.0)
    g2.set(0, 0, 0, 1.0)
    g2.set(9, 0, 0, 2.0)
    g2.deactivate(9, 0, 0)  # a leaf with no active voxel
    return [("sphere128", sphere128(orc)), ("sphere_neg", orc.sphere(5.0, (0.0, 0.0, 0.0), 0.5, 3.0)),
            ("specific", g1), ("empty_leaf", g2), ("empty", orc.new_grid(42.0, 1.0))]


def test_vol_round_trips(orc, ctx):
    announce("vol round trips")
    for name, og in vol_cases(orc):
        want = orc.vol_write(og)
        g = ctx.upload(og.to_nano())
        assert g.download_vol() == want, f"{name}: write_vol bytes"
        back = ctx.upload_vol(want)
        ref, err = orc.vol_read(want)
        assert err == ""
        assert_nano_equal(back.download(), ref.to_nano(), f"{name}: read_vol")
        assert back.download_vol() == want, f"{name}: vol -> device -> vol"


def test_vol_read_semantics(orc, ctx):
    """Records out of order, a leaf given twice (the later wins), origins
    that are not multiples of 8, bytes after the last record."""
    announce("vol read semantics")
    rng = np.random.default_rng(1)
    vals = [rng.normal(size=512).astype(np.float32) for _ in range(5)]
    m = [[1, 0, 0, 0, 0, 0, 0, 1 << 63], [0, 2, 0, 0, 0, 0, 0, 0], [0] * 8, [5] * 8, [1 << 40] + [0] * 7]
    recs = [((16, 0, -8), m[0], vals[0]), ((3, 5, 7), m[1], vals[1]), ((17, 1, -3), m[3], vals[2]),
            ((-130, 250, 8), m[2], vals[3]), ((-129, 255, 15), m[4], vals[4])]
    b = vol_bytes(1, 0.5, 1.5, recs) + b"trailing bytes"
    ref, err = orc.vol_read(b)
    assert err == ""
    g = ctx.upload_vol(b)
    assert_nano_equal(g.download(), ref.to_nano(), "vol read semantics")
    assert g.counts() == (ref.leaf_count(), ref.active_count())


@pytest.mark.parametrize("case", ["short_header", "bad_magic", "bad_version", "bad_class", "truncated_record"])
def test_vol_errors(orc, ctx, vr, case):
    announce(f"vol error {case}")
    good = vol_bytes(1, 1.0, 3.0, [((0, 0, 0), [1] + [0] * 7, np.zeros(512))])
    b = {"short_header": good[:20], "bad_magic": b"NOPE" + good[4:],
         "bad_version": good[:4] + struct.pack("<I", 99) + good[8:],
         "bad_class": good[:8] + struct.pack("<I", 3) + good[12:], "truncated_record": good[:-1]}[case]
    _, want = orc.vol_read(b)
    assert want
    with pytest.raises(vr.VolrusError) as ei:
        ctx.upload_vol(b)
    assert ei.value.code == vr.VR_ERR_INVALID_VOL and str(ei.value).endswith(want)


def test_vol_large_round_trip(orc, ctx, vr):
    """A 1/512 slotted-sphere scene (3206 leaves) generated on the device."""
    announce("vol large")
    g = ctx.from_primitives([(vr.PRIM_SPHERE, 0, (0.5, 0.75, 0.5, 0.15)),
                             (vr.PRIM_BOX, vr.CSG_DIFFERENCE, (0.475, 0.55, 0.3, 0.525, 0.725, 0.7))], 1.0 / 512, 3.0)
    nano = g.download()
    vb = g.download_vol()
    assert vb == orc.vol_write(orc.from_nano(nano))
    back = ctx.upload_vol(vb)
    ref, _ = orc.vol_read(vb)
    assert_nano_equal(back.download(), ref.to_nano(), "large vol")


# ---------------------------------------------------------------- extend_velocity
def velocity_grids(orc, c, fraction, seed, bg=0.25, vs=1.0):
    rng = np.random.default_rng(seed)
    vals = rng.normal(size=(3, len(c))).astype(np.float32)
    keep = rng.uniform(size=len(c)) < fraction
    grids = [orc.new_grid(bg, vs) for _ in range(3)]
    for k in range(3):
        for (x, y, z), val in zip(c[keep].tolist(), vals[k][keep].tolist()):
            grids[k].set(x, y, z, val)
    return grids


def run_extend(orc, ctx, sdf, vel, hw, what):
    dev = [ctx.upload(g.to_nano()) for g in vel]
    ctx.extend_velocity(ctx.upload(sdf.to_nano()), *dev, hw)
    orc.extend_velocity(sdf, *vel, hw)
    for k in range(3):
        assert_nano_equal(dev[k].download(), vel[k].to_nano(), f"{what}: component {k}")


def test_extend_velocity_reference_case(orc, ctx):
    """level_set_track.rs:272: velocity set only at the interface of a
    sphere r = 5 (vs = 1, band 3): the band voxels gain it, new leaves
    included."""
    announce("extend velocity reference case")
    sdf = orc.sphere(5.0, (0.0, 0.0, 0.0), 1.0, 3.0)
    c, v = active_voxels(sdf.to_nano())
    vel = [orc.new_grid(0.0, 1.0) for _ in range(3)]
    for (x, y, z), val in zip(c.tolist(), v.tolist()):
        if abs(val) < 1.0:
            for k in range(3):
                vel[k].set(x, y, z, float(k + 1))
    before = vel[0].leaf_count()
    run_extend(orc, ctx, sdf, vel, 3.0, "reference case")
    assert vel[0].leaf_count() >= before


@pytest.mark.parametrize("case", ["sphere", "two_nodes", "noisy_values", "sphere64", "far_band"])
def test_extend_velocity_matches_oracle(orc, ctx, case):
    announce(f"extend velocity {case}")
    vs, hw = 1.0, 3.0
    if case == "sphere":
        sdf = orc.sphere(6.0, (0.3, -0.2, 0.1), vs, 4.0)
    elif case == "two_nodes":
        sdf = orc.sphere(5.0, (128.0, 3.0, -1.0), vs, 4.0)
    elif case == "noisy_values":
        sdf = orc.sphere(5.0, (0.0, 0.0, 0.0), vs, 4.0)
        rng = np.random.default_rng(3)
        c0, _ = active_voxels(sdf.to_nano())
        for (x, y, z) in c0[rng.choice(len(c0), 60, replace=False)].tolist():
            sdf.set(x, y, z, float(rng.uniform(-2.5, 2.5)))
    elif case == "sphere64":  # voxel size 1/64, band 3
        vs = 1.0 / 64
        sdf = orc.sphere(0.2, (0.4, 0.5, 0.55), vs, 4.0)
    else:  # band voxels far from every interface voxel: the brute-force pass
        sdf = orc.sphere(4.0, (0.0, 0.0, 0.0), vs, 3.0)
        for p in [(40, 3, 5), (-37, 20, 0), (25, 25, 25), (60, 0, 0)]:
            sdf.set(*p, 2.0)
        sdf.set(61, 0, 0, -1.5)
    c, _ = active_voxels(sdf.to_nano())
    vel = velocity_grids(orc, c, 0.7, 7, vs=vs)
    run_extend(orc, ctx, sdf, vel, hw, case)


def test_extend_velocity_no_interface_is_noop(orc, ctx):
    announce("extend velocity no interface")
    sdf = orc.new_level_set(1.0, 3.0)
    for p in [(0, 0, 0), (3, 4, 5)]:
        sdf.set(*p, 2.0)  # band voxels only
    vel = velocity_grids(orc, np.array([[0, 0, 0], [8, 8, 8]]), 1.0, 2)
    run_extend(orc, ctx, sdf, vel, 3.0, "no interface")
    run_extend(orc, ctx, orc.new_level_set(1.0, 3.0), vel, 3.0, "empty sdf")
