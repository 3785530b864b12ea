"""Pins the oracle's .vol reader/writer (src/io/vol_format.rs:53-147) and
extend_velocity (src/tools/level_set_track.rs:127-169): the reference's own
tests transcribed (vol_format.rs:175-296, level_set_track.rs:272-309), an
independent parse of the byte layout, read-side edge cases, and an
independent brute force of the nearest-interface rule."""
import os
import struct
import tempfile

import numpy as np
import pytest

from util import active_voxels, leaf_array

VOL_HEADER, VOL_LEAF = 32, 12 + 64 + 2048


def parse_vol(b):
    """Independent reader of the documented layout (vol_format.rs:3-20)."""
    magic, version, cls, vs, bg, n, reserved = struct.unpack_from("<4sIIdfI4s", b, 0)
    leaves = []
    for i in range(n):
        at = VOL_HEADER + i * VOL_LEAF
        o = struct.unpack_from("<3i", b, at)
        mask = np.frombuffer(b, dtype="<u8", count=8, offset=at + 12)
        vals = np.frombuffer(b, dtype="<f4", count=512, offset=at + 76)
        leaves.append((o, mask, vals))
    return {"magic": magic, "version": version, "class": cls, "vs": vs, "bg": bg, "n": n, "reserved": reserved,
            "leaves": leaves, "len": len(b)}


# ---------------------------------------------------------------- reference tests (vol_format.rs)
def test_roundtrip_sphere(orc):  # vol_format.rs:175
    g = orc.sphere(5.0, (0.0, 0.0, 0.0), 0.5, 3.0)
    buf = orc.vol_write(g)
    loaded, err = orc.vol_read(buf)
    assert err == ""
    assert loaded.leaf_count() == g.leaf_count() and loaded.active_count() == g.active_count()
    assert loaded.background() == g.background()
    # every leaf value and mask, and the header fields .vol carries (read_vol builds Grid::new, whose
    # inverse translation is +0.0 where make_level_set_sphere's set_transform gives -0.0)
    a, b = loaded.to_nano(), g.to_nano()
    assert a[320:] == b[320:] and a[:56] == b[:56]
    c, v = active_voxels(g.to_nano())
    for (x, y, z), val in list(zip(c.tolist(), v.tolist()))[:50]:
        assert loaded.get(x, y, z) == val


def test_roundtrip_empty_grid(orc):  # vol_format.rs:201
    g = orc.new_grid(42.0, 1.0)
    buf = orc.vol_write(g)
    assert len(buf) == VOL_HEADER
    loaded, err = orc.vol_read(buf)
    assert err == "" and loaded.leaf_count() == 0 and loaded.active_count() == 0 and loaded.background() == 42.0


def test_roundtrip_specific_voxels(orc):  # vol_format.rs:215
    g = orc.new_grid(-1.0, 0.25)
    orc.lib().or_set_class(g.h, 2)  # FogVolume
    g.set(0, 0, 0, 1.0)
    g.set(7, 7, 7, 2.0)
    g.set(100, 200, 50, 3.14)
    loaded, err = orc.vol_read(orc.vol_write(g))
    assert err == ""
    assert parse_vol(orc.vol_write(loaded))["class"] == 2
    assert loaded.get(0, 0, 0) == 1.0 and loaded.get(7, 7, 7) == 2.0
    assert abs(loaded.get(100, 200, 50) - 3.14) < 1e-6
    assert loaded.get(99, 99, 99) == -1.0


def test_verify_magic_bytes(orc):  # vol_format.rs:237
    buf = orc.vol_write(orc.new_grid(0.0, 1.0))
    assert buf[0:4] == b"VOLR" and struct.unpack_from("<I", buf, 4)[0] == 1


def test_bad_magic_returns_error(orc):  # vol_format.rs:247
    g, err = orc.vol_read(b"NOPE____________________________")
    assert g is None and "bad magic" in err


def test_bad_version_returns_error(orc):  # vol_format.rs:257
    g, err = orc.vol_read(b"VOLR" + struct.pack("<I", 99) + bytes(24))
    assert g is None and "version" in err


def test_roundtrip_preserves_active_mask(orc):  # vol_format.rs:268
    g = orc.new_grid(0.0, 1.0)
    g.set(3, 3, 3, 10.0)
    loaded, _ = orc.vol_read(orc.vol_write(g))
    assert loaded.is_active(3, 3, 3) and not loaded.is_active(4, 4, 4)


def test_file_roundtrip(orc):  # vol_format.rs:283
    g = orc.new_grid(0.0, 1.0)
    g.set(1, 2, 3, 99.0)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "volrus_test.vol")
        with open(p, "wb") as f:
            f.write(orc.vol_write(g))
        with open(p, "rb") as f:
            loaded, err = orc.vol_read(f.read())
    assert err == "" and loaded.get(1, 2, 3) == 99.0 and loaded.active_count() == 1


# ---------------------------------------------------------------- layout and read semantics
def test_vol_layout_matches_documented_format(orc):
    """Header fields at 0/4/8/12/20/24 (reserved zero), leaf records of 2124
    bytes (origin, mask, values), the same data as the NanoGrid of the grid."""
    g = orc.sphere(4.0, (1.5, -2.0, 0.25), 0.5, 3.0)
    orc.lib().or_set_class(g.h, 1)
    v = parse_vol(orc.vol_write(g))
    nano = g.to_nano()
    L = leaf_array(nano)
    assert (v["magic"], v["version"], v["class"], v["vs"], v["reserved"]) == (b"VOLR", 1, 1, 0.5, bytes(4))
    assert v["bg"] == np.float32(1.5) and v["n"] == len(L) and v["len"] == VOL_HEADER + len(L) * VOL_LEAF
    for (o, m, vals), leaf in zip(v["leaves"], L):
        assert tuple(o) == tuple(leaf["origin"])
        assert np.array_equal(m, leaf["mask"]) and np.array_equal(vals.view(np.uint32), leaf["values"].view(np.uint32))


def vol_bytes(cls, vs, bg, leaves):
    b = b"VOLR" + struct.pack("<IIdfI", 1, cls, vs, bg, len(leaves)) + bytes(4)
    for o, mask, vals in leaves:
        b += struct.pack("<3i", *o) + np.asarray(mask, "<u8").tobytes() + np.asarray(vals, "<f4").tobytes()
    return b


def test_vol_read_order_duplicates_and_unaligned_origins(orc):
    """read_vol places each record at origin & ~7 (set() then leaf_mut,
    vol_format.rs:128-137); a later record with the same leaf replaces the
    earlier one; records may come in any order; bytes past the last record
    are ignored."""
    rng = np.random.default_rng(1)
    m1 = [1, 0, 0, 0, 0, 0, 0, 1 << 63]
    m2 = [0, 2, 0, 0, 0, 0, 0, 0]
    v1, v2, v3 = (rng.normal(size=512).astype(np.float32) for _ in range(3))
    recs = [((16, 0, -8), m1, v1), ((3, 5, 7), m2, v2), ((17, 1, -3), m2, v3)]  # (17,1,-3) -> (16,0,-8)
    g, err = orc.vol_read(vol_bytes(0, 1.0, 0.5, recs) + b"trailing")
    assert err == ""
    assert g.leaf_count() == 2
    nano = g.to_nano()
    L = leaf_array(nano)
    assert [tuple(o) for o in L["origin"]] == [(0, 0, 0), (16, 0, -8)]
    assert np.array_equal(L["values"][1].view(np.uint32), v3.view(np.uint32)) and list(L["mask"][1]) == m2
    assert np.array_equal(L["values"][0].view(np.uint32), v2.view(np.uint32))


@pytest.mark.parametrize("cut", [0, 10, 31, 32 + 100, 32 + 2124 + 5])
def test_vol_truncated_is_an_error(orc, cut):
    b = vol_bytes(1, 1.0, 3.0, [((0, 0, 0), [1] + [0] * 7, np.zeros(512)), ((8, 0, 0), [1] + [0] * 7, np.ones(512))])
    g, err = orc.vol_read(b[:cut])
    assert g is None and err == "failed to fill whole buffer"


def test_vol_unknown_class_is_an_error(orc):
    g, err = orc.vol_read(vol_bytes(7, 1.0, 3.0, []))
    assert g is None and err == "unknown grid class: 7"


# ---------------------------------------------------------------- extend_velocity
def test_extend_velocity_fills_band_voxels(orc):  # level_set_track.rs:272
    sdf = orc.sphere(5.0, (0.0, 0.0, 0.0), 1.0, 3.0)
    vel = [orc.new_grid(0.0, 1.0) for _ in range(3)]
    c, v = active_voxels(sdf.to_nano())
    for (x, y, z), val in zip(c.tolist(), v.tolist()):
        if abs(val) < 1.0:
            for k, g in enumerate(vel):
                g.set(x, y, z, float(k + 1))
    before = vel[0].active_count()
    orc.extend_velocity(sdf, *vel, 3.0)
    assert vel[0].active_count() >= before
    band = int(((np.abs(v) >= 1.0) & (np.abs(v) < 3.0)).sum())
    assert vel[0].active_count() == before + band
    for (x, y, z), val in zip(c.tolist(), v.tolist()):
        if 1.0 <= abs(val) < 3.0:
            assert (vel[0].get(x, y, z), vel[1].get(x, y, z), vel[2].get(x, y, z)) == (1.0, 2.0, 3.0)


def test_extend_velocity_empty_sdf_noop(orc):  # level_set_track.rs:301
    sdf = orc.new_level_set(1.0, 3.0)
    vel = [orc.new_grid(0.0, 1.0) for _ in range(3)]
    orc.extend_velocity(sdf, *vel, 3.0)
    assert vel[0].active_count() == 0


def iteration_key(c):
    """The visiting order of the oracle (and of the reference within one
    128^3 node): internal-node origin, then leaf slot, then voxel offset."""
    c = np.asarray(c, dtype=np.int64)
    root = c >> 7
    leaf = (c >> 3) & 15
    vox = c & 7
    cols = [root[:, 0], root[:, 1], root[:, 2], leaf[:, 0], leaf[:, 1], leaf[:, 2], vox[:, 0], vox[:, 1], vox[:, 2]]
    return np.lexsort(cols[::-1])


@pytest.mark.parametrize("case", ["sphere", "two_nodes", "noisy_values"])
def test_extend_velocity_brute_force(orc, case):
    """Every band voxel gets the velocity of the interface voxel at minimum
    squared distance, the first in visiting order among equals (numpy brute
    force over all pairs); other voxels keep theirs; new voxels are active."""
    vs, hw = 1.0, 3.0
    if case == "sphere":
        sdf = orc.sphere(6.0, (0.3, -0.2, 0.1), vs, 4.0)
    elif case == "two_nodes":  # straddles internal nodes at x = 128
        sdf = orc.sphere(5.0, (128.0, 3.0, -1.0), vs, 4.0)
    else:
        sdf = orc.sphere(5.0, (0.0, 0.0, 0.0), vs, 4.0)
        rng = np.random.default_rng(3)
        c0, _ = active_voxels(sdf.to_nano())
        for (x, y, z) in c0[rng.choice(len(c0), 60, replace=False)].tolist():
            sdf.set(x, y, z, float(rng.uniform(-2.5, 2.5)))
    c, v = active_voxels(sdf.to_nano())
    order = iteration_key(c)
    c, v = c[order], v[order]
    rng = np.random.default_rng(7)
    vel = [orc.new_grid(0.25, vs) for _ in range(3)]
    vals = rng.normal(size=(3, len(c))).astype(np.float32)
    keep = rng.uniform(size=len(c)) < 0.7  # velocity defined at 70 % of the sdf's voxels
    for k in range(3):
        for (x, y, z), val in zip(c[keep].tolist(), vals[k][keep].tolist()):
            vel[k].set(x, y, z, val)
    want = [{tuple(p): vel[k].get(*p) for p in c.tolist()} for k in range(3)]
    iface = np.abs(v) < np.float32(vs)
    band = (np.abs(v) >= np.float32(vs)) & (np.abs(v) < np.float32(hw) * np.float32(vs))
    ic, bc = c[iface], c[band]
    d2 = ((bc[:, None, :].astype(np.int64) - ic[None, :, :].astype(np.int64)) ** 2).sum(axis=2)
    src = ic[np.argmin(d2, axis=1)]  # argmin keeps the first minimum in visiting order
    for k in range(3):
        for p, s in zip(bc.tolist(), src.tolist()):
            want[k][tuple(p)] = want[k][tuple(s)]
    orc.extend_velocity(sdf, *vel, hw)
    for k in range(3):
        for p in c.tolist():
            assert vel[k].get(*p) == want[k][tuple(p)], (k, p)
        for p in bc.tolist():
            assert vel[k].is_active(*p)
    if case == "two_nodes":
        assert len(set(map(tuple, (c >> 7).tolist()))) > 1
