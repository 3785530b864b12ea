"""The reference's own hot-path tests, transcribed against the CPU oracle.

Each test names the Rust test it restates (file:line in revred/Sharc.Volrus).
Exact known answers are asserted exactly; qualitative ones as written there.
"""
import math
import struct

import numpy as np

NANO_HEADER = 320
NANO_LEAF = 2128


def leaves(b):
    n = struct.unpack_from("<I", b, 28)[0]
    out = []
    for i in range(n):
        off = NANO_HEADER + i * NANO_LEAF
        o = struct.unpack_from("<3i", b, off)
        m = struct.unpack_from("<8Q", b, off + 16)
        v = np.frombuffer(b, dtype=np.float32, count=512, offset=off + 80)
        out.append((o, m, v))
    return out


# ---------------------------------------------------------------- math/coord.rs, tree/leaf.rs
def test_coord_aligned_rounds_down(orc):  # coord.rs:142, leaf.rs:221
    g = orc.new_grid(0.0, 1.0)
    g.set(13, 7, 20, 1.0)
    (o, m, v), = leaves(g.to_nano())
    assert o == (8, 0, 16)


def test_coord_offset_in_tile_zyx(orc):  # coord.rs:150
    g = orc.new_grid(0.0, 1.0)
    g.set(2, 3, 5, 7.0)
    (o, m, v), = leaves(g.to_nano())
    off = 2 * 64 + 3 * 8 + 5
    assert v[off] == 7.0
    assert m[off // 64] == 1 << (off % 64)


def test_leaf_active_mask_512_bits(orc):  # leaf.rs:227
    g = orc.new_grid(0.0, 1.0)
    for x in range(8):
        for y in range(8):
            for z in range(8):
                g.set(x, y, z, 1.0)
    (o, m, v), = leaves(g.to_nano())
    assert m == (2 ** 64 - 1,) * 8
    assert g.active_count() == 512


def test_leaf_set_inactive_resets(orc):  # leaf.rs:210
    g = orc.new_grid(-1.0, 1.0)
    g.set(1, 1, 1, 5.0)
    assert g.is_active(1, 1, 1)
    g.deactivate(1, 1, 1)
    assert not g.is_active(1, 1, 1)
    assert g.get(1, 1, 1) == -1.0


# ---------------------------------------------------------------- tools/interpolation.rs
def test_trilinear_at_voxel_center_returns_exact(orc):  # interpolation.rs:56
    g = orc.new_grid(0.0, 1.0)
    g.set(5, 5, 5, 42.0)
    assert g.sample(5.0, 5.0, 5.0) == 42.0


def test_trilinear_midpoint_interpolates(orc):  # interpolation.rs:69
    g = orc.new_grid(0.0, 1.0)
    g.set(0, 0, 0, 0.0)
    g.set(1, 0, 0, 10.0)
    assert abs(g.sample(0.5, 0.0, 0.0) - 5.0) < 1e-5


def test_trilinear_respects_voxel_size(orc):  # interpolation.rs:80
    g = orc.new_grid(0.0, 0.5)
    g.set(0, 0, 0, 0.0)
    g.set(1, 0, 0, 10.0)
    assert abs(g.sample(0.25, 0.0, 0.0) - 5.0) < 1e-5


def test_trilinear_on_sphere_sdf_smooth(orc):  # interpolation.rs:93
    g = orc.sphere(5.0, (0, 0, 0), 0.5, 3.0)
    assert abs(g.sample(5.0, 0.0, 0.0)) < 1.0
    assert g.sample(4.0, 0.0, 0.0) < 0.0


# ---------------------------------------------------------------- tools/advection.rs
def uniform_velocity(orc, vs, vel, bmin, bmax):
    """advection.rs:116-134: the field lives in the background of three grids."""
    gs = []
    for c in range(3):
        g = orc.new_grid(vel[c], vs)
        g.set(*bmin, vel[c])
        g.set(*bmax, vel[c])
        gs.append(g)
    return gs


def test_uniform_velocity_moves_sphere_sdf(orc):  # advection.rs:152
    grid = orc.sphere(5.0, (0, 0, 0), 0.5, 3.0)
    vx, vy, vz = uniform_velocity(orc, 0.5, (1.0, 0.0, 0.0), (-30, -30, -30), (30, 30, 30))
    adv = orc.advect_grids3(grid, vx, vy, vz, 1.0)
    assert abs(adv.sample(6.0, 0.0, 0.0)) < 1.0
    assert adv.sample(5.0, 0.0, 0.0) < 0.0
    assert adv.sample(7.0, 0.0, 0.0) > 0.0


# ---------------------------------------------------------------- tools/level_set_track.rs
def test_find_zero_crossings_sphere_nonempty(orc):  # level_set_track.rs:186
    g = orc.sphere(5.0, (0, 0, 0), 1.0, 3.0)
    c = orc.zero_crossings(g)
    assert len(c) > 0
    r = np.sqrt((c.astype(np.float64) ** 2).sum(axis=1))
    assert np.all(np.abs(r - 5.0) < 2.5)


def test_find_zero_crossings_empty_grid(orc):  # level_set_track.rs:205
    g = orc.new_level_set(1.0, 3.0)
    assert len(orc.zero_crossings(g)) == 0


def test_trim_narrow_band_shrinks_count(orc):  # level_set_track.rs:211
    g = orc.sphere(5.0, (0, 0, 0), 1.0, 5.0)
    before = g.active_count()
    orc.trim(g, 2.0)
    after = g.active_count()
    assert after < before
    for o, m, v in leaves(g.to_nano()):
        for off in range(512):
            if (m[off // 64] >> (off % 64)) & 1:
                assert abs(v[off]) < 2.0


def test_trim_narrow_band_empty_grid_noop(orc):  # level_set_track.rs:235
    g = orc.new_level_set(1.0, 3.0)
    orc.trim(g, 2.0)
    assert g.active_count() == 0


def test_rebuild_narrow_band_preserves_sign(orc):  # level_set_track.rs:242
    g = orc.sphere(5.0, (0, 0, 0), 1.0, 3.0)
    before = {}
    for o, m, v in leaves(g.to_nano()):
        for off in range(512):
            if (m[off // 64] >> (off % 64)) & 1:
                before[(o[0] + (off >> 6), o[1] + ((off >> 3) & 7), o[2] + (off & 7))] = float(v[off])
    orc.rebuild(g, 3.0)
    for c, old in before.items():
        if g.is_active(*c) and abs(old) > 0.1:
            assert math.copysign(1.0, old) == math.copysign(1.0, g.get(*c))


def test_rebuild_narrow_band_empty_grid_noop(orc):  # level_set_track.rs:265
    g = orc.new_level_set(1.0, 3.0)
    orc.rebuild(g, 3.0)
    assert g.active_count() == 0


# ---------------------------------------------------------------- tools/volume_to_mesh.rs
def test_sphere_to_mesh_produces_triangles(orc):  # volume_to_mesh.rs:511
    v, t, wt = orc.mesh(orc.sphere(5.0, (0, 0, 0), 1.0, 3.0), 0.0)
    assert len(t) > 0 and len(v) > 0


def test_sphere_mesh_vertex_count_reasonable(orc):  # volume_to_mesh.rs:525
    v, t, wt = orc.mesh(orc.sphere(5.0, (0, 0, 0), 1.0, 3.0), 0.0)
    assert 50 < len(t) < 100_000


def test_sphere_mesh_is_watertight(orc):  # volume_to_mesh.rs:543
    v, t, wt = orc.mesh(orc.sphere(5.0, (0, 0, 0), 1.0, 3.0), 0.0)
    assert wt


def test_sphere_mesh_vertices_near_surface(orc):  # volume_to_mesh.rs:555
    v, t, wt = orc.mesh(orc.sphere(5.0, (0, 0, 0), 0.5, 5.0), 0.0)
    d = np.sqrt((v ** 2).sum(axis=1))
    assert (np.abs(d - 5.0) < 1.0).mean() > 0.75


def test_empty_grid_produces_no_mesh(orc):  # volume_to_mesh.rs:582
    v, t, wt = orc.mesh(orc.new_grid(3.0, 1.0), 0.0)
    assert len(v) == 0 and len(t) == 0


# ---------------------------------------------------------------- tools/sphere.rs
def test_sphere_signs_and_band(orc):  # sphere.rs:66-108
    g = orc.sphere(5.0, (0, 0, 0), 1.0, 3.0)
    assert g.active_count() > 0
    assert g.get(3, 0, 0) < 0.0
    assert g.get(7, 0, 0) > 0.0
    assert abs(g.get(5, 0, 0)) < 0.5
    g10 = orc.sphere(10.0, (0, 0, 0), 1.0, 3.0)
    assert g10.get(20, 0, 0) == g10.background()


def test_sphere_with_offset_center(orc):  # sphere.rs:111
    g = orc.sphere(3.0, (10.0, 20.0, 30.0), 0.5, 3.0)
    assert g.active_count() > 0
    assert abs(g.get(26, 40, 60)) < 1.0


# ---------------------------------------------------------------- tools/csg.rs
def test_union_of_two_spheres_has_more_voxels(orc):  # csg.rs:86
    s1 = orc.sphere(5.0, (0, 0, 0), 1.0, 3.0)
    s2 = orc.sphere(5.0, (6.0, 0, 0), 1.0, 3.0)
    u = orc.csg(s1, s2, 0)
    assert u.active_count() >= s1.active_count() and u.active_count() >= s2.active_count()


def test_intersection_of_overlapping_spheres(orc):  # csg.rs:97
    s1 = orc.sphere(5.0, (0, 0, 0), 1.0, 3.0)
    s2 = orc.sphere(5.0, (4.0, 0, 0), 1.0, 3.0)
    i = orc.csg(s1, s2, 1)
    assert 0 < i.active_count() <= s1.active_count() + s2.active_count()


def test_difference_removes_overlap(orc):  # csg.rs:108
    s1 = orc.sphere(5.0, (0, 0, 0), 1.0, 3.0)
    s2 = orc.sphere(5.0, (4.0, 0, 0), 1.0, 3.0)
    d = orc.csg(s1, s2, 2)
    assert d.get(3, 0, 0) < 0.0
    assert d.get(6, 0, 0) > 0.0


def test_union_is_commutative(orc):  # csg.rs:149
    s1 = orc.sphere(3.0, (0, 0, 0), 1.0, 2.0)
    s2 = orc.sphere(3.0, (4.0, 0, 0), 1.0, 2.0)
    assert orc.csg(s1, s2, 0).active_count() == orc.csg(s2, s1, 0).active_count()


# ---------------------------------------------------------------- nano/lookup.rs
def test_round_trip_grid_to_nano_to_grid(orc):  # lookup.rs:109
    g = orc.new_grid(-1.0, 0.5)
    g.set(0, 0, 0, 1.0)
    g.set(1, 2, 3, 2.0)
    g.set(10, 20, 5, 3.0)
    g2 = orc.from_nano(g.to_nano())
    assert g2.get(0, 0, 0) == 1.0 and g2.get(1, 2, 3) == 2.0 and g2.get(10, 20, 5) == 3.0
    assert g2.get(99, 99, 99) == -1.0 and g2.background() == -1.0


def test_header_fields_match_source_grid(orc):  # lookup.rs:127
    g = orc.new_grid(0.5, 0.25)
    orc.lib().or_set_class(g.h, 1)
    g.set(5, 10, 15, 42.0)
    g.set(-3, -7, 2, -1.0)
    b = g.to_nano()
    assert b[0:4] == b"NANO"
    assert struct.unpack_from("<I", b, 4)[0] == 1
    assert struct.unpack_from("<I", b, 8)[0] == 1
    assert abs(struct.unpack_from("<d", b, 16)[0] - 0.25) < 1e-10
    assert struct.unpack_from("<f", b, 24)[0] == 0.5
    assert struct.unpack_from("<3i", b, 32) == (-3, -7, 2)
    assert struct.unpack_from("<3i", b, 44) == (5, 10, 15)


def test_get_returns_correct_values_and_background(orc):  # lookup.rs:146,158
    g = orc.new_grid(0.0, 1.0)
    g.set(1, 2, 3, 42.0)
    g.set(4, 5, 6, 99.0)
    b = g.to_nano()
    assert orc.nano_get(b, 1, 2, 3) == 42.0 and orc.nano_get(b, 4, 5, 6) == 99.0
    g = orc.new_grid(-7.5, 1.0)
    g.set(0, 0, 0, 1.0)
    assert orc.nano_get(g.to_nano(), 100, 100, 100) == -7.5


def test_is_active_matches_source(orc):  # lookup.rs:169
    g = orc.new_grid(0.0, 1.0)
    g.set(3, 3, 3, 5.0)
    b = g.to_nano()
    assert orc.nano_is_active(b, 3, 3, 3)
    assert not orc.nano_is_active(b, 0, 0, 0)
    assert not orc.nano_is_active(b, 100, 100, 100)


def test_as_bytes_length_matches_expected(orc):  # lookup.rs:181
    g = orc.new_grid(0.0, 1.0)
    g.set(0, 0, 0, 1.0)
    g.set(8, 0, 0, 2.0)
    assert len(g.to_nano()) == NANO_HEADER + 2 * NANO_LEAF


def test_from_bytes_rejects_invalid_and_truncated(orc):  # lookup.rs:195,208
    ok, msg = orc.nano_validate(b"XYZW" + b"\0" * (NANO_HEADER - 4))
    assert not ok and "Invalid magic" in msg
    ok, msg = orc.nano_validate(b"\0" * 10)
    assert not ok


def test_from_bytes_round_trip(orc):  # lookup.rs:215
    g = orc.new_grid(0.0, 1.0)
    g.set(1, 2, 3, 10.0)
    g.set(9, 10, 11, 20.0)
    b = g.to_nano()
    ok, _ = orc.nano_validate(b)
    assert ok
    assert orc.nano_get(b, 1, 2, 3) == 10.0 and orc.nano_get(b, 9, 10, 11) == 20.0


def test_leaf_ordering_is_sorted(orc):  # lookup.rs:230
    g = orc.new_grid(0.0, 1.0)
    for c, v in [((100, 0, 0), 1.0), ((0, 0, 0), 2.0), ((50, 50, 50), 3.0), ((-10, -10, -10), 4.0)]:
        g.set(*c, v)
    origins = [o for o, m, v in leaves(g.to_nano())]
    assert origins == sorted(origins) and len(set(origins)) == len(origins)


def test_empty_grid_produces_valid_nanogrid(orc):  # lookup.rs:278
    g = orc.new_grid(5.0, 1.0)
    b = g.to_nano()
    assert struct.unpack_from("<I", b, 28)[0] == 0 and struct.unpack_from("<f", b, 24)[0] == 5.0
    assert orc.nano_get(b, 0, 0, 0) == 5.0 and not orc.nano_is_active(b, 0, 0, 0)
    g2 = orc.from_nano(b)
    assert g2.background() == 5.0 and g2.active_count() == 0


def test_many_leaves_and_negative_coordinates(orc):  # lookup.rs:311,330
    g = orc.new_grid(0.0, 1.0)
    for i in range(20):
        g.set(i * 8, i * 8, i * 8, float(i))
    b = g.to_nano()
    assert struct.unpack_from("<I", b, 28)[0] == 20
    for i in range(20):
        assert orc.nano_get(b, i * 8, i * 8, i * 8) == float(i)
    g = orc.new_grid(-1.0, 1.0)
    g.set(-5, -10, -3, 77.0)
    g.set(-100, -200, -300, 88.0)
    b = g.to_nano()
    assert orc.nano_get(b, -5, -10, -3) == 77.0 and orc.nano_get(b, -100, -200, -300) == 88.0
    g2 = orc.from_nano(b)
    assert g2.get(-5, -10, -3) == 77.0 and g2.get(-100, -200, -300) == 88.0


def test_affine_transform_preserved(orc):  # lookup.rs:349
    g = orc.new_grid(0.0, 0.25)
    g.set(4, 8, 12, 7.0)
    b = g.to_nano()
    mat = struct.unpack_from("<16d", b, 56)
    assert abs(mat[0] - 0.25) < 1e-10 and abs(mat[5] - 0.25) < 1e-10 and abs(mat[10] - 0.25) < 1e-10
