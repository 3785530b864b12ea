"""HIP path vs the CPU oracle, through the C ABI, on the same inputs.

Bit-exact for masks, leaf topology, values (NanoGrid bytes) and MC triangle
indices; vertex positions within 1e-5 voxel (they come out bit-exact).
"""
import math
import struct

import numpy as np
import pytest

from util import (NANO_HEADER, active_voxels, assert_nano_equal, sphere128_rigid_dt, sphere128, header, leaf_array,
                  mesh_equal, rigid_velocity_grids, slotted_sphere, white_noise_block)

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[gpu] {msg}", flush=True)


# ---------------------------------------------------------------- NanoGrid boundary
def test_upload_download_round_trip(orc, ctx):
    announce("nano round trip")
    cases = [sphere128(orc), orc.new_grid(5.0, 1.0)]
    g = orc.new_grid(-1.0, 1.0)
    g.set(-5, -10, -3, 77.0)
    g.set(-100, -200, -300, 88.0)
    g.set(130, 130, 130, 4.0)
    cases.append(g)
    e = orc.new_grid(0.0, 1.0)
    e.set(0, 0, 0, 1.0)
    e.set(9, 0, 0, 2.0)
    e.deactivate(9, 0, 0)  # leaf with no active voxel stays allocated
    cases.append(e)
    for og in cases:
        b = og.to_nano()
        got = ctx.upload(b).download()
        assert_nano_equal(got, b, "round trip")


def test_upload_rejects_invalid_nano(orc, ctx, vr):
    announce("nano validation")
    b = bytearray(sphere128(orc).to_nano())
    bad = [bytes(b[:10]), b"XYZW" + bytes(b[4:]), bytes(b[:4]) + struct.pack("<I", 2) + bytes(b[8:]),
           bytes(b[:-1])]
    sw = bytearray(b)  # swap two leaves: not sorted
    a0, a1 = NANO_HEADER, NANO_HEADER + 2128
    sw[a0:a1], sw[a1:a1 + 2128] = b[a1:a1 + 2128], b[a0:a1]
    bad.append(bytes(sw))
    for x in bad:
        with pytest.raises(vr.VolrusError) as ei:
            ctx.upload(x)
        assert ei.value.code == vr.VR_ERR_INVALID_NANO


# ---------------------------------------------------------------- advection
def test_advect_rk1_grids3_sphere128(orc, ctx):
    """Sphere r=0.3 at 1/128, rigid rotation through the reference VelocityField path (vx/vy/vz grids), one RK1 step at CFL 0.5."""
    announce("advect rk1 grids3 sphere128")
    sdf = sphere128(orc)
    b = sdf.to_nano()
    vxyz = rigid_velocity_grids(orc, b, 1.0 / 128)
    want = orc.advect_grids3(sdf, *vxyz, sphere128_rigid_dt()).to_nano()
    g = ctx.upload(b)
    vg = [ctx.upload(v.to_nano()) for v in vxyz]
    vel = ctx.velocity_grids3(*vg)
    got = g.advect(vel, sphere128_rigid_dt(), 0.0, 1).download()
    assert_nano_equal(got, want, "advect grids3")


def test_uniform_velocity_moves_sphere_sdf(orc, ctx):
    """advection.rs:152 through the GPU path (background-encoded field)."""
    announce("advection.rs:152")
    grid = orc.sphere(5.0, (0, 0, 0), 0.5, 3.0)
    vs = []
    for c in range(3):
        vv = (1.0, 0.0, 0.0)[c]
        gg = orc.new_grid(vv, 0.5)
        gg.set(-30, -30, -30, vv)
        gg.set(30, 30, 30, vv)
        vs.append(gg)
    want = orc.advect_grids3(grid, *vs, 1.0).to_nano()
    vel = ctx.velocity_grids3(*[ctx.upload(v.to_nano()) for v in vs])
    got = ctx.upload(grid.to_nano()).advect(vel, 1.0).download()
    assert_nano_equal(got, want, "uniform")
    adv = orc.from_nano(got)
    assert abs(adv.sample(6.0, 0, 0)) < 1.0 and adv.sample(5.0, 0, 0) < 0.0 and adv.sample(7.0, 0, 0) > 0.0


@pytest.mark.parametrize("kind,params,scheme,period", [
    (3, (0, 0, 2 * math.pi, 0.5, 0.5, 0.5), 1, 0.0),
    (3, (0, 0, 2 * math.pi, 0.5, 0.5, 0.5), 3, 0.0),
    (2, (0.3, -0.7, 0.2), 3, 0.0),
    (4, (), 1, 8.0),
    (4, (), 3, 8.0),
])
def test_advect_analytic(orc, ctx, kind, params, scheme, period):
    announce(f"advect analytic kind={kind} scheme={scheme}")
    sdf = sphere128(orc)
    dt, t = sphere128_rigid_dt(), 0.37
    tf = orc.stage_factors(period, t, dt, scheme)
    want = orc.advect_analytic(sdf, kind, params, dt, scheme, tf).to_nano()
    vel = ctx.velocity_analytic(kind, params, period)
    got = ctx.upload(sdf.to_nano()).advect(vel, dt, t, scheme).download()
    assert_nano_equal(got, want, f"analytic {kind}/{scheme}")


def test_advect_dense_velocity_large_dt(orc, ctx):
    """Dense field, dt large enough that departure cells leave the leaf halo."""
    announce("advect dense")
    rng = np.random.default_rng(5)
    n = 16
    u = rng.normal(0.0, 0.5, size=(3, n, n, n)).astype(np.float32)
    # r = 0.27: 137 leaves, not a multiple of 8 (XCD ranges of unequal length)
    for r in (0.3, 0.27):
        sdf = orc.sphere(r, (0.5, 0.5, 0.5), 1.0 / 64, 3.0)
        for dt, scheme in [(0.5 / 64 / 1.5, 1), (6.0 / 64, 1), (3.0 / 64, 3)]:
            want = orc.advect_dense(sdf, u, n, 1.0 / n, dt, scheme).to_nano()
            vel = ctx.velocity_dense(u, n, 1.0 / n)
            got = ctx.upload(sdf.to_nano()).advect(vel, dt, 0.0, scheme).download()
            assert_nano_equal(got, want, f"dense r={r} dt={dt} scheme={scheme}")


def test_advect_departure_saturates_to_int_min(orc, ctx):
    """Departure points far below the i32 range: Rust's saturating `f64 as i32`
    (interpolation.rs:33-35) puts the cell at x = i32::MIN, where values were planted."""
    announce("advect saturating departure")
    g = orc.sphere(3.0, (0.0, 0.0, 0.0), 1.0, 3.0)  # voxel size 1: index = world
    for y in range(-8, 9):
        for z in range(-8, 9):
            g.set(-2 ** 31, y, z, 0.25 * y - 0.125 * z + 0.5)
    params = (1e14, 0.0, 0.0)
    want = orc.advect_analytic(g, 2, params, 1.0).to_nano()
    got = ctx.upload(g.to_nano()).advect(ctx.velocity_analytic(2, params), 1.0).download()
    assert_nano_equal(got, want, "saturating departure")
    # voxel (0, 1, 0) of the band sampled the planted plane
    assert orc.from_nano(got).get(0, 1, 0) == 0.75


def test_advect_input_with_empty_leaf(orc, ctx):
    announce("advect empty leaf")
    g = orc.sphere(5.0, (0, 0, 0), 1.0, 3.0)
    g.set(40, 40, 40, 1.0)
    g.deactivate(40, 40, 40)
    want = orc.advect_analytic(g, 2, (0.25, 0.0, -0.5), 1.0).to_nano()
    got = ctx.upload(g.to_nano()).advect(ctx.velocity_analytic(2, (0.25, 0.0, -0.5)), 1.0).download()
    assert_nano_equal(got, want, "empty leaf")


@pytest.mark.parametrize("scheme", [1, 3])
@pytest.mark.parametrize("disp", [(0.5, -0.5, 0.0), (-0.3, 0.0, 0.9), (1.5, -1.25, 0.75), (-2.5, 2.5, -0.01)])
def test_advect_uniform_displacements(orc, ctx, scheme, disp):
    """Uniform velocity, whose departure cells are tabulated per leaf and axis
    and whose brick is staged only on the sides they reach: displacements of
    either sign, zero, and past one voxel (cells outside the brick go through
    the tree)."""
    announce(f"advect uniform disp={disp} scheme={scheme}")
    vs = 1.0 / 64
    sdf = orc.sphere(0.3, (0.5, 0.5, 0.5), vs, 3.0)
    params = tuple(d * vs for d in disp)  # voxels per unit time, dt = 1
    tf = orc.stage_factors(0.0, 0.0, 1.0, scheme)
    want = orc.advect_analytic(sdf, 2, params, 1.0, scheme, tf).to_nano()
    got = ctx.upload(sdf.to_nano()).advect(ctx.velocity_analytic(2, params), 1.0, 0.0, scheme).download()
    assert_nano_equal(got, want, f"uniform disp={disp} scheme={scheme}")


# ---------------------------------------------------------------- band tools
def crossing_set(a):
    return set(map(tuple, np.asarray(a).tolist()))


def test_zero_crossings_match(orc, ctx):
    announce("zero crossings")
    for og in [orc.sphere(5.0, (0, 0, 0), 1.0, 3.0), sphere128(orc), slotted_sphere(orc, 1.0 / 128)]:
        want = crossing_set(orc.zero_crossings(og))
        got = ctx.upload(og.to_nano()).zero_crossings()
        assert len(got) == len(want) and crossing_set(got) == want
    assert len(ctx.upload(orc.new_level_set(1.0, 3.0).to_nano()).zero_crossings()) == 0


def test_trim_matches(orc, ctx):
    announce("trim")
    for hw in (2.0, 0.5, 10.0):
        og = orc.sphere(5.0, (0, 0, 0), 1.0, 5.0)
        g = ctx.upload(og.to_nano())
        orc.trim(og, hw)
        g.trim(hw)
        assert_nano_equal(g.download(), og.to_nano(), f"trim {hw}")


def test_rebuild_matches_brute_force(orc, ctx):
    announce("rebuild brute force")
    for og in [orc.sphere(5.0, (0, 0, 0), 1.0, 3.0), orc.sphere(4.0, (1.5, -2.0, 0.25), 0.5, 3.0)]:
        for hw in (3.0, 1.5, 5.0):
            src = og.to_nano()
            ref = orc.from_nano(src)
            orc.rebuild(ref, hw)
            g = ctx.upload(src)
            g.rebuild(hw)
            assert_nano_equal(g.download(), ref.to_nano(), f"rebuild hw={hw}")


def test_rebuild_after_advect_sphere128(orc, ctx):
    announce("rebuild sphere128")
    sdf = sphere128(orc)
    dt = sphere128_rigid_dt()
    params = (0, 0, 2 * math.pi, 0.5, 0.5, 0.5)
    adv = orc.advect_analytic(sdf, 3, params, dt, 1)
    ref = orc.from_nano(adv.to_nano())
    orc.rebuild_bounded(ref, 3.0)
    g = ctx.upload(sdf.to_nano()).advect(ctx.velocity_analytic(3, params), dt)
    g.rebuild(3.0)
    assert_nano_equal(g.download(), ref.to_nano(), "rebuild sphere128")


@pytest.mark.parametrize("op", ["trim", "rebuild"])
def test_lookups_after_in_place_compaction(orc, ctx, op):
    """The neighbour lists, internal-node slots and origins that an in-place
    compaction updates alongside the leaves: an advect whose departures leave
    the leaf's brick (a lookup through the tree per corner) and marching
    cubes, right after the compaction, against the oracle."""
    announce(f"lookups after {op}")
    og = orc.sphere(6.0, (0.5, -1.0, 2.0), 0.5, 5.0)
    g = ctx.upload(og.to_nano())
    if op == "trim":
        orc.trim(og, 1.5)
        g.trim(1.5)
    else:
        orc.rebuild(og, 2.0)
        g.rebuild(2.0)
    assert_nano_equal(g.download(), og.to_nano(), op)
    params = (0.0, 0.0, 2 * math.pi, 0.0, 0.0, 0.0)
    dt = 0.04  # up to ~3 voxels of rotation: corners outside the brick
    want = orc.advect_analytic(og, 3, params, dt, 1)
    got = g.advect(ctx.velocity_analytic(3, params), dt)
    assert_nano_equal(got.download(), want.to_nano(), f"advect after {op}")
    vw, tw, _ = orc.mesh(og, 0.0)
    vg, tg = g.mesh(0.0)
    assert np.array_equal(tg, tw) and np.array_equal(vg.view(np.uint64), vw.view(np.uint64)), f"mesh after {op}"


def test_rebuild_noop_cases(orc, ctx):
    announce("rebuild no-op")
    e = orc.new_level_set(1.0, 3.0)
    g = ctx.upload(e.to_nano())
    g.rebuild(3.0)
    assert g.counts() == (0, 0)
    # all-positive band: no crossings -> grid unchanged, empty leaves kept
    p = orc.new_grid(3.0, 1.0)
    for i in range(10):
        p.set(i, 0, 0, 1.0 + i)
    p.set(20, 0, 0, 1.0)
    p.deactivate(20, 0, 0)
    b = p.to_nano()
    g = ctx.upload(b)
    g.rebuild(3.0)
    assert_nano_equal(g.download(), b, "no crossings")


def test_rebuild_rejects_wide_band(ctx, orc, vr):
    g = ctx.upload(orc.sphere(5.0, (0, 0, 0), 1.0, 3.0).to_nano())
    with pytest.raises(vr.VolrusError) as ei:
        g.rebuild(8.5)
    assert ei.value.code == vr.VR_ERR_UNSUPPORTED


# ---------------------------------------------------------------- marching cubes
@pytest.mark.parametrize("case", ["r5", "sphere128", "sphere128_advected", "slotted", "offset_negative"])
def test_marching_cubes_matches(orc, ctx, case):
    announce(f"marching cubes {case}")
    if case == "r5":
        og, vs = orc.sphere(5.0, (0, 0, 0), 1.0, 3.0), 1.0
    elif case == "sphere128":
        og, vs = sphere128(orc), 1.0 / 128
    elif case == "sphere128_advected":
        vs = 1.0 / 128
        og = orc.advect_analytic(sphere128(orc), 3, (0, 0, 2 * math.pi, 0.5, 0.5, 0.5), sphere128_rigid_dt())
    elif case == "slotted":
        og, vs = slotted_sphere(orc, 1.0 / 128), 1.0 / 128
    else:
        og, vs = orc.sphere(3.0, (-10.0, 20.0, -30.0), 0.5, 3.0), 0.5
    vw, tw, wt = orc.mesh(og, 0.0)
    vg, tg = ctx.upload(og.to_nano()).mesh(0.0)
    mesh_equal(vg, tg, vw, tw, vs, case)
    assert np.array_equal(vg.view(np.uint64), vw.view(np.uint64)), f"{case}: vertices not bit-exact"


def test_marching_cubes_white_noise(orc, ctx):
    """Random values instead of a distance field: rounds too full for the
    per-round lists (written cell by cell) and tiles with more edges owned by
    earlier tiles than the side table holds (resolved again per corner)."""
    announce("marching cubes white noise")
    og = white_noise_block(orc, 7)
    vw, tw, _ = orc.mesh(og, 0.0)
    vg, tg = ctx.upload(og.to_nano()).mesh(0.0)
    assert len(tw) > 40 * 512  # dense: ~3 triangles per cell
    mesh_equal(vg, tg, vw, tw, 1.0, "white noise")
    assert np.array_equal(vg.view(np.uint64), vw.view(np.uint64)), "white noise: vertices not bit-exact"


def test_marching_cubes_empty(orc, ctx):
    v, t = ctx.upload(orc.new_grid(3.0, 1.0).to_nano()).mesh(0.0)
    assert len(v) == 0 and len(t) == 0


def test_counts(orc, ctx):
    og = sphere128(orc)
    assert ctx.upload(og.to_nano()).counts() == (og.leaf_count(), og.active_count())


# ---------------------------------------------------------------- device fixture generation
def test_fixture_sphere_matches_reference_generator(orc, ctx, vr):
    """vr_grid_from_primitives(sphere) == make_level_set_sphere (sphere.rs:17-59)."""
    announce("fixture sphere")
    for r, c, vs, hw in [(0.3, (0.5, 0.5, 0.5), 1.0 / 128, 3.0), (5.0, (0, 0, 0), 1.0, 3.0),
                         (3.0, (10.0, -20.0, 30.0), 0.5, 2.0)]:
        want = orc.sphere(r, c, vs, hw).to_nano()
        got = ctx.from_primitives([(vr.PRIM_SPHERE, 0, (*c, r))], vs, hw).download()
        assert_nano_equal(got, want, f"sphere r={r}")


def test_fixture_scenes_match_csg_chain(orc, ctx, vr):
    """Slotted sphere (csg_difference) and sphere unions (csg_union chain)."""
    announce("fixture scenes")
    vs = 1.0 / 128
    want = slotted_sphere(orc, vs).to_nano()
    got = ctx.from_primitives([(vr.PRIM_SPHERE, 0, (0.5, 0.75, 0.5, 0.15)),
                               (vr.PRIM_BOX, vr.CSG_DIFFERENCE, (0.475, 0.55, 0.3, 0.525, 0.725, 0.7))],
                              vs, 3.0).download()
    assert_nano_equal(got, want, "slotted")
    rng = np.random.default_rng(11)
    prims = [(0, tuple(rng.uniform(0.2, 0.8, 3)) + (float(np.exp(rng.uniform(np.log(0.02), np.log(0.2)))),))
             for _ in range(5)]
    want = orc.scene([(k, p) for k, p in prims], [0] * len(prims), vs, 3.0).to_nano()
    got = ctx.from_primitives([(k, vr.CSG_UNION, p) for k, p in prims], vs, 3.0).download()
    assert_nano_equal(got, want, "union")
    noisy = [(2, (0.5, 0.5, 0.5, 0.3, 0.02, 4.0, 7.0))]
    want = orc.scene(noisy, [0], vs, 3.0).to_nano()
    got = ctx.from_primitives([(vr.PRIM_NOISY_SPHERE, 0, noisy[0][1])], vs, 3.0).download()
    assert_nano_equal(got, want, "noisy")


# ---------------------------------------------------------------- rollout
def oracle_rollout(orc, g, kind, params, period, t0, dt, scheme, steps, rebuild_every, hw, mesh_every):
    """advect -> rebuild every k -> marching cubes every m, composed from the oracle's tools."""
    updates, mesh = 0, None
    for s in range(steps):
        tf = orc.stage_factors(period, t0 + s * dt, dt, scheme)
        updates += g.active_count()
        g = orc.advect_analytic(g, kind, params, dt, scheme, tf)
        if rebuild_every and (s + 1) % rebuild_every == 0:
            orc.rebuild(g, hw)
        if mesh_every and (s + 1) % mesh_every == 0:
            mesh = orc.mesh(g, 0.0)
    return g, updates, mesh


@pytest.mark.parametrize("shared", [False, True])
def test_rollout_matches_oracle_composition(orc, ctx, vr, shared):
    """vr_rollout == the oracle's advect/rebuild/MC chain, over repeated calls
    on one context (reused ping buffer) and on a grid whose lookup arrays are
    shared with another grid (an advect output)."""
    announce(f"rollout shared={shared}")
    vs, hw, period = 1.0 / 64, 3.0, 8.0
    dt = 1.0 * vs
    steps, rb, me = 9, 3, 4
    og = slotted_sphere(orc, vs)
    vel = ctx.velocity_analytic(vr.VEL_VORTEX, (), period)
    src = ctx.upload(og.to_nano())
    base, t0 = src, 0.0
    if shared:
        base = src.advect(vel, dt, 0.0, 3)  # shares src's lookup arrays
        og = orc.advect_analytic(og, 4, (), dt, 3, orc.stage_factors(period, 0.0, dt, 3))
        t0 = dt
    want, updates, wmesh = oracle_rollout(orc, og, 4, (), period, t0, dt, 3, steps, rb, hw, me)
    want_b = want.to_nano()
    mesh = vr.Mesh(ctx)
    for rep in range(3):
        g = base.clone() if rep < 2 else base
        st = g.rollout(vel, t0, dt, vr.RK3, steps, rb, hw, me, 0.0, mesh)
        assert_nano_equal(g.download(), want_b, f"rollout rep {rep}")
        assert st.voxel_updates == updates
        assert (st.advect_steps, st.rebuilds, st.meshes) == (steps, steps // rb, steps // me)
        assert (st.final_leaf_count, st.final_active_count) == (want.leaf_count(), want.active_count())
        vg, tg = mesh.arrays()
        mesh_equal(vg, tg, wmesh[0], wmesh[1], vs, f"rollout mesh rep {rep}")
        assert (st.last_vertex_count, st.last_triangle_count) == (len(wmesh[0]), len(wmesh[1]))


def test_empty_grid_through_every_pass(orc, ctx, vr):
    """A level set without leaves (level_set_track.rs:265, volume_to_mesh.rs:582
    test it per tool): advect (RK1 and RK3, analytic and dense fields),
    rebuild, trim and a rollout with meshes all give the oracle's empty grid."""
    announce("empty grid every pass")
    og = orc.new_level_set(1.0 / 64, 3.0)
    nano = og.to_nano()
    assert (len(nano) - 320) // 2128 == 0
    n = 8
    u = np.random.default_rng(3).normal(0.0, 0.5, size=(3, n, n, n)).astype(np.float32)
    dense = ctx.velocity_dense(u, n, 1.0 / n)
    vortex = ctx.velocity_analytic(vr.VEL_VORTEX, (), 8.0)
    for scheme in (1, 3):
        want = orc.advect_dense(og, u, n, 1.0 / n, 0.01, scheme).to_nano()
        assert_nano_equal(ctx.upload(nano).advect(dense, 0.01, 0.0, scheme).download(), want, f"dense {scheme}")
        tf = orc.stage_factors(8.0, 0.0, 0.01, scheme)
        want = orc.advect_analytic(og, 4, (), 0.01, scheme, tf).to_nano()
        assert_nano_equal(ctx.upload(nano).advect(vortex, 0.01, 0.0, scheme).download(), want, f"vortex {scheme}")
    for hw in (3.0, 1.5):
        ref = orc.from_nano(nano)
        orc.rebuild(ref, hw)
        g = ctx.upload(nano)
        g.rebuild(hw)
        assert_nano_equal(g.download(), ref.to_nano(), f"rebuild {hw}")
        ref = orc.from_nano(nano)
        orc.trim(ref, hw)
        g = ctx.upload(nano)
        g.trim(hw)
        assert_nano_equal(g.download(), ref.to_nano(), f"trim {hw}")
    g = ctx.upload(nano)
    st = g.rollout(vortex, 0.0, 0.01, vr.RK3, 6, 2, 3.0, 3, 0.0, vr.Mesh(ctx))
    assert_nano_equal(g.download(), nano, "rollout")
    assert st.voxel_updates == 0 and st.final_leaf_count == 0
    assert (st.last_vertex_count, st.last_triangle_count) == (0, 0)
    assert g.counts() == (0, 0)


def test_rollout_timing_records(orc, ctx, vr):
    """Event timing of a rollout: every advect launch is counted (the runs of
    back-to-back advects between rebuild and mesh passes share one event
    pair), and their algorithmic bytes are 2 x 2128 per leaf of the segment."""
    announce("rollout timing")
    vs, hw, period = 1.0 / 64, 3.0, 8.0
    g = ctx.upload(slotted_sphere(orc, vs).to_nano())
    leaves0 = g.counts()[0]
    vel = ctx.velocity_analytic(vr.VEL_VORTEX, (), period)
    ctx.set_timing(True)
    try:
        st = g.rollout(vel, 0.0, vs, vr.RK3, 9, 3, hw, 4, 0.0, vr.Mesh(ctx))
        ms, n, b = ctx.timing(vr.TIMING_ADVECT)
        rms, rn, _ = ctx.timing(vr.TIMING_REBUILD)
    finally:
        ctx.set_timing(False)
    assert n == st.advect_steps == 9 and rn == st.rebuilds == 3
    assert ms > 0.0 and rms > 0.0
    assert b % (2 * 2128) == 0
    assert 9 * 2 * 2128 * st.final_leaf_count <= b <= 9 * 2 * 2128 * leaves0


def test_rollout_mesh_capacity_regrowth(orc, ctx, vr):
    """A mesh larger than a fresh mesh's first capacity guess (16 vertices per
    leaf): the rollout grows the buffers and meshes the final state again, and
    the next rollout fits; both equal the oracle."""
    announce("rollout mesh regrowth")
    og = sphere128(orc)
    nano = og.to_nano()
    leaves = (len(nano) - 320) // 2128
    params = (0.0, 0.0, 2 * math.pi, 0.5, 0.5, 0.5)
    dt = sphere128_rigid_dt()
    steps, rb, me = 4, 2, 2
    want, updates, wmesh = oracle_rollout(orc, og, 3, params, 0.0, 0.0, dt, 3, steps, rb, 3.0, me)
    assert len(wmesh[0]) > 16 * max(leaves, 256)
    vel = ctx.velocity_analytic(vr.VEL_RIGID, params, 0.0)
    base = ctx.upload(nano)
    mesh = vr.Mesh(ctx)
    for rep in range(2):
        g = base.clone()
        st = g.rollout(vel, 0.0, dt, vr.RK3, steps, rb, 3.0, me, 0.0, mesh)
        assert_nano_equal(g.download(), want.to_nano(), f"regrowth rollout rep {rep}")
        assert st.voxel_updates == updates
        vg, tg = mesh.arrays()
        mesh_equal(vg, tg, wmesh[0], wmesh[1], 1.0 / 128, f"regrowth mesh rep {rep}")
        assert (st.last_vertex_count, st.last_triangle_count) == (len(wmesh[0]), len(wmesh[1]))


def test_rollout_last_mesh_before_last_step(orc, ctx, vr):
    """Meshes at steps 2 and 4 of a 5-step rollout into a fresh mesh, whose
    first capacity guess the mesh exceeds: the returned mesh is the state
    after step 4 (meshed again at that step), not the final state."""
    announce("rollout last mesh before last step")
    og = sphere128(orc)
    nano = og.to_nano()
    leaves = (len(nano) - 320) // 2128
    params = (0.0, 0.0, 2 * math.pi, 0.5, 0.5, 0.5)
    dt = sphere128_rigid_dt()
    steps, rb, me = 5, 2, 2
    want, updates, wmesh = oracle_rollout(orc, og, 3, params, 0.0, 0.0, dt, 3, steps, rb, 3.0, me)
    assert len(wmesh[0]) > 16 * max(leaves, 256)
    vel = ctx.velocity_analytic(vr.VEL_RIGID, params, 0.0)
    base = ctx.upload(nano)
    for rep in range(2):
        mesh = vr.Mesh(ctx)
        g = base.clone()
        st = g.rollout(vel, 0.0, dt, vr.RK3, steps, rb, 3.0, me, 0.0, mesh)
        assert_nano_equal(g.download(), want.to_nano(), f"rollout rep {rep}")
        assert st.voxel_updates == updates and st.meshes == 2
        vg, tg = mesh.arrays()
        mesh_equal(vg, tg, wmesh[0], wmesh[1], 1.0 / 128, f"last mesh rep {rep}")
        assert (st.last_vertex_count, st.last_triangle_count) == (len(wmesh[0]), len(wmesh[1]))


# ---------------------------------------------------------------- FP32 error against the FP64 oracle
@pytest.mark.parametrize("kind,params,scheme,steps", [
    (3, (0, 0, 2 * math.pi, 0.5, 0.5, 0.5), 1, 1),   # rigid rotation, one RK1 step
    (4, (), 3, 4),                                    # single vortex, four RK3 steps
])
def test_fp32_error_against_fp64_oracle(orc, ctx, vr, kind, params, scheme, steps):
    """The GPU's f32 values after `steps` advection steps against the same
    steps carried out in f64 by the oracle: the error, in voxels, stays below
    1e-5 (the stated tolerance) and is printed."""
    announce(f"fp32 vs fp64 kind={kind} scheme={scheme} steps={steps}")
    vs = 1.0 / 128
    og = sphere128(orc)
    dt, period = sphere128_rigid_dt(), (8.0 if kind == 4 else 0.0)
    vel = ctx.velocity_analytic(kind, params, period)
    g = ctx.upload(og.to_nano())
    g64 = orc.to_f64(og)
    for s in range(steps):
        tf = orc.stage_factors(period, s * dt, dt, scheme)
        g = g.advect(vel, dt, s * dt, scheme)
        g64 = orc.advect_analytic_f64(g64, kind, params, dt, scheme, tf)
    c, v = active_voxels(g.download())
    ref = np.array([g64.get(*xyz) for xyz in c.tolist()])
    err = float(np.abs(v.astype(np.float64) - ref).max()) / vs
    print(f"[gpu] max |f32 - f64| = {err:.3e} voxel over {len(v)} active voxels", flush=True)
    assert err < 1e-5
