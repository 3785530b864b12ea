"""Pins the CPU oracle where no reference test does, by mathematics:
closed forms, exact reductions, invariants, brute force and finite
differences, each chosen so a plausible slip (a dropped term, a wrong sign or
index, a transposed operand) fails.
"""
import math
import os
import re

import numpy as np
import pytest

from conftest import ROOT
from util import active_voxels, sphere128_rigid_dt, sphere128, leaf_array, slotted_sphere

REF_MC = "/root/reference/src/tools/volume_to_mesh.rs"


def decoded_tri_table():
    text = open(os.path.join(ROOT, "sharc.volrus_amd", "csrc", "mc_tables.h")).read()
    body = text[text.index("MC_TRI[256] = {"):]
    body = body[: body.index("};")]
    strs = re.findall(r'"([0-9a-b]*)"', body)
    assert len(strs) == 256
    return [[int(ch, 16) for ch in s] for s in strs]


MC_CORNER = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
MC_EDGE = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]


def edge_mask(case):
    return sum(1 << e for e, (a, b) in enumerate(MC_EDGE) if ((case >> a) & 1) != ((case >> b) & 1))


# ---------------------------------------------------------------- MC tables
def test_mc_tables_self_consistent():
    """Triangle edges of every case are exactly the sign-change edges, each
    triangle has three distinct edges, and at most 5 triangles per case."""
    tri = decoded_tri_table()
    hist = {}
    for c in range(256):
        t = tri[c]
        assert len(t) % 3 == 0 and len(t) <= 15
        used = 0
        for i in range(0, len(t), 3):
            assert len(set(t[i:i + 3])) == 3
            for e in t[i:i + 3]:
                used |= 1 << e
        assert used == edge_mask(c), f"case {c}"
        hist[len(t) // 3] = hist.get(len(t) // 3, 0) + 1
    assert hist == {0: 2, 1: 16, 2: 50, 3: 80, 4: 76, 5: 32}


def test_mc_tables_orientation_consistent():
    """Every triangle winds the same way relative to the below-iso corners:
    its normal (right-hand rule over edge midpoints) points out of the region
    where v < iso, for all 256 cases (a transposed triangle fails)."""
    tri = decoded_tri_table()
    mid = [np.add(MC_CORNER[a], MC_CORNER[b]) / 2.0 for a, b in MC_EDGE]

    def grad(vals, p):
        # gradient of the trilinear interpolant of the corner values at p
        gx = gy = gz = 0.0
        for k, (a, b, c) in enumerate(MC_CORNER):
            wx, wy, wz = (p[0] if a else 1 - p[0]), (p[1] if b else 1 - p[1]), (p[2] if c else 1 - p[2])
            gx += vals[k] * (1 if a else -1) * wy * wz
            gy += vals[k] * wx * (1 if b else -1) * wz
            gz += vals[k] * wx * wy * (1 if c else -1)
        return np.array([gx, gy, gz])

    signs = []
    for c in range(1, 255):
        vals = [-1.0 if (c >> k) & 1 else 1.0 for k in range(8)]
        t = tri[c]
        for i in range(0, len(t), 3):
            p0, p1, p2 = (mid[e] for e in t[i:i + 3])
            n = np.cross(p1 - p0, p2 - p0)
            s = np.dot(n, grad(vals, (p0 + p1 + p2) / 3.0))
            signs.append(np.sign(s))
    signs = np.array(signs)
    # all triangles face the same way relative to the field gradient, except
    # where an ambiguous face puts the centroid on a saddle (|s| tiny)
    assert (signs == signs[0]).mean() > 0.97


@pytest.mark.skipif(not os.path.exists(REF_MC), reason="reference source not mounted")
def test_mc_tables_equal_reference_text():
    """The re-encoded table equals the reference's TRI_TABLE and the derived
    edge masks equal its EDGE_TABLE (src/tools/volume_to_mesh.rs:221-503)."""
    src = open(REF_MC).read()
    e0 = src.index("= [", src.index("const EDGE_TABLE"))
    edge_tab = [int(x, 16) for x in re.findall(r"0x([0-9a-f]+)", src[e0:src.index("];", e0)])]
    assert edge_tab == [edge_mask(c) for c in range(256)]
    t0 = src.index("= [", src.index("const TRI_TABLE"))
    rows = re.findall(r"\[([-0-9, ]+)\]", src[t0:src.index("\n];", t0)])
    ref = [[int(x) for x in r.split(",") if int(x) >= 0] for r in rows]
    assert ref == decoded_tri_table()


# ---------------------------------------------------------------- marching cubes invariants
def mesh_stats(v, t):
    edges = set()
    for a, b, c in t.tolist():
        for x, y in ((a, b), (b, c), (c, a)):
            edges.add((min(x, y), max(x, y)))
    return len(v) - len(edges) + len(t)


@pytest.mark.parametrize("r,vs", [(5.0, 1.0), (0.3, 1.0 / 128)])
def test_mc_sphere_band_two_closed_shells(orc, r, vs):
    """The +background outside the band makes a second, inner shell: two
    closed genus-0 surfaces, so the Euler characteristic is 4 and every
    edge is shared by two triangles."""
    g = orc.sphere(r, (0.5, 0.5, 0.5) if vs < 1 else (0, 0, 0), vs, 3.0)
    v, t, wt = orc.mesh(g, 0.0)
    assert wt
    assert mesh_stats(v, t) == 4


def test_mc_vertices_on_lattice_edges_at_zero(orc):
    """Each vertex lies on a lattice edge (two coordinates on the lattice) and
    the trilinear interpolant of the grid vanishes there (iso = 0)."""
    vs = 1.0 / 128
    g = sphere128(orc)
    v, t, _ = orc.mesh(g, 0.0)
    idx = v / vs
    on = np.abs(idx - np.round(idx)) < 1e-9
    assert np.all(on.sum(axis=1) >= 2)
    for p in v[:: max(1, len(v) // 400)]:
        assert abs(g.sample(*p)) < 1e-6


def test_mc_single_cell_cases(orc):
    """Brute force over all 256 corner sign patterns of one cell: the number
    of triangles equals the table's, every vertex lies on a sign-change edge
    at the linear zero, and triangles only use the cell's own vertices."""
    tri = decoded_tri_table()
    for c in range(256):
        g = orc.new_grid(5.0, 1.0)  # background far above iso
        vals = []
        for k, (x, y, z) in enumerate(MC_CORNER):
            val = -1.0 - 0.25 * k if (c >> k) & 1 else 1.0 + 0.5 * k
            vals.append(val)
            g.set(8 + x, 8 + y, 8 + z, val)
        v, t, _ = orc.mesh(g, 0.0)
        # the cell at base (8,8,8) is one of many processed; pick its vertices
        sel = np.all((v >= 8 - 1e-12) & (v <= 9 + 1e-12), axis=1)
        assert sel.sum() == bin(edge_mask(c)).count("1"), f"case {c}"
        for p in v[sel]:
            for e, (a, b) in enumerate(MC_EDGE):
                pa, pb = np.add(8, MC_CORNER[a]), np.add(8, MC_CORNER[b])
                d = pb - pa
                tpar = np.dot(p - pa, d)
                if np.allclose(pa + tpar * d, p) and 0 <= tpar <= 1:
                    lin = vals[a] + tpar * (vals[b] - vals[a])
                    assert abs(lin) < 1e-6
        in_cell = set(np.nonzero(sel)[0].tolist())
        n_cell = sum(1 for tr in t.tolist() if set(tr) <= in_cell)
        assert n_cell >= len(tri[c]) // 3


# ---------------------------------------------------------------- trilinear / advection
def test_sincospi_matches_libm(orc):
    xs = np.concatenate([np.linspace(-3, 3, 2001), [0.0, 0.25, 0.5, 1.0, -0.5, 1e-9, 2.75]])
    for x in xs:
        s, c = orc.sincospi(float(x))
        assert abs(s - math.sin(math.pi * x)) < 2e-15 + 4e-15 * abs(x)
        assert abs(c - math.cos(math.pi * x)) < 2e-15 + 4e-15 * abs(x)
    assert orc.sincospi(0.5) == (1.0, 0.0) or abs(orc.sincospi(0.5)[1]) < 1e-300


def test_integer_shift_is_exact_copy(orc):
    """Uniform velocity moving exactly one voxel per step: all trilinear
    weights are 0 or 1, so the advected value is the neighbour's value."""
    vs = 1.0 / 64
    g = orc.sphere(0.3, (0.5, 0.5, 0.5), vs, 3.0)
    dt = 0.25
    u = (vs / dt, 0.0, -2 * vs / dt)  # one voxel in +x, two in -z
    out = orc.advect_analytic(g, 2, u, dt)
    c, val = active_voxels(out.to_nano())
    for (x, y, z), vv in list(zip(c.tolist(), val.tolist()))[::37]:
        assert vv == g.get(x - 1, y, z + 2)


def planar_grid(orc, vs, lo, hi, axis):
    g = orc.new_grid(0.0, vs)
    for x in range(lo, hi):
        for y in range(lo, hi):
            for z in range(lo, hi):
                g.set(x, y, z, float((x, y, z)[axis]) * vs)
    return g


def test_planar_field_gives_departure_points(orc):
    """A linear field is reproduced exactly by trilinear interpolation, so
    advecting phi = x returns the x coordinate of the departure point; for
    RK1 that is x - dt u(x) (advection.rs:40-48)."""
    vs = 1.0 / 32
    g = orc.to_f64(planar_grid(orc, vs, 8, 24, 0))
    w = (0.0, 0.0, 2 * math.pi, 0.5, 0.5, 0.5)
    dt = 0.2 * vs
    out = orc.advect_analytic_f64(g, 3, w, dt, 1)
    for x, y, z in [(12, 13, 14), (16, 16, 16), (20, 11, 15)]:
        wx, wy = x * vs, y * vs
        ux = -2 * math.pi * (wy - 0.5)
        assert abs(out.get(x, y, z) - (wx - dt * ux)) < 1e-14


def departure_error(orc, scheme, dt, vs=1.0 / 32):
    gx = orc.to_f64(planar_grid(orc, vs, 4, 28, 0))
    gy = orc.to_f64(planar_grid(orc, vs, 4, 28, 1))
    w = (0.0, 0.0, 2 * math.pi, 0.5, 0.5, 0.5)
    ox = orc.advect_analytic_f64(gx, 3, w, dt, scheme)
    oy = orc.advect_analytic_f64(gy, 3, w, dt, scheme)
    err = 0.0
    for x, y, z in [(10, 12, 16), (20, 18, 16), (14, 21, 9)]:
        px, py = x * vs - 0.5, y * vs - 0.5
        a = -2 * math.pi * dt  # exact back-rotation
        ex = 0.5 + px * math.cos(a) - py * math.sin(a)
        ey = 0.5 + px * math.sin(a) + py * math.cos(a)
        err = max(err, abs(ox.get(x, y, z) - ex), abs(oy.get(x, y, z) - ey))
    return err


def test_rk3_is_third_order_rk1_first_order(orc):
    """Local error of the back-trace on rigid rotation: RK1 O(dt^2), RK3 O(dt^4)."""
    dt = 0.004
    e1a, e1b = departure_error(orc, 1, dt), departure_error(orc, 1, dt / 2)
    e3a, e3b = departure_error(orc, 3, dt), departure_error(orc, 3, dt / 2)
    assert 3.5 < e1a / e1b < 4.5
    assert 12.0 < e3a / e3b < 20.0
    assert e3a < 1e-3 * e1a


def test_rk3_uniform_velocity_reduces_to_rk1(orc):
    g = sphere128(orc)
    u = (0.3, -0.7, 0.2)
    dt = sphere128_rigid_dt()
    a = active_voxels(orc.advect_analytic(g, 2, u, dt, 1).to_nano())[1]
    b = active_voxels(orc.advect_analytic(g, 2, u, dt, 3).to_nano())[1]
    assert np.max(np.abs(a - b)) < 2e-7


def test_dense_velocity_equals_reference_velocity_field(orc):
    """The dense field samples exactly like VelocityField of three grids with
    background 0 (advection.rs:10-24 + interpolation.rs:11-48)."""
    n = 8
    rng = np.random.default_rng(3)
    u = rng.normal(0, 0.5, size=(3, n, n, n)).astype(np.float32)
    grids = []
    for c in range(3):
        vg = orc.new_grid(0.0, 1.0 / n)
        for i in range(n):
            for j in range(n):
                for k in range(n):
                    vg.set(i, j, k, float(u[c, i, j, k]))
        grids.append(vg)
    g = orc.sphere(0.3, (0.5, 0.5, 0.5), 1.0 / 32, 3.0)
    dt = 0.05
    a = orc.advect_dense(g, u, n, 1.0 / n, dt, 1).to_nano()
    b = orc.advect_grids3(g, *grids, dt).to_nano()
    assert a == b


def test_time_factor(orc):
    tf = orc.stage_factors(8.0, 1.0, 0.5, 3)
    assert abs(tf[0] - math.cos(math.pi * 1.5 / 8)) < 1e-15
    assert abs(tf[1] - math.cos(math.pi * 1.0 / 8)) < 1e-15
    assert abs(tf[2] - math.cos(math.pi * 1.25 / 8)) < 1e-15
    assert list(orc.stage_factors(0.0, 1.0, 0.5, 3)) == [1.0, 1.0, 1.0]


# ---------------------------------------------------------------- band rebuild
@pytest.mark.parametrize("case", ["r5", "r4_off", "slotted64", "advected"])
def test_rebuild_bounded_equals_brute_force(orc, case):
    """The bounded-radius window (the device algorithm) gives exactly the
    literal O(A x M) rebuild (level_set_track.rs:80-114)."""
    if case == "r5":
        g = orc.sphere(5.0, (0, 0, 0), 1.0, 3.0)
    elif case == "r4_off":
        g = orc.sphere(4.0, (1.5, -2.0, 0.25), 0.5, 3.0)
    elif case == "slotted64":
        g = slotted_sphere(orc, 1.0 / 64)
    else:
        g = orc.advect_analytic(orc.sphere(0.3, (0.5, 0.5, 0.5), 1.0 / 64, 3.0), 4, (), 0.5 / 64, 3,
                                orc.stage_factors(8.0, 0.0, 0.5 / 64, 3))
    for hw in (3.0, 2.5, 1.0):
        b = g.to_nano()
        a1, a2 = orc.from_nano(b), orc.from_nano(b)
        orc.rebuild(a1, hw)
        orc.rebuild_bounded(a2, hw)
        assert a1.to_nano() == a2.to_nano(), f"{case} hw={hw}"


def test_rebuild_values_are_lattice_distances(orc):
    """Kept values are sign * sqrt(integer) * vs with |value| < hw * vs, and
    crossing voxels become +-0."""
    g = orc.sphere(5.0, (0, 0, 0), 1.0, 3.0)
    cross = set(map(tuple, orc.zero_crossings(g).tolist()))
    orc.rebuild(g, 3.0)
    c, v = active_voxels(g.to_nano())
    d2 = (v.astype(np.float64)) ** 2
    assert np.all(np.abs(d2 - np.round(d2)) < 1e-4)
    assert np.all(np.abs(v) < 3.0)
    for xyz, val in zip(map(tuple, c.tolist()), v.tolist()):
        if xyz in cross:
            assert val == 0.0


# ---------------------------------------------------------------- adjoint
@pytest.mark.parametrize("scheme", [1, 3])
def test_adjoint_matches_central_differences(orc, scheme):
    """FP64 adjoint of a K-step RK1/RK3 rollout through a dense trilinear
    velocity vs central differences (h = 1e-3) on random velocity entries."""
    n = 6
    rng = np.random.default_rng(7)
    u = rng.normal(0, 0.5, size=3 * n ** 3)
    vs = 1.0 / 32
    g = orc.sphere(0.3, (0.5, 0.5, 0.5), vs, 3.0)
    target = orc.sphere(0.3, (0.6, 0.5, 0.5), vs, 3.0)
    dt = 0.5 * vs / np.abs(u).max()
    steps = 4
    loss, grad = orc.rollout_adjoint_f64(g, u, n, 1.0 / n, dt, scheme, steps, target)
    assert abs(loss - orc.rollout_loss_f64(g, u, n, 1.0 / n, dt, scheme, steps, target)) < 1e-12 * max(1.0, loss)
    h = 1e-3
    idx = rng.choice(np.nonzero(np.abs(grad) > 1e-6 * np.abs(grad).max())[0], 12, replace=False)
    for i in idx:
        up, dn = u.copy(), u.copy()
        up[i] += h
        dn[i] -= h
        fd = (orc.rollout_loss_f64(g, up, n, 1.0 / n, dt, scheme, steps, target) -
              orc.rollout_loss_f64(g, dn, n, 1.0 / n, dt, scheme, steps, target)) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-3 * abs(grad).max() + 1e-2 * abs(fd), (i, fd, grad[i])
