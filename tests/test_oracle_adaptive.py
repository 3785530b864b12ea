"""Pins the oracle's adaptive_surface (src/tools/adaptive_mesh.rs:100-241) and
the differential operators it uses (src/tools/gradient.rs:13-135): the
reference's own tests transcribed (gradient.rs:137-247,
adaptive_mesh.rs:273-406), central differences that are exact on quadratics
checked against the closed-form mean curvature, and count invariants of the
1-to-4 subdivision with its shared-edge midpoints."""
import numpy as np
import pytest


def ramp(orc):
    g = orc.new_grid(0.0, 1.0)
    for x in range(-5, 6):
        for y in range(-5, 6):
            for z in range(-5, 6):
                g.set(x, y, z, 3.0 * x + 5.0 * y + 7.0 * z)
    return g


def quadric(orc, coef, n=8):
    """f = a x^2 + b y^2 + c z^2 + d xy + e yz + f xz sampled at integer points."""
    a, b, c, d, e, f = coef
    g = orc.new_grid(0.0, 1.0)
    for x in range(-n, n + 1):
        for y in range(-n, n + 1):
            for z in range(-n, n + 1):
                g.set(x, y, z, float(np.float32(a * x * x + b * y * y + c * z * z + d * x * y + e * y * z + f * x * z)))
    return g


# ---------------------------------------------------------------- reference tests (gradient.rs)
def test_gradient_of_linear_ramp_is_constant(orc):  # gradient.rs:172
    g = ramp(orc)
    for p in [(0, 0, 0), (2, -3, 1)]:
        assert np.allclose(orc.gradient_at(g, *p), (3.0, 5.0, 7.0), atol=1e-5)


def test_laplacian_of_quadratic_is_constant(orc):  # gradient.rs:190
    g = quadric(orc, (1, 2, 3, 0, 0, 0), 10)
    assert abs(orc.laplacian_at(g, 0, 0, 0) - 12.0) < 1e-4 and abs(orc.laplacian_at(g, 3, -2, 1) - 12.0) < 1e-4


def test_laplacian_of_linear_is_zero(orc):  # gradient.rs:203
    assert abs(orc.laplacian_at(ramp(orc), 0, 0, 0)) < 1e-5


def test_mean_curvature_of_sphere(orc):  # gradient.rs:210
    g = orc.sphere(10.0, (0.0, 0.0, 0.0), 0.5, 3.0)
    assert abs(orc.mean_curvature_at(g, 20, 0, 0) - 2.0 / 20.0) < 0.05


# ---------------------------------------------------------------- closed form on quadrics
@pytest.mark.parametrize("coef", [(1, 2, 3, 0, 0, 0), (1, 2, 3, 1, 0.5, -0.75), (0.5, -1, 2, 2, 0, 1)])
def test_mean_curvature_closed_form_on_quadrics(orc, coef):
    """Central differences are exact on quadratics, so mean_curvature_at
    returns div(grad f / |grad f|) of the quadric up to f32 rounding: a
    dropped term or cross derivative, a wrong sign or factor fails."""
    a, b, c, d, e, f = coef
    g = quadric(orc, coef)
    for x, y, z in [(1, 2, 3), (-3, 1, 2), (2, -2, -4), (4, 0, 1)]:
        fx, fy, fz = 2 * a * x + d * y + f * z, 2 * b * y + d * x + e * z, 2 * c * z + e * y + f * x
        fxx, fyy, fzz, fxy, fyz, fxz = 2 * a, 2 * b, 2 * c, d, e, f
        num = (fxx * (fy * fy + fz * fz) + fyy * (fx * fx + fz * fz) + fzz * (fx * fx + fy * fy)
               - 2 * (fx * fy * fxy + fx * fz * fxz + fy * fz * fyz))
        want = num / (fx * fx + fy * fy + fz * fz) ** 1.5
        got = orc.mean_curvature_at(g, x, y, z)
        assert abs(got - want) <= 2e-5 * max(1.0, abs(want)), (x, y, z, got, want)
        assert np.allclose(orc.gradient_at(g, x, y, z), (fx, fy, fz), rtol=1e-6, atol=1e-5)


# ---------------------------------------------------------------- reference tests (adaptive_mesh.rs)
def sphere_grid(orc, r, vs):
    return orc.sphere(r, (0.0, 0.0, 0.0), vs, 3.0)


def test_adaptive_sphere_more_tris_than_uniform(orc):  # adaptive_mesh.rs:278
    g = sphere_grid(orc, 5.0, 1.0)
    u = orc.adaptive_surface(g, 0.0, 2.0, 0.01, 0)
    a = orc.adaptive_surface(g, 0.0, 2.0, 0.01, 2)
    assert len(a[2]) > len(u[2])


def test_normals_are_unit_length(orc):  # adaptive_mesh.rs:307
    _, n, _, _ = orc.adaptive_surface(sphere_grid(orc, 5.0, 1.0), 0.0, 1.0, 0.01, 1)
    ln = np.linalg.norm(n, axis=1)
    assert np.all((np.abs(ln - 1.0) < 0.01) | (ln < 1e-10))


def test_curvatures_on_sphere_approx_inverse_radius(orc):  # adaptive_mesh.rs:327
    _, _, _, c = orc.adaptive_surface(sphere_grid(orc, 10.0, 0.5), 0.0, 1.0, 0.5, 0)
    expected = 2.0 / 10.0
    assert np.mean(np.abs(np.abs(c) - expected) < expected * 2.0) > 0.3


def test_depth_zero_same_as_standard_mc(orc):  # adaptive_mesh.rs:355
    g = sphere_grid(orc, 5.0, 1.0)
    v, n, t, c = orc.adaptive_surface(g, 0.0, 1.0, 0.5, 0)
    vm, tm, _ = orc.mesh(g, 0.0)
    assert np.array_equal(t, tm) and np.array_equal(v, vm)


def test_render_mesh_counts_consistent(orc):  # adaptive_mesh.rs:375
    v, n, t, c = orc.adaptive_surface(sphere_grid(orc, 5.0, 1.0), 0.0, 1.0, 0.01, 1)
    assert len(v) == len(n) == len(c) and len(t) > 0 and t.max() < len(v)


def test_empty_grid_produces_empty_mesh(orc):  # adaptive_mesh.rs:396
    v, n, t, c = orc.adaptive_surface(orc.new_grid(3.0, 1.0), 0.0, 1.0, 0.5, 3)
    assert len(v) == len(n) == len(t) == len(c) == 0


# ---------------------------------------------------------------- subdivision invariants
def test_full_subdivision_counts(orc):
    """Threshold 0 and a tiny edge bound split every triangle of positive
    curvature: 4 triangles each, and one new vertex per distinct edge of the
    split triangles (shared edges get one midpoint)."""
    g = sphere_grid(orc, 5.0, 1.0)
    vm, tm, _ = orc.mesh(g, 0.0)
    v, n, t, c = orc.adaptive_surface(g, 0.0, 1e-6, 0.0, 1)
    rnd = np.sign(vm) * np.floor(np.abs(vm) + 0.5)  # f64::round: half away from zero
    curv = np.abs(np.array([orc.mean_curvature_at(g, *p.astype(int).tolist()) for p in rnd]))
    split = np.max(curv[tm], axis=1) > 0.0
    edges = set()
    for tri in tm[split]:
        for a, b in [(tri[0], tri[1]), (tri[1], tri[2]), (tri[2], tri[0])]:
            edges.add((min(a, b), max(a, b)))
    assert len(t) == 4 * int(split.sum()) + int((~split).sum())
    assert len(v) == len(vm) + len(edges)
    assert np.array_equal(v[: len(vm)], vm)
    # the first children of the first split triangle, in the reference's order
    i = int(np.nonzero(split)[0][0])
    k = int(np.nonzero(~split[:i])[0].size)  # unsplit triangles before it keep one slot each
    t0, t1, t2 = tm[i]
    m01, m12, m20 = t[k + 3]
    assert list(t[k]) == [t0, m01, m20] and list(t[k + 1]) == [t1, m12, m01] and list(t[k + 2]) == [t2, m20, m12]


def test_projected_midpoints_approach_the_surface(orc):
    """New vertices are edge midpoints moved along the SDF gradient by the
    sampled value: on a sphere band they sit closer to the zero level than
    the plain midpoints."""
    g = sphere_grid(orc, 8.0, 1.0)
    vm, tm, _ = orc.mesh(g, 0.0)
    v, n, t, c = orc.adaptive_surface(g, 0.0, 1e-6, 0.0, 1)
    new = v[len(vm):]
    r = np.linalg.norm(new, axis=1)
    outer = r > 6.5  # the outer shell (the inner one is the band's spurious shell)
    assert np.median(np.abs(r[outer] - 8.0)) < 0.05
    # radial normals on the outer shell
    d = np.abs(np.sum(n[len(vm):][outer] * (new[outer] / r[outer, None]), axis=1))
    assert np.median(d) > 0.99
