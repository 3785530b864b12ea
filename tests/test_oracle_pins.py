"""Closed-form pins of the oracle's builder-defined functions (the parts with
no reference counterpart): the single-vortex velocity and its time factor
through the RK1/RK3 back-trace, the exact box distance, 4-octave value noise
and SplitMix64.  Each check is chosen so that a plausible slip (a swapped
axis, a dropped square, a wrong sign, a linear fade, a wrong octave weight, a
wrong constant) fails; none restates the oracle's own formula.
"""
import importlib.util
import math
import os

import numpy as np
import pytest

from conftest import ROOT

# ---------------------------------------------------------------- SplitMix64
# Published outputs of Vigna's reference splitmix64.c (the generator of
# Steele, Lea and Flood, OOPSLA 2014).
SPLITMIX_SEED0 = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC,
                  0x1B39896A51A8749B]
SPLITMIX_SEED1234567 = [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
                        16408922859458223821]


def test_splitmix64_published_vectors(orc):
    assert orc.splitmix64(0, 5) == SPLITMIX_SEED0
    assert orc.splitmix64(1234567, 5) == SPLITMIX_SEED1234567


def test_bench_scene_generator_is_splitmix64():
    """The batched scene generator in bench.py draws from the same published sequence."""
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    r = bench.SplitMix64(1234567)
    assert [r.next() for _ in range(5)] == SPLITMIX_SEED1234567
    r = bench.SplitMix64(0)
    assert r.uniform() == (SPLITMIX_SEED0[0] >> 11) / 2.0 ** 53


# ---------------------------------------------------------------- single vortex
def vortex(x, y):
    """u = (sin^2(pi x) sin(2 pi y), -sin^2(pi y) sin(2 pi x), 0) from libm."""
    return (math.sin(math.pi * x) ** 2 * math.sin(2 * math.pi * y),
            -math.sin(math.pi * y) ** 2 * math.sin(2 * math.pi * x))


def planar_f64(orc, vs, lo, hi, axis):
    g = orc.new_grid(0.0, vs)
    for x in range(lo, hi):
        for y in range(lo, hi):
            for z in range(lo, hi):
                g.set(x, y, z, float((x, y, z)[axis]) * vs)
    return orc.to_f64(g)


VORTEX_PTS = [(12, 13, 14), (16, 16, 16), (20, 11, 15), (9, 22, 12), (21, 21, 9)]


def test_vortex_rk1_departure_closed_form(orc):
    """phi = x and phi = y are reproduced exactly by trilinear sampling, so
    one RK1 step returns the departure point x - dt u(x) of the vortex field
    (advection.rs:40-48 with the builder's analytic field, kind 4)."""
    vs = 1.0 / 32
    gx, gy = planar_f64(orc, vs, 6, 26, 0), planar_f64(orc, vs, 6, 26, 1)
    dt = 0.3 * vs
    ox = orc.advect_analytic_f64(gx, 4, (), dt, 1)
    oy = orc.advect_analytic_f64(gy, 4, (), dt, 1)
    for x, y, z in VORTEX_PTS:
        ux, uy = vortex(x * vs, y * vs)
        assert abs(ox.get(x, y, z) - (x * vs - dt * ux)) < 1e-14
        assert abs(oy.get(x, y, z) - (y * vs - dt * uy)) < 1e-14


def exact_departure(x, y, t, dt, period, sub=400):
    """Backward characteristic of u(x, tau) = vortex(x) cos(pi tau / period)
    from tau = t + dt to tau = t, by classical RK4 with `sub` substeps."""
    def f(p, tau):
        ux, uy = vortex(p[0], p[1])
        c = math.cos(math.pi * tau / period)
        return (-ux * c, -uy * c)

    h = dt / sub
    p = (x, y)
    for i in range(sub):
        tau = t + dt - i * h
        k1 = f(p, tau)
        k2 = f((p[0] + 0.5 * h * k1[0], p[1] + 0.5 * h * k1[1]), tau - 0.5 * h)
        k3 = f((p[0] + 0.5 * h * k2[0], p[1] + 0.5 * h * k2[1]), tau - 0.5 * h)
        k4 = f((p[0] + h * k3[0], p[1] + h * k3[1]), tau - h)
        p = (p[0] + h / 6 * (k1[0] + 2 * k2[0] + 2 * k3[0] + k4[0]),
             p[1] + h / 6 * (k1[1] + 2 * k2[1] + 2 * k3[1] + k4[1]))
    return p


def vortex_step_error(orc, scheme, dt, t=0.3, period=0.5):
    vs = 1.0 / 32
    gx, gy = planar_f64(orc, vs, 6, 26, 0), planar_f64(orc, vs, 6, 26, 1)
    tf = orc.stage_factors(period, t, dt, scheme)
    ox = orc.advect_analytic_f64(gx, 4, (), dt, scheme, tf)
    oy = orc.advect_analytic_f64(gy, 4, (), dt, scheme, tf)
    err = 0.0
    for x, y, z in VORTEX_PTS:
        ex, ey = exact_departure(x * vs, y * vs, t, dt, period)
        err = max(err, abs(ox.get(x, y, z) - ex), abs(oy.get(x, y, z) - ey))
    return err


def test_vortex_time_dependent_rk3_order(orc):
    """With a fast time factor cos(pi tau / 0.5) the stage times matter:
    the RK3 back-trace keeps local error O(dt^4) (ratio 16 per halving) and
    RK1 O(dt^2) (ratio 4) against an RK4 integration of the exact
    characteristic (stage times t + dt, t, t + dt/2 on the backward path)."""
    dt = 0.02
    e1a, e1b = vortex_step_error(orc, 1, dt), vortex_step_error(orc, 1, dt / 2)
    e3a, e3b = vortex_step_error(orc, 3, dt), vortex_step_error(orc, 3, dt / 2)
    assert 3.3 < e1a / e1b < 4.7, (e1a, e1b)
    assert 12.0 < e3a / e3b < 20.0, (e3a, e3b)
    assert e3a < 1e-2 * e1a


# ---------------------------------------------------------------- box distance
BOX = (0.2, 0.3, 0.4, 0.6, 0.9, 1.0)  # min (x, y, z), max (x, y, z)


@pytest.mark.parametrize("p,want", [
    ((0.4, 0.6, 0.7), -0.2),                                  # centre: nearest faces are x (0.2 away)
    ((0.25, 0.6, 0.7), -0.05),                                # inside, nearest face x = 0.2
    ((0.4, 0.85, 0.7), -0.05),                                # inside, nearest face y = 0.9
    ((0.4, 0.6, 0.41), -0.01),                                # inside, nearest face z = 0.4
    ((0.6, 0.6, 0.7), 0.0),                                   # on a face
    ((0.9, 0.6, 0.7), 0.3),                                   # face region
    ((0.4, 0.1, 0.7), 0.2),                                   # face region, -y side
    ((0.9, 1.3, 0.7), math.hypot(0.3, 0.4)),                  # edge region
    ((0.0, 0.0, 1.5), math.sqrt(0.2 ** 2 + 0.3 ** 2 + 0.5 ** 2)),  # corner region
])
def test_box_sdf_closed_forms(orc, p, want):
    assert abs(orc.primitive_sdf(1, BOX, *p) - want) < 1e-15


def test_sphere_and_noisy_sphere_sdf(orc):
    c, r = (0.5, 0.25, 0.75), 0.3
    for p in [(0.5, 0.25, 0.75), (0.9, 0.25, 0.75), (0.1, 0.7, 0.2)]:
        want = math.dist(p, c) - r
        assert abs(orc.primitive_sdf(0, c + (r,), *p) - want) < 1e-15
        noisy = orc.primitive_sdf(2, c + (r, 0.02, 8.0, 7), *p)
        assert abs(noisy - (want + orc.value_noise4(*p, 0.02, 8.0, 7))) < 1e-15


# ---------------------------------------------------------------- value noise
AMP, F0, SEED = 0.02, 8.0, 1000
WSUM = 1.0 + 0.5 + 0.25 + 0.125


def test_value_noise_bounded_by_amplitude(orc):
    rng = np.random.default_rng(0)
    vals = np.array([orc.value_noise4(*p, AMP, F0, SEED) for p in rng.uniform(-1.0, 2.0, size=(3000, 3))])
    assert np.abs(vals).max() <= AMP
    assert np.abs(vals).max() > 0.4 * AMP  # not degenerate
    lat = np.array([orc.value_noise_lattice(int(i), int(j), int(k), SEED)
                    for i, j, k in rng.integers(-1000, 1000, size=(4000, 3))])
    assert lat.min() >= -1.0 and lat.max() < 1.0
    assert abs(lat.mean()) < 0.05 and abs(lat.std() - 1 / math.sqrt(3)) < 0.03  # uniform on [-1, 1)


def octave_nodes(orc, i, j, k):
    """Noise at a node of every octave: each octave contributes its lattice value."""
    s = sum(w * orc.value_noise_lattice(i << o, j << o, k << o, SEED + o)
            for o, w in enumerate([1.0, 0.5, 0.25, 0.125]))
    return AMP * s / WSUM


@pytest.mark.parametrize("ijk", [(0, 0, 0), (3, -2, 5), (-7, 4, 1)])
def test_value_noise_at_lattice_nodes(orc, ijk):
    i, j, k = ijk
    assert abs(orc.value_noise4(i / F0, j / F0, k / F0, AMP, F0, SEED) - octave_nodes(orc, i, j, k)) < 1e-17


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_value_noise_fade_midpoint_and_flat_nodes(orc, axis):
    """At the midpoint of an octave-0 cell edge the cubic fade is 1/2 and the
    finer octaves sit on their nodes; at a node of every octave the gradient
    is zero (fade'(0) = fade'(1) = 0: a linear fade fails both)."""
    i, j, k = 2, -3, 4
    base = [i, j, k]
    nxt = list(base)
    nxt[axis] += 1
    w = [c / F0 for c in base]
    w[axis] += 0.5 / F0
    o0 = 0.5 * (orc.value_noise_lattice(*base, SEED) + orc.value_noise_lattice(*nxt, SEED))
    finer = sum(wt * orc.value_noise_lattice(*[(2 * c + (1 if a == axis else 0)) << (o - 1)
                                                for a, c in enumerate(base)], SEED + o)
                for o, wt in [(1, 0.5), (2, 0.25), (3, 0.125)])
    assert abs(orc.value_noise4(*w, AMP, F0, SEED) - AMP * (o0 + finer) / WSUM) < 1e-17
    h = 1e-7 / F0
    p = [c / F0 for c in base]
    pp, pm = list(p), list(p)
    pp[axis] += h
    pm[axis] -= h
    grad = (orc.value_noise4(*pp, AMP, F0, SEED) - orc.value_noise4(*pm, AMP, F0, SEED)) / (2 * h)
    # a linear fade would give |grad| ~ AMP * F0 ~ 0.1
    assert abs(grad) < 1e-4 * AMP * F0
