"""ctypes wrapper over oracle/liboracle.so — TEST INFRASTRUCTURE ONLY.

Imported by tests/, by the smoke() check and by bench.py's cpu_baseline leg.
The product library never imports it.  The oracle is a C++ restatement of
the reference's CPU path (see volrus_oracle.hpp for the file:line map).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# VR_ORACLE_LIB: another build of the same oracle (the sanitizer build).
LIB_PATH = os.environ.get("VR_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")

_P = ctypes.c_void_p
_D = ctypes.c_double
_F = ctypes.c_float
_I = ctypes.c_int
_S = ctypes.c_size_t
_U64 = ctypes.c_uint64

_SIGS = {
    "or_new_grid": (_P, [_F, _D]),
    "or_new_level_set": (_P, [_D, _D]),
    "or_free": (None, [_P]),
    "or_free_f64": (None, [_P]),
    "or_set": (None, [_P, _I, _I, _I, _F]),
    "or_get": (_F, [_P, _I, _I, _I]),
    "or_is_active": (_I, [_P, _I, _I, _I]),
    "or_deactivate": (None, [_P, _I, _I, _I]),
    "or_set_class": (None, [_P, _I]),
    "or_set_transform_uniform": (None, [_P, _D]),
    "or_active_count": (_U64, [_P]),
    "or_leaf_count": (_U64, [_P]),
    "or_background": (_F, [_P]),
    "or_sample_trilinear": (_D, [_P, _D, _D, _D]),
    "or_to_nano": (_S, [_P, _P, _S]),
    "or_nano_validate": (_I, [_P, _S, ctypes.c_char_p, _S]),
    "or_from_nano": (_P, [_P, _S]),
    "or_nano_get": (_F, [_P, _I, _I, _I]),
    "or_nano_is_active": (_I, [_P, _I, _I, _I]),
    "or_sphere": (_P, [_D] * 6),
    "or_csg": (_P, [_P, _P, _I]),
    "or_scene": (_P, [_I, _P, _P, _P, _D, _D]),
    "or_primitive_sdf": (_D, [_I, _P, _D, _D, _D]),
    "or_sincospi": (None, [_D, _P, _P]),
    "or_splitmix_next": (_U64, [_P]),
    "or_value_noise4": (_D, [_D, _D, _D, _D, _D, _U64]),
    "or_value_noise_lattice": (_D, [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _U64]),
    "or_advect_grids3": (_P, [_P, _P, _P, _P, _D]),
    "or_advect_grids3_scheme": (_P, [_P, _P, _P, _P, _D, _I, _P]),
    "or_advect_analytic": (_P, [_P, _I, _P, _D, _I, _P]),
    "or_advect_dense": (_P, [_P, _I, _D, _P, _D, _I, _P]),
    "or_stage_factors": (None, [_D, _D, _D, _I, _P]),
    "or_to_f64": (_P, [_P]),
    "or_advect_analytic_f64": (_P, [_P, _I, _P, _D, _I, _P]),
    "or_advect_dense_f64": (_P, [_P, _I, _D, _P, _D, _I, _P]),
    "or_get_f64": (_D, [_P, _I, _I, _I]),
    "or_zero_crossings": (_S, [_P, _P, _S]),
    "or_trim": (None, [_P, _F]),
    "or_rebuild": (None, [_P, _F]),
    "or_rebuild_bounded": (None, [_P, _F]),
    "or_mesh": (_P, [_P, _F]),
    "or_mesh_nv": (_S, [_P]),
    "or_mesh_nt": (_S, [_P]),
    "or_mesh_copy": (None, [_P, _P, _P]),
    "or_mesh_watertight": (_I, [_P]),
    "or_mesh_free": (None, [_P]),
    "or_mc_tiles": (_S, [_P, _P, _S]),
    "or_rollout_loss_f64": (_D, [_P, _I, _D, _P, _D, _I, _I, _P]),
    "or_rollout_adjoint_f64": (_D, [_P, _I, _D, _P, _D, _I, _I, _P, _P]),
    "or_rollout_dense": (_P, [_P, _I, _D, _P, _D, _I, _I]),
    "or_time_advect_mc": (_D, [_P, _I, _P, _D, _F, _P]),
    "or_vol_write": (_S, [_P, _P, _S]),
    "or_vol_read": (_P, [_P, _S, ctypes.c_char_p, _S]),
    "or_extend_velocity": (None, [_P, _P, _P, _P, _F]),
    "or_gradient_at": (None, [_P, _I, _I, _I, _P]),
    "or_laplacian_at": (_F, [_P, _I, _I, _I]),
    "or_mean_curvature_at": (_F, [_P, _I, _I, _I]),
    "or_adaptive_surface": (_P, [_P, _F, _D, _D, ctypes.c_uint32]),
    "or_render_nv": (_S, [_P]),
    "or_render_nt": (_S, [_P]),
    "or_render_copy": (None, [_P, _P, _P, _P, _P]),
    "or_render_free": (None, [_P]),
    "or_time_step": (_D, [_P, _I, _P, _D, _I, _P, _I, _F, _I, _F]),
}

_lib = None


def build():
    """Compiles liboracle.so with the committed Makefile (g++, no GPU)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _dvec(v, n):
    a = np.zeros(n, dtype=np.float64)
    a[: len(v)] = v
    return a


class OGrid:
    """Owning handle of an oracle Grid<f32>."""

    def __init__(self, h):
        if not h:
            raise ValueError("oracle returned a null grid")
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_free(self.h)
            self.h = None

    def to_nano(self):
        n = lib().or_to_nano(self.h, None, 0)
        out = np.empty(n, dtype=np.uint8)
        lib().or_to_nano(self.h, _ptr(out), n)
        return out.tobytes()

    def active_count(self):
        return lib().or_active_count(self.h)

    def leaf_count(self):
        return lib().or_leaf_count(self.h)

    def get(self, x, y, z):
        return lib().or_get(self.h, x, y, z)

    def is_active(self, x, y, z):
        return bool(lib().or_is_active(self.h, x, y, z))

    def set(self, x, y, z, v):
        lib().or_set(self.h, x, y, z, v)

    def deactivate(self, x, y, z):
        lib().or_deactivate(self.h, x, y, z)

    def background(self):
        return lib().or_background(self.h)

    def sample(self, x, y, z):
        return lib().or_sample_trilinear(self.h, x, y, z)


class OGrid64:
    def __init__(self, h):
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_free_f64(self.h)
            self.h = None

    def get(self, x, y, z):
        return lib().or_get_f64(self.h, x, y, z)


def new_grid(bg, vs):
    return OGrid(lib().or_new_grid(bg, vs))


def new_level_set(vs, hw):
    return OGrid(lib().or_new_level_set(vs, hw))


def sphere(r, center, vs, hw):
    return OGrid(lib().or_sphere(r, center[0], center[1], center[2], vs, hw))


def csg(a, b, op):
    return OGrid(lib().or_csg(a.h, b.h, op))


def scene(prims, ops, vs, hw):
    """prims: list of (kind, params<=8); ops[i] for i>=1: 0 union, 1 intersection, 2 difference."""
    n = len(prims)
    kinds = np.array([p[0] for p in prims], dtype=np.int32)
    params = np.zeros((n, 8), dtype=np.float64)
    for i, p in enumerate(prims):
        params[i, : len(p[1])] = p[1]
    o = np.array(ops, dtype=np.int32)
    return OGrid(lib().or_scene(n, _ptr(kinds), _ptr(params), _ptr(o), vs, hw))


def from_nano(b):
    a = np.frombuffer(b, dtype=np.uint8)
    h = lib().or_from_nano(_ptr(a), a.size)
    return OGrid(h) if h else None


def nano_validate(b):
    a = np.frombuffer(b, dtype=np.uint8) if len(b) else np.zeros(1, dtype=np.uint8)
    msg = ctypes.create_string_buffer(256)
    rc = lib().or_nano_validate(_ptr(a), len(b), msg, 256)
    return rc == 0, msg.value.decode()


def nano_get(b, x, y, z):
    a = np.frombuffer(b, dtype=np.uint8)
    return lib().or_nano_get(_ptr(a), x, y, z)


def nano_is_active(b, x, y, z):
    a = np.frombuffer(b, dtype=np.uint8)
    return bool(lib().or_nano_is_active(_ptr(a), x, y, z))


def sincospi(x):
    s, c = ctypes.c_double(), ctypes.c_double()
    lib().or_sincospi(x, ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def splitmix64(seed, n):
    """The first n outputs of SplitMix64 from state `seed`."""
    s = ctypes.c_uint64(seed)
    return [lib().or_splitmix_next(ctypes.byref(s)) for _ in range(n)]


def primitive_sdf(kind, params, x, y, z):
    return lib().or_primitive_sdf(kind, _ptr(_dvec(params, 8)), x, y, z)


def value_noise4(x, y, z, amp, f0, seed):
    return lib().or_value_noise4(x, y, z, amp, f0, seed)


def value_noise_lattice(i, j, k, seed):
    return lib().or_value_noise_lattice(i, j, k, seed)


def stage_factors(period, t, dt, scheme):
    tf = np.zeros(3, dtype=np.float64)
    lib().or_stage_factors(period, t, dt, scheme, _ptr(tf))
    return tf


def advect_grids3(g, vx, vy, vz, dt):
    return OGrid(lib().or_advect_grids3(g.h, vx.h, vy.h, vz.h, dt))


def advect_grids3_scheme(g, vx, vy, vz, dt, scheme, tf=(1.0, 1.0, 1.0)):
    t = _dvec(tf, 3)
    return OGrid(lib().or_advect_grids3_scheme(g.h, vx.h, vy.h, vz.h, dt, scheme, _ptr(t)))


def advect_analytic(g, kind, params, dt, scheme=1, tf=(1.0, 1.0, 1.0)):
    p, t = _dvec(params, 6), _dvec(tf, 3)
    return OGrid(lib().or_advect_analytic(g.h, kind, _ptr(p), dt, scheme, _ptr(t)))


def advect_dense(g, uvw, n, h, dt, scheme=1, tf=(1.0, 1.0, 1.0)):
    u = np.ascontiguousarray(uvw, dtype=np.float32).reshape(-1)
    t = _dvec(tf, 3)
    return OGrid(lib().or_advect_dense(g.h, n, h, _ptr(u), dt, scheme, _ptr(t)))


def to_f64(g):
    return OGrid64(lib().or_to_f64(g.h))


def advect_analytic_f64(g64, kind, params, dt, scheme=1, tf=(1.0, 1.0, 1.0)):
    p, t = _dvec(params, 6), _dvec(tf, 3)
    return OGrid64(lib().or_advect_analytic_f64(g64.h, kind, _ptr(p), dt, scheme, _ptr(t)))


def advect_dense_f64(g64, uvw, n, h, dt, scheme=1, tf=(1.0, 1.0, 1.0)):
    u = np.ascontiguousarray(uvw, dtype=np.float32).reshape(-1)
    t = _dvec(tf, 3)
    return OGrid64(lib().or_advect_dense_f64(g64.h, n, h, _ptr(u), dt, scheme, _ptr(t)))


def zero_crossings(g):
    n = lib().or_zero_crossings(g.h, None, 0)
    out = np.empty((max(n, 1), 3), dtype=np.int32)
    lib().or_zero_crossings(g.h, _ptr(out), n)
    return out[:n]


def trim(g, hw):
    lib().or_trim(g.h, hw)


def rebuild(g, hw):
    lib().or_rebuild(g.h, hw)


def rebuild_bounded(g, hw):
    lib().or_rebuild_bounded(g.h, hw)


def mesh(g, iso=0.0):
    m = lib().or_mesh(g.h, iso)
    try:
        nv, nt = lib().or_mesh_nv(m), lib().or_mesh_nt(m)
        v = np.empty((max(nv, 1), 3), dtype=np.float64)
        t = np.empty((max(nt, 1), 3), dtype=np.uint32)
        lib().or_mesh_copy(m, _ptr(v), _ptr(t))
        wt = bool(lib().or_mesh_watertight(m))
        return v[:nv], t[:nt], wt
    finally:
        lib().or_mesh_free(m)


def mc_tiles(g):
    n = lib().or_mc_tiles(g.h, None, 0)
    out = np.empty((max(n, 1), 3), dtype=np.int32)
    lib().or_mc_tiles(g.h, _ptr(out), n)
    return out[:n]


def rollout_loss_f64(g, u64, n, h, dt, scheme, steps, target):
    u = np.ascontiguousarray(u64, dtype=np.float64).reshape(-1)
    return lib().or_rollout_loss_f64(g.h, n, h, _ptr(u), dt, scheme, steps, target.h)


def rollout_adjoint_f64(g, u64, n, h, dt, scheme, steps, target):
    u = np.ascontiguousarray(u64, dtype=np.float64).reshape(-1)
    grad = np.zeros(3 * n ** 3, dtype=np.float64)
    loss = lib().or_rollout_adjoint_f64(g.h, n, h, _ptr(u), dt, scheme, steps, target.h, _ptr(grad))
    return loss, grad


def rollout_dense(g, uvw, n, h, dt, scheme, steps):
    u = np.ascontiguousarray(uvw, dtype=np.float32).reshape(-1)
    return OGrid(lib().or_rollout_dense(g.h, n, h, _ptr(u), dt, scheme, steps))


def time_step(g, kind, params, dt, scheme, tf, do_rebuild, hw, do_mc, iso=0.0):
    """Times one oracle step in place on g (advect [+ rebuild] [+ MC]); seconds."""
    p, t = _dvec(params, 6), _dvec(tf, 3)
    return lib().or_time_step(g.h, kind, _ptr(p), dt, scheme, _ptr(t), int(do_rebuild), hw, int(do_mc), iso)


def vol_write(g):
    """write_vol (src/io/vol_format.rs:53-85), leaves in ascending origin order."""
    n = lib().or_vol_write(g.h, None, 0)
    out = np.empty(n, dtype=np.uint8)
    lib().or_vol_write(g.h, _ptr(out), n)
    return out.tobytes()


def vol_read(b):
    """read_vol (src/io/vol_format.rs:88-147): (grid, "") or (None, error text)."""
    a = np.frombuffer(bytes(b), dtype=np.uint8) if len(b) else np.zeros(1, dtype=np.uint8)
    msg = ctypes.create_string_buffer(256)
    h = lib().or_vol_read(_ptr(a), len(b), msg, 256)
    return (OGrid(h) if h else None), msg.value.decode()


def extend_velocity(sdf, vx, vy, vz, half_width):
    """extend_velocity (src/tools/level_set_track.rs:127-169), in place on vx, vy, vz."""
    lib().or_extend_velocity(sdf.h, vx.h, vy.h, vz.h, half_width)


def gradient_at(g, x, y, z):
    out = np.zeros(3, dtype=np.float32)
    lib().or_gradient_at(g.h, x, y, z, _ptr(out))
    return out


def laplacian_at(g, x, y, z):
    return lib().or_laplacian_at(g.h, x, y, z)


def mean_curvature_at(g, x, y, z):
    return lib().or_mean_curvature_at(g.h, x, y, z)


def adaptive_surface(g, iso=0.0, max_edge_length=1.0, curvature_threshold=0.5, max_depth=3):
    """adaptive_surface (src/tools/adaptive_mesh.rs:100-241): (vertices, normals, triangles, curvatures)."""
    m = lib().or_adaptive_surface(g.h, iso, max_edge_length, curvature_threshold, max_depth)
    try:
        nv, nt = lib().or_render_nv(m), lib().or_render_nt(m)
        v = np.empty((max(nv, 1), 3)); n = np.empty((max(nv, 1), 3))
        t = np.empty((max(nt, 1), 3), dtype=np.uint32); c = np.empty(max(nv, 1))
        lib().or_render_copy(m, _ptr(v), _ptr(n), _ptr(t), _ptr(c))
        return v[:nv], n[:nv], t[:nt], c[:nv]
    finally:
        lib().or_render_free(m)
