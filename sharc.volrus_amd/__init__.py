"""ctypes binding of libvolrus_amd.so (include/volrus.h) for Python callers.

The library runs the narrow-band level-set path on an MI355X (gfx950):
NanoGrid upload/download, semi-Lagrangian advection (RK1 as in
src/tools/advection.rs:32-56, plus RK3), zero crossings, band trim/rebuild
(src/tools/level_set_track.rs:14-114) and marching cubes
(src/tools/volume_to_mesh.rs:46-201).  There is no CPU fallback: loading
fails loudly when the library is missing, and calls fail when no GPU is
present.

The directory name contains a dot, so import it by path::

    spec = importlib.util.spec_from_file_location(
        "volrus_amd", "<repo>/sharc.volrus_amd/__init__.py")
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvolrus_amd.so")

VR_OK = 0
VR_ERR_INVALID_ARGUMENT = 1
VR_ERR_INVALID_NANO = 2
VR_ERR_OUT_OF_MEMORY = 3
VR_ERR_HIP = 4
VR_ERR_UNSUPPORTED = 5
VR_ERR_CAPACITY = 6
VR_ERR_INVALID_VOL = 7

RK1 = 1
RK3 = 3
VEL_UNIFORM = 2
VEL_RIGID = 3
VEL_VORTEX = 4

PRIM_SPHERE = 0
PRIM_BOX = 1
PRIM_NOISY_SPHERE = 2
CSG_UNION = 0
CSG_INTERSECTION = 1
CSG_DIFFERENCE = 2

PRIM_DTYPE = [("kind", "<i4"), ("op", "<i4"), ("p", "<f8", 8)]

NANO_HEADER = 320
NANO_LEAF = 2128

# Every symbol include/volrus.h declares, with its ctypes signature.
_P = ctypes.c_void_p
_SIGS = {
    "vr_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_P)]),
    "vr_ctx_destroy": (ctypes.c_int, [_P]),
    "vr_last_error": (ctypes.c_char_p, [_P]),
    "vr_ctx_synchronize": (ctypes.c_int, [_P]),
    "vr_ctx_stream": (_P, [_P]),
    "vr_grid_upload_nano": (ctypes.c_int, [_P, _P, ctypes.c_size_t, ctypes.POINTER(_P)]),
    "vr_grid_download_nano": (ctypes.c_int, [_P, _P, _P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    "vr_grid_upload_vol": (ctypes.c_int, [_P, _P, ctypes.c_size_t, ctypes.POINTER(_P)]),
    "vr_grid_download_vol": (ctypes.c_int, [_P, _P, _P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    "vr_extend_velocity": (ctypes.c_int, [_P, _P, _P, _P, _P, ctypes.c_float]),
    "vr_adaptive_surface": (ctypes.c_int, [_P, _P, ctypes.c_float, ctypes.c_double, ctypes.c_double, ctypes.c_uint32,
                                           ctypes.POINTER(_P)]),
    "vr_render_mesh_counts": (ctypes.c_int, [_P, _P, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]),
    "vr_render_mesh_copy": (ctypes.c_int, [_P, _P, _P, _P, _P, ctypes.c_size_t, _P, ctypes.c_size_t]),
    "vr_render_mesh_destroy": (ctypes.c_int, [_P, _P]),
    "vr_grid_counts": (ctypes.c_int, [_P, _P, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]),
    "vr_grid_destroy": (ctypes.c_int, [_P, _P]),
    "vr_grid_from_primitives": (ctypes.c_int, [_P, _P, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                               ctypes.POINTER(_P)]),
    "vr_velocity_grids3": (ctypes.c_int, [_P, _P, _P, _P, ctypes.POINTER(_P)]),
    "vr_velocity_dense": (ctypes.c_int, [_P, _P, ctypes.c_int, ctypes.c_double, ctypes.POINTER(_P)]),
    "vr_velocity_analytic": (ctypes.c_int, [_P, ctypes.c_int, _P, ctypes.POINTER(_P)]),
    "vr_velocity_set_period": (ctypes.c_int, [_P, ctypes.c_double]),
    "vr_velocity_destroy": (ctypes.c_int, [_P, _P]),
    "vr_advect": (ctypes.c_int, [_P, _P, _P, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.POINTER(_P)]),
    "vr_zero_crossings": (ctypes.c_int, [_P, _P, _P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    "vr_trim_band": (ctypes.c_int, [_P, _P, ctypes.c_float]),
    "vr_rebuild_band": (ctypes.c_int, [_P, _P, ctypes.c_float]),
    "vr_marching_cubes": (ctypes.c_int, [_P, _P, ctypes.c_float, ctypes.POINTER(_P)]),
    "vr_mesh_counts": (ctypes.c_int, [_P, _P, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]),
    "vr_mesh_copy": (ctypes.c_int, [_P, _P, _P, ctypes.c_size_t, _P, ctypes.c_size_t]),
    "vr_mesh_destroy": (ctypes.c_int, [_P, _P]),
    "vr_mesh_create": (ctypes.c_int, [_P, ctypes.POINTER(_P)]),
    "vr_marching_cubes_noalloc": (ctypes.c_int, [_P, _P, ctypes.c_float, _P, ctypes.c_uint64, _P, ctypes.c_uint64,
                                                 _P]),
    "vr_grid_clone": (ctypes.c_int, [_P, _P, ctypes.POINTER(_P)]),
    "vr_batch_create": (ctypes.c_int, [_P, _P, _P, ctypes.c_int, ctypes.POINTER(_P)]),
    "vr_batch_set_tuning": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int]),
    "vr_batch_set_dt": (ctypes.c_int, [_P, _P, _P]),
    "vr_batch_autotune": (ctypes.c_int, [_P, _P, ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int)]),
    "vr_batch_step": (ctypes.c_int, [_P, _P, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_float,
                                     ctypes.c_int]),
    "vr_batch_mesh_counts": (ctypes.c_int, [_P, _P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64),
                                            ctypes.POINTER(ctypes.c_uint64)]),
    "vr_batch_mesh_copy": (ctypes.c_int, [_P, _P, ctypes.c_int, ctypes.c_int, _P, ctypes.c_size_t, _P,
                                          ctypes.c_size_t]),
    "vr_batch_output": (ctypes.c_int, [_P, ctypes.c_int, ctypes.POINTER(_P)]),
    "vr_batch_destroy": (ctypes.c_int, [_P, _P]),
    "vr_adjoint_rollout": (ctypes.c_int, [_P, _P, _P, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          _P, _P, _P]),
    "vr_adjoint_rollout_batch": (ctypes.c_int, [_P, ctypes.c_int, _P, _P, _P, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_int, _P, _P, _P]),
    "vr_rollout": (ctypes.c_int, [_P, _P, _P, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_float, _P, _P]),
    "vr_ctx_set_timing": (ctypes.c_int, [_P, ctypes.c_int]),
    "vr_ctx_timing": (ctypes.c_int, [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                     ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_double)]),
}

TIMING_ADVECT = 0
TIMING_CROSSINGS = 1
TIMING_REBUILD = 2
TIMING_COMPACT = 3
TIMING_MC = 4
TIMING_MC_EMIT = 6


class AdjointStats(ctypes.Structure):
    _fields_ = [("loss", ctypes.c_double), ("forward_ms", ctypes.c_double), ("adjoint_ms", ctypes.c_double),
                ("steps", ctypes.c_uint64), ("active_voxels", ctypes.c_uint64)]


class RolloutStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "voxel_updates", "advect_steps", "rebuilds", "meshes", "last_vertex_count", "last_triangle_count",
        "final_leaf_count", "final_active_count")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}

_lib = None


def load_library(path=LIB_PATH):
    """Loads the built library and binds every C ABI symbol (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is not built; run `python sharc.volrus_amd/build.py` (there is no CPU fallback)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class VolrusError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"volrus error {code}: {msg}")
        self.code = code


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


class Context:
    """One device, one HIP stream (vr_ctx)."""

    def __init__(self, device=0):
        self.lib = load_library()
        h = ctypes.c_void_p()
        rc = self.lib.vr_ctx_create(device, ctypes.byref(h))
        if rc != VR_OK:
            raise VolrusError(rc, "vr_ctx_create failed (is a GPU present?)")
        self.h = h

    def check(self, rc):
        if rc != VR_OK:
            raise VolrusError(rc, self.lib.vr_last_error(self.h).decode())

    def synchronize(self):
        self.check(self.lib.vr_ctx_synchronize(self.h))

    def set_timing(self, on):
        self.check(self.lib.vr_ctx_set_timing(self.h, 1 if on else 0))

    def timing(self, kernel_class):
        """(total device ms, launches, algorithmic bytes) since set_timing(True)."""
        ms, n, b = ctypes.c_double(), ctypes.c_uint64(), ctypes.c_double()
        self.check(self.lib.vr_ctx_timing(self.h, kernel_class, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(b)))
        return ms.value, n.value, b.value

    def stream(self):
        return self.lib.vr_ctx_stream(self.h)

    def close(self):
        if self.h:
            self.lib.vr_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- grids
    def upload(self, nano_bytes):
        b = np.frombuffer(bytes(nano_bytes), dtype=np.uint8)
        h = ctypes.c_void_p()
        self.check(self.lib.vr_grid_upload_nano(self.h, _ptr(b), b.size, ctypes.byref(h)))
        return Grid(self, h)

    def upload_vol(self, vol_bytes):
        """A grid from the bytes of a .vol file (read_vol)."""
        b = np.frombuffer(bytes(vol_bytes), dtype=np.uint8) if len(vol_bytes) else np.zeros(1, dtype=np.uint8)
        h = ctypes.c_void_p()
        self.check(self.lib.vr_grid_upload_vol(self.h, _ptr(b), len(vol_bytes), ctypes.byref(h)))
        return Grid(self, h)

    def extend_velocity(self, sdf, vx, vy, vz, half_width):
        """extend_velocity, in place on the velocity grids vx, vy, vz."""
        self.check(self.lib.vr_extend_velocity(self.h, sdf.h, vx.h, vy.h, vz.h, half_width))

    def from_primitives(self, prims, voxel_size, half_width):
        """prims: list of (kind, op, params); op is ignored for the first."""
        a = np.zeros(len(prims), dtype=PRIM_DTYPE)
        for i, (kind, op, params) in enumerate(prims):
            a[i]["kind"], a[i]["op"] = kind, op
            a[i]["p"][: len(params)] = params
        h = ctypes.c_void_p()
        self.check(self.lib.vr_grid_from_primitives(self.h, _ptr(a), len(prims), voxel_size, half_width,
                                                    ctypes.byref(h)))
        return Grid(self, h)

    def velocity_grids3(self, vx, vy, vz):
        h = ctypes.c_void_p()
        self.check(self.lib.vr_velocity_grids3(self.h, vx.h, vy.h, vz.h, ctypes.byref(h)))
        return Velocity(self, h, keep=(vx, vy, vz))

    def velocity_dense(self, uvw, n, h_spacing):
        u = np.ascontiguousarray(uvw, dtype=np.float32).reshape(-1)
        assert u.size == 3 * n ** 3
        h = ctypes.c_void_p()
        self.check(self.lib.vr_velocity_dense(self.h, _ptr(u), n, h_spacing, ctypes.byref(h)))
        return Velocity(self, h)

    def velocity_analytic(self, kind, params=(0, 0, 0, 0, 0, 0), period=0.0):
        p = np.zeros(6, dtype=np.float64)
        p[: len(params)] = params
        h = ctypes.c_void_p()
        self.check(self.lib.vr_velocity_analytic(self.h, kind, _ptr(p), ctypes.byref(h)))
        v = Velocity(self, h)
        if period:
            self.check(self.lib.vr_velocity_set_period(h, period))
        return v


class Grid:
    """A level set resident in HBM (vr_grid)."""

    def __init__(self, ctx, h):
        self.ctx, self.h = ctx, h

    def download(self):
        n = ctypes.c_size_t()
        self.ctx.check(self.ctx.lib.vr_grid_download_nano(self.ctx.h, self.h, None, 0, ctypes.byref(n)))
        out = np.empty(n.value, dtype=np.uint8)
        self.ctx.check(self.ctx.lib.vr_grid_download_nano(self.ctx.h, self.h, _ptr(out), out.size, ctypes.byref(n)))
        return out.tobytes()

    def download_vol(self):
        """The grid as the bytes of a .vol file (write_vol), leaves in ascending origin order."""
        n = ctypes.c_size_t()
        self.ctx.check(self.ctx.lib.vr_grid_download_vol(self.ctx.h, self.h, None, 0, ctypes.byref(n)))
        out = np.empty(n.value, dtype=np.uint8)
        self.ctx.check(self.ctx.lib.vr_grid_download_vol(self.ctx.h, self.h, _ptr(out), out.size, ctypes.byref(n)))
        return out.tobytes()

    def clone(self):
        h = ctypes.c_void_p()
        self.ctx.check(self.ctx.lib.vr_grid_clone(self.ctx.h, self.h, ctypes.byref(h)))
        return Grid(self.ctx, h)

    def counts(self):
        lc, ac = ctypes.c_uint64(), ctypes.c_uint64()
        self.ctx.check(self.ctx.lib.vr_grid_counts(self.ctx.h, self.h, ctypes.byref(lc), ctypes.byref(ac)))
        return lc.value, ac.value

    def advect(self, vel, dt, t=0.0, scheme=RK1):
        h = ctypes.c_void_p()
        self.ctx.check(self.ctx.lib.vr_advect(self.ctx.h, self.h, vel.h, dt, t, scheme, ctypes.byref(h)))
        return Grid(self.ctx, h)

    def rollout(self, vel, t0, dt, scheme, steps, rebuild_every=0, half_width=3.0, mesh_every=0, iso=0.0,
                mesh=None):
        """In-place multi-step rollout; returns RolloutStats."""
        st = RolloutStats()
        self.ctx.check(self.ctx.lib.vr_rollout(self.ctx.h, self.h, vel.h, t0, dt, scheme, steps, rebuild_every,
                                               half_width, mesh_every, iso, mesh.h if mesh else None,
                                               ctypes.byref(st)))
        return st

    def adjoint_rollout(self, vel, n, dt, scheme, steps, checkpoint_every, target, want_grad=True):
        """Differentiable rollout through a dense field of n^3 nodes; returns (AdjointStats, grad[3, n, n, n])."""
        st = AdjointStats()
        grad = np.zeros(3 * n ** 3, dtype=np.float32) if want_grad else None
        self.ctx.check(self.ctx.lib.vr_adjoint_rollout(self.ctx.h, self.h, vel.h, dt, scheme, steps,
                                                       checkpoint_every, target.h, ctypes.byref(st), _ptr(grad)))
        return st, (grad.reshape(3, n, n, n) if want_grad else None)

    def adaptive_surface(self, iso=0.0, max_edge_length=1.0, curvature_threshold=0.5, max_depth=3):
        """adaptive_surface: (vertices, normals, triangles, curvatures)."""
        h = ctypes.c_void_p()
        self.ctx.check(self.ctx.lib.vr_adaptive_surface(self.ctx.h, self.h, iso, max_edge_length, curvature_threshold,
                                                        max_depth, ctypes.byref(h)))
        try:
            nv, nt = ctypes.c_uint64(), ctypes.c_uint64()
            self.ctx.check(self.ctx.lib.vr_render_mesh_counts(self.ctx.h, h, ctypes.byref(nv), ctypes.byref(nt)))
            v = np.empty((max(nv.value, 1), 3))
            n = np.empty((max(nv.value, 1), 3))
            c = np.empty(max(nv.value, 1))
            t = np.empty((max(nt.value, 1), 3), dtype=np.uint32)
            self.ctx.check(self.ctx.lib.vr_render_mesh_copy(self.ctx.h, h, _ptr(v), _ptr(n), _ptr(c), nv.value, _ptr(t),
                                                            nt.value))
            return v[: nv.value], n[: nv.value], t[: nt.value], c[: nv.value]
        finally:
            self.ctx.lib.vr_render_mesh_destroy(self.ctx.h, h)

    def zero_crossings(self):
        n = ctypes.c_size_t()
        self.ctx.check(self.ctx.lib.vr_zero_crossings(self.ctx.h, self.h, None, 0, ctypes.byref(n)))
        out = np.empty((max(n.value, 1), 3), dtype=np.int32)
        self.ctx.check(self.ctx.lib.vr_zero_crossings(self.ctx.h, self.h, _ptr(out), n.value, ctypes.byref(n)))
        return out[: n.value]

    def trim(self, half_width):
        self.ctx.check(self.ctx.lib.vr_trim_band(self.ctx.h, self.h, half_width))

    def rebuild(self, half_width):
        self.ctx.check(self.ctx.lib.vr_rebuild_band(self.ctx.h, self.h, half_width))

    def mesh(self, iso=0.0):
        h = ctypes.c_void_p()
        self.ctx.check(self.ctx.lib.vr_marching_cubes(self.ctx.h, self.h, iso, ctypes.byref(h)))
        try:
            nv, nt = ctypes.c_uint64(), ctypes.c_uint64()
            self.ctx.check(self.ctx.lib.vr_mesh_counts(self.ctx.h, h, ctypes.byref(nv), ctypes.byref(nt)))
            v = np.empty((max(nv.value, 1), 3), dtype=np.float64)
            t = np.empty((max(nt.value, 1), 3), dtype=np.uint32)
            self.ctx.check(self.ctx.lib.vr_mesh_copy(self.ctx.h, h, _ptr(v), nv.value, _ptr(t), nt.value))
            return v[: nv.value], t[: nt.value]
        finally:
            self.ctx.lib.vr_mesh_destroy(self.ctx.h, h)

    def close(self):
        if self.h:
            self.ctx.lib.vr_grid_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def adjoint_rollout_batch(ctx, phis, vels, ns, dts, scheme, steps, checkpoint_every, targets, want_grad=True):
    """vr_adjoint_rollout_batch: one launch per step over all level sets.
    Returns ([AdjointStats per grid], [grad[3, n, n, n] or None per grid])."""
    G = len(phis)
    ph = (ctypes.c_void_p * G)(*[g.h.value if hasattr(g.h, "value") else g.h for g in phis])
    vh = (ctypes.c_void_p * G)(*[v.h.value if hasattr(v.h, "value") else v.h for v in vels])
    th = (ctypes.c_void_p * G)(*[t.h.value if hasattr(t.h, "value") else t.h for t in targets])
    d = np.ascontiguousarray(dts, dtype=np.float64)
    stats = (AdjointStats * G)()
    grads = [np.zeros(3 * n ** 3, dtype=np.float32) for n in ns] if want_grad else None
    gp = (ctypes.c_void_p * G)(*[g.ctypes.data for g in grads]) if want_grad else None
    ctx.check(ctx.lib.vr_adjoint_rollout_batch(ctx.h, G, ph, vh, _ptr(d), scheme, steps, checkpoint_every, th,
                                                stats, gp))
    return list(stats), ([g.reshape(3, n, n, n) for g, n in zip(grads, ns)] if want_grad else None)


BATCH_LOOP = 0
BATCH_EAGER = 1
BATCH_GRAPH = 2
TIMING_GRAPH = 5


class Batch:
    """Many level sets advected and meshed in one step (vr_batch)."""

    def __init__(self, ctx, grids, vels):
        self.ctx, self._keep = ctx, (list(grids), list(vels))
        gh = (ctypes.c_void_p * len(grids))(*[g.h.value if hasattr(g.h, "value") else g.h for g in grids])
        vh = (ctypes.c_void_p * len(vels))(*[v.h.value if hasattr(v.h, "value") else v.h for v in vels])
        h = ctypes.c_void_p()
        ctx.check(ctx.lib.vr_batch_create(ctx.h, gh, vh, len(grids), ctypes.byref(h)))
        self.h = h
        self.n = len(grids)

    def set_dt(self, dts):
        """Per-grid time steps (None: use the dt passed to step)."""
        if dts is None:
            self.ctx.check(self.ctx.lib.vr_batch_set_dt(self.ctx.h, self.h, None))
            return
        a = np.ascontiguousarray(dts, dtype=np.float64)
        assert a.size == self.n
        self.ctx.check(self.ctx.lib.vr_batch_set_dt(self.ctx.h, self.h, _ptr(a)))

    def set_tuning(self, leaves_per_grab, waves_per_cu):
        self.ctx.check(self.ctx.lib.vr_batch_set_tuning(self.h, leaves_per_grab, waves_per_cu))

    def autotune(self, dt, scheme):
        a, b = ctypes.c_int(), ctypes.c_int()
        self.ctx.check(self.ctx.lib.vr_batch_autotune(self.ctx.h, self.h, dt, scheme, ctypes.byref(a),
                                                      ctypes.byref(b)))
        return a.value, b.value

    def step(self, dt, t, scheme, iso=0.0, mode=BATCH_GRAPH):
        self.ctx.check(self.ctx.lib.vr_batch_step(self.ctx.h, self.h, dt, t, scheme, iso, mode))

    def mesh_counts(self, i, mode=BATCH_GRAPH):
        """(vertex count, triangle count) of grid i's mesh from the last step."""
        nv, nt = ctypes.c_uint64(), ctypes.c_uint64()
        self.ctx.check(self.ctx.lib.vr_batch_mesh_counts(self.ctx.h, self.h, i, mode, ctypes.byref(nv),
                                                         ctypes.byref(nt)))
        return nv.value, nt.value

    def mesh(self, i, mode=BATCH_GRAPH):
        nv, nt = ctypes.c_uint64(), ctypes.c_uint64()
        self.ctx.check(self.ctx.lib.vr_batch_mesh_counts(self.ctx.h, self.h, i, mode, ctypes.byref(nv),
                                                         ctypes.byref(nt)))
        v = np.empty((max(nv.value, 1), 3), dtype=np.float64)
        t = np.empty((max(nt.value, 1), 3), dtype=np.uint32)
        self.ctx.check(self.ctx.lib.vr_batch_mesh_copy(self.ctx.h, self.h, i, mode, _ptr(v), nv.value, _ptr(t),
                                                       nt.value))
        return v[: nv.value], t[: nt.value]

    def output_nano(self, i):
        h = ctypes.c_void_p()
        self.ctx.check(self.ctx.lib.vr_batch_output(self.h, i, ctypes.byref(h)))
        g = Grid(self.ctx, h)
        try:
            return g.download()
        finally:
            g.h = None  # borrowed

    def close(self):
        if self.h:
            self.ctx.lib.vr_batch_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Mesh:
    """A reusable device mesh (vr_mesh) that vr_rollout fills."""

    def __init__(self, ctx):
        self.ctx = ctx
        h = ctypes.c_void_p()
        ctx.check(ctx.lib.vr_mesh_create(ctx.h, ctypes.byref(h)))
        self.h = h

    def arrays(self):
        nv, nt = ctypes.c_uint64(), ctypes.c_uint64()
        self.ctx.check(self.ctx.lib.vr_mesh_counts(self.ctx.h, self.h, ctypes.byref(nv), ctypes.byref(nt)))
        v = np.empty((max(nv.value, 1), 3), dtype=np.float64)
        t = np.empty((max(nt.value, 1), 3), dtype=np.uint32)
        self.ctx.check(self.ctx.lib.vr_mesh_copy(self.ctx.h, self.h, _ptr(v), nv.value, _ptr(t), nt.value))
        return v[: nv.value], t[: nt.value]

    def close(self):
        if self.h:
            self.ctx.lib.vr_mesh_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Velocity:
    def __init__(self, ctx, h, keep=()):
        self.ctx, self.h, self._keep = ctx, h, keep

    def close(self):
        if self.h:
            self.ctx.lib.vr_velocity_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
