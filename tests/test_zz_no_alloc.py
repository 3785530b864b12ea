"""vr_marching_cubes_noalloc (the allocation-free marching cubes) against
vr_marching_cubes through the C ABI: bit-exact vertices and triangles on two
calls in a row (the second with a smaller grid, on the same device buffers and
the same scratch array), capacity clamping, the empty grid and the leaf limit."""
import ctypes

import numpy as np
import pytest

from util import sphere128

pytestmark = pytest.mark.gpu

HIP_H2D, HIP_D2H = 1, 2


def announce(msg):
    print(f"[gpu] {msg}", flush=True)


class DeviceBuffers:
    """Device memory for the outputs, from the HIP runtime the library uses."""

    def __init__(self, nbytes_list):
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.ptrs = []
        for nb in nbytes_list:
            p = ctypes.c_void_p()
            assert self.hip.hipMalloc(ctypes.byref(p), nb) == 0
            self.ptrs.append(p)

    def fill(self, i, arr):
        assert self.hip.hipMemcpy(self.ptrs[i], arr.ctypes.data, arr.nbytes, HIP_H2D) == 0

    def read(self, i, dtype, count):
        out = np.empty(count, dtype=dtype)
        if count:
            assert self.hip.hipMemcpy(out.ctypes.data, self.ptrs[i], out.nbytes, HIP_D2H) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def noalloc_mesh(ctx, vr, g, iso, bufs, vcap, tcap):
    rc = ctx.lib.vr_marching_cubes_noalloc(ctx.h, g.h, iso, bufs.ptrs[0], vcap, bufs.ptrs[1], tcap, bufs.ptrs[2])
    ctx.check(rc)
    ctx.synchronize()
    nv, nt = (int(c) for c in bufs.read(2, np.int32, 2))
    v = bufs.read(0, np.float64, 3 * min(nv, vcap)).reshape(-1, 3)
    t = bufs.read(1, np.uint32, 3 * min(nt, tcap)).reshape(-1, 3)
    return nv, nt, v, t


def test_noalloc_marching_cubes_matches(orc, ctx, vr):
    announce("noalloc marching cubes")
    big = ctx.upload(sphere128(orc).to_nano())                                  # 602 leaves
    small = ctx.upload(orc.sphere(0.2, (0.4, 0.5, 0.55), 1.0 / 64, 3.0).to_nano())
    assert small.counts()[0] < big.counts()[0]
    vcap, tcap = 200000, 400000
    bufs = DeviceBuffers([24 * vcap, 12 * tcap, 8])
    try:
        for g, iso in [(big, 0.0), (small, 0.0), (big, 0.5 / 128), (small, -0.25 / 64)]:
            vw, tw = g.mesh(iso)
            assert len(tw) > 0
            bufs.fill(2, np.array([-7, -7], dtype=np.int32))  # stale counts must be overwritten
            nv, nt, v, t = noalloc_mesh(ctx, vr, g, iso, bufs, vcap, tcap)
            assert (nv, nt) == (len(vw), len(tw))
            assert np.array_equal(t, tw), "triangles"
            assert np.array_equal(v.view(np.uint64), vw.view(np.uint64)), "vertices"
        # capacity: the counts are the full ones, writes stop at the capacity
        vw, tw = big.mesh(0.0)
        nv, nt, v, t = noalloc_mesh(ctx, vr, big, 0.0, bufs, 1000, 1500)
        assert (nv, nt) == (len(vw), len(tw))
        assert np.array_equal(v.view(np.uint64), vw[:1000].view(np.uint64)) and np.array_equal(t, tw[:1500])
        # a grid without leaves: zero counts
        empty = ctx.upload(orc.new_level_set(1.0 / 64, 3.0).to_nano())
        bufs.fill(2, np.array([5, 5], dtype=np.int32))
        assert noalloc_mesh(ctx, vr, empty, 0.0, bufs, vcap, tcap)[:2] == (0, 0)
        # above the leaf limit: a clear error, nothing launched
        huge = ctx.upload(orc.sphere(0.3, (0.5, 0.5, 0.5), 1.0 / 512, 3.0).to_nano())
        assert huge.counts()[0] > 4096
        rc = ctx.lib.vr_marching_cubes_noalloc(ctx.h, huge.h, 0.0, bufs.ptrs[0], vcap, bufs.ptrs[1], tcap,
                                               bufs.ptrs[2])
        assert rc == vr.VR_ERR_CAPACITY
        assert b"at most 4096" in ctx.lib.vr_last_error(ctx.h)
    finally:
        ctx.synchronize()
        bufs.free()
