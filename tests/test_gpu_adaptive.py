"""adaptive_surface (src/tools/adaptive_mesh.rs:100-241) on the GPU against
the oracle: vertices, triangles and curvatures bit for bit; normals (the only
place acos enters) within 1e-12."""
import numpy as np
import pytest

from util import slotted_sphere, sphere128

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[gpu] {msg}", flush=True)


CASES = {
    "sphere5_depth2": (lambda orc: orc.sphere(5.0, (0.0, 0.0, 0.0), 1.0, 3.0), (0.0, 2.0, 0.01, 2)),
    "sphere5_depth0": (lambda orc: orc.sphere(5.0, (0.0, 0.0, 0.0), 1.0, 3.0), (0.0, 1.0, 0.5, 0)),
    "sphere10_default": (lambda orc: orc.sphere(10.0, (0.0, 0.0, 0.0), 0.5, 3.0), (0.0, 1.0, 0.5, 3)),
    "sphere128_depth2": (sphere128, (0.0, 0.02, 2.0, 2)),
    "slotted128_all": (lambda orc: slotted_sphere(orc, 1.0 / 128), (0.0, 1e-6, 0.0, 1)),
    "offset_iso": (lambda orc: orc.sphere(6.0, (-13.5, 7.25, -3.0), 0.75, 3.0), (0.3, 1.5, 0.05, 3)),
}


@pytest.mark.parametrize("case", list(CASES))
def test_adaptive_surface_matches_oracle(orc, ctx, case):
    announce(f"adaptive surface {case}")
    make, (iso, max_edge, thr, depth) = CASES[case]
    og = make(orc)
    vw, nw, tw, cw = orc.adaptive_surface(og, iso, max_edge, thr, depth)
    vg, ng, tg, cg = ctx.upload(og.to_nano()).adaptive_surface(iso, max_edge, thr, depth)
    assert tg.shape == tw.shape and np.array_equal(tg, tw), f"{case}: triangles"
    assert vg.shape == vw.shape and np.array_equal(vg.view(np.uint64), vw.view(np.uint64)), f"{case}: vertices"
    assert np.array_equal(cg.view(np.uint64), cw.view(np.uint64)), f"{case}: curvatures"
    err = float(np.abs(ng - nw).max()) if len(nw) else 0.0
    print(f"[gpu] {case}: {len(vg)} vertices, {len(tg)} triangles, normal max |diff| {err:.2e}", flush=True)
    assert err <= 1e-12, f"{case}: normals"
    if depth > 0 and thr < 1.0:
        assert len(tg) > len(orc.mesh(og, iso)[1])  # something was refined


def test_adaptive_surface_empty_grid(orc, ctx):
    announce("adaptive surface empty")
    v, n, t, c = ctx.upload(orc.new_grid(3.0, 1.0).to_nano()).adaptive_surface(0.0, 1.0, 0.5, 3)
    assert len(v) == len(n) == len(t) == len(c) == 0
