#!/usr/bin/env python3
"""Benchmark of the narrow-band level-set path on one MI355X.

Workload (one bench step = one full rollout, inputs resident in HBM):
Zalesak slotted sphere (r = 0.15 at (0.5, 0.75, 0.5), slot 0.05 wide, 0.125
deep, cut by csg_difference), voxel size 1/512, band +-3 voxels, f32; single
vortex velocity u = (sin^2(pi x) sin(2 pi y), -sin^2(pi y) sin(2 pi x), 0)
* cos(pi t / 8); 100 RK3 steps at CFL 0.5 (dt = 0.5 vs, max |u| = 1), band
rebuild every 5 steps, marching cubes every 10 steps.

Metric: active-voxel updates per second (sum over steps of the active voxels
advected, divided by wall time of the rollouts), plus the advect kernel's
achieved HBM bandwidth against the 8 TB/s roofline, measured with HIP events
on the library's stream.  The CPU baseline is the C++ restatement of the
reference's CPU path (oracle/), timed single-threaded on a bounded sample.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import importlib.util
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PEAK_GBPS = 8000.0

VS = 1.0 / 512
HW = 3.0
PERIOD = 8.0
STEPS = 100
REBUILD_EVERY = 5
MESH_EVERY = 10
SPHERE = (0.5, 0.75, 0.5, 0.15)
SLOT = (0.475, 0.55, 0.3, 0.525, 0.725, 0.7)


def load_package():
    name = "volrus_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(ROOT, "sharc.volrus_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def traffic_from_profiles():
    """HBM bytes per advect launch from the committed rocprofv3 counter summary, if any."""
    p = os.path.join(ROOT, "profiles", "advect_vortex_slotted512_traffic.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f).get("bytes_per_launch")
    return None


class SplitMix64:
    """Steele, Lea, Flood (OOPSLA 2014) generator, for the seeded scene."""

    MASK = (1 << 64) - 1

    def __init__(self, seed):
        self.s = seed & self.MASK

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & self.MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)


def batched_sphere_scenes(vr, count=64, seed=0):
    """64 level sets, each the csg_union of 1..32 spheres with centres uniform
    in [0,1]^3 and radii log-uniform in [0.01, 0.2], and a uniform velocity
    with components uniform in [-1, 1] (SplitMix64, seed 0)."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(count):
        ns = 1 + min(31, int(rng.uniform() * 32))
        prims = []
        for _ in range(ns):
            c = (rng.uniform(), rng.uniform(), rng.uniform())
            r = math.exp(math.log(0.01) + rng.uniform() * (math.log(0.2) - math.log(0.01)))
            prims.append((vr.PRIM_SPHERE, vr.CSG_UNION, (c[0], c[1], c[2], r)))
        vel = tuple(2.0 * rng.uniform() - 1.0 for _ in range(3))
        out.append((prims, vel))
    return out


def batched_extra(vr, ctx, reps=3):
    """64 sphere unions at 1/1024, one RK3 step + MC each: batched persistent graph step vs the
    per-grid loop, one GPU."""
    vs = 1.0 / 1024
    t0 = time.perf_counter()
    spec = batched_sphere_scenes(vr)
    grids = [ctx.from_primitives(p, vs, HW) for p, _ in spec]
    vels = [ctx.velocity_analytic(vr.VEL_UNIFORM, v) for _, v in spec]
    dts = [0.5 * vs / max(abs(c) for c in v) for _, v in spec]
    counts = [g.counts() for g in grids]
    active = sum(a for _, a in counts)
    log(f"batched sphere unions: 64 grids, {sum(l for l, _ in counts)} leaves, {active} active voxels "
        f"(min {min(a for _, a in counts)}, max {max(a for _, a in counts)}), built in {time.perf_counter() - t0:.1f} s")
    b = vr.Batch(ctx, grids, vels)
    b.set_dt(dts)
    lpg, wpc = b.autotune(dts[0], vr.RK3)
    res = {"grids": len(grids), "active_voxels": active, "leaves_per_grab": lpg, "waves_per_cu": wpc}
    for name, mode in [("loop", vr.BATCH_LOOP), ("graph", vr.BATCH_GRAPH)]:
        b.step(0.0, 0.0, vr.RK3, 0.0, mode)  # warm (sizes, captures)
        ctx.synchronize()
        t1 = time.perf_counter()
        for _ in range(reps):
            b.step(0.0, 0.0, vr.RK3, 0.0, mode)
        ctx.synchronize()
        dtw = (time.perf_counter() - t1) / reps
        res[f"{name}_ms"] = round(1e3 * dtw, 3)
        res[f"{name}_updates_per_s"] = round(active / dtw, 1)
    res["speedup_graph_vs_loop"] = round(res["loop_ms"] / res["graph_ms"], 3)
    counts = [b.mesh_counts(i) for i in range(len(grids))]
    res["vertices"] = sum(v for v, _ in counts)
    res["triangles"] = sum(t for _, t in counts)
    b.close()
    return res


def adjoint_extra(vr, ctx):
    """Differentiable rollout: 20 RK1 steps of a 1/256 sphere through a dense 64^3 field
    (seeded Gaussian, sigma 0.5), loss vs the sphere moved by (0.1, 0, 0),
    gradient by the adjoint with a checkpoint at every step."""
    import numpy as np

    vs, n, steps = 1.0 / 256, 64, 20
    rng = np.random.default_rng(0)
    u = rng.normal(0.0, 0.5, size=(3, n, n, n)).astype(np.float32)
    phi = ctx.from_primitives([(vr.PRIM_SPHERE, 0, (0.5, 0.5, 0.5, 0.3))], vs, HW)
    target = ctx.from_primitives([(vr.PRIM_SPHERE, 0, (0.6, 0.5, 0.5, 0.3))], vs, HW)
    vel = ctx.velocity_dense(u, n, 1.0 / n)
    dt = 0.5 * vs / float(np.abs(u).max())
    phi.adjoint_rollout(vel, n, dt, vr.RK1, steps, 1, target)  # warm
    st, _ = phi.adjoint_rollout(vel, n, dt, vr.RK1, steps, 1, target)
    return {"steps": steps, "active_voxels": st.active_voxels, "forward_ms": round(st.forward_ms, 3),
            "adjoint_ms": round(st.adjoint_ms, 3), "adjoint_over_forward": round(st.adjoint_ms / st.forward_ms, 3),
            "loss": st.loss}


def adjoint_noisy_extra(vr, ctx, count=8):
    """8 level sets at 1/2048, each a sphere r = 0.35 displaced by seeded 4-octave value noise (amplitude
    0.02, base frequency 8), band 3; per grid a dense 128^3 field of seeded Gaussian noise (sigma 0.5);
    10 RK3 steps forward + adjoint with a checkpoint every 2 steps; loss against the undisplaced sphere.
    Grids run one after another through vr_adjoint_rollout."""
    import numpy as np

    vs, n, steps, ckpt = 1.0 / 2048, 128, 10, 2
    target = ctx.from_primitives([(vr.PRIM_SPHERE, 0, (0.5, 0.5, 0.5, 0.35))], vs, HW)
    fwd = adj = 0.0
    active = leaves = 0
    for i in range(count):
        phi = ctx.from_primitives([(vr.PRIM_NOISY_SPHERE, 0, (0.5, 0.5, 0.5, 0.35, 0.02, 8.0, 1000 + i))], vs, HW)
        lc, ac = phi.counts()
        u = np.random.default_rng(100 + i).normal(0.0, 0.5, size=(3, n, n, n)).astype(np.float32)
        vel = ctx.velocity_dense(u, n, 1.0 / n)
        dt = 0.5 * vs / float(np.abs(u).max())
        if i == 0:
            phi.adjoint_rollout(vel, n, dt, vr.RK3, steps, ckpt, target, False)  # warm
        st, _ = phi.adjoint_rollout(vel, n, dt, vr.RK3, steps, ckpt, target)
        fwd += st.forward_ms
        adj += st.adjoint_ms
        active += ac
        leaves += lc
        log(f"noisy sphere {i}: {lc} leaves, {ac} active voxels, forward {st.forward_ms:.2f} ms, "
            f"adjoint {st.adjoint_ms:.2f} ms")
        del phi, vel
    return {"grids": count, "leaves": leaves, "active_voxels": active, "steps": steps, "checkpoint_every": ckpt,
            "forward_ms": round(fwd, 3), "adjoint_ms": round(adj, 3), "adjoint_over_forward": round(adj / fwd, 3)}


def cpu_baseline(seconds_budget):
    """Oracle (literal C++ restatement of the reference CPU path) on a bounded
    sample of the same workload: the first steps of the rollout, single thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    g = orc.scene([(0, SPHERE), (1, SLOT)], [0, 2], VS, HW)
    dt = 0.5 * VS
    total_s, updates, steps = 0.0, 0, 0
    while total_s < seconds_budget and steps < STEPS:
        a = g.active_count()
        tf = orc.stage_factors(PERIOD, steps * dt, dt, 3)
        do_rb = (steps + 1) % REBUILD_EVERY == 0
        do_mc = (steps + 1) % MESH_EVERY == 0
        total_s += orc.time_step(g, 4, (), dt, 3, tf, do_rb, HW, do_mc, 0.0)
        updates += a
        steps += 1
    return {
        "value": updates / total_s,
        "unit": "active-voxel updates/s",
        "cores": 1,
        "kind": "port",
        "sample": f"first {steps} of {STEPS} rollout steps (RK3 advect, rebuild every {REBUILD_EVERY} in the "
                  f"bounded-radius form proven equal to the brute force, MC every {MESH_EVERY}) on the same "
                  f"slotted sphere, {total_s:.1f} s of single-thread CPU time",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the batched and adjoint lines")
    ap.add_argument("--only", choices=["batched", "adjoint", "adjoint_noisy"],
                    help="run one side workload alone (for profiling)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    vr = load_package()
    ctx = vr.Context(0)
    if args.only:
        res = {"batched": batched_extra, "adjoint": adjoint_extra, "adjoint_noisy": adjoint_noisy_extra}[args.only](vr, ctx)
        print(json.dumps({args.only: res}), flush=True)
        ctx.close()
        return
    prims = [(vr.PRIM_SPHERE, 0, SPHERE), (vr.PRIM_BOX, vr.CSG_DIFFERENCE, SLOT)]
    base = ctx.from_primitives(prims, VS, HW)
    leaves0, active0 = base.counts()
    log(f"slotted sphere at 1/512: {leaves0} leaves, {active0} active voxels")
    vel = ctx.velocity_analytic(vr.VEL_VORTEX, (), PERIOD)
    dt = 0.5 * VS
    mesh = vr.Mesh(ctx)
    grids = [base.clone() for _ in range(args.warmup + args.steps + 1)]
    ctx.synchronize()

    def rollout(g):
        return g.rollout(vel, 0.0, dt, vr.RK3, STEPS, REBUILD_EVERY, HW, MESH_EVERY, 0.0, mesh)

    for i in range(args.warmup):
        rollout(grids[i])
    ctx.synchronize()
    log(f"warmup done ({args.warmup})")
    updates = 0
    stats = None
    t0 = time.perf_counter()
    for i in range(args.steps):
        stats = rollout(grids[args.warmup + i])
        updates += stats.voxel_updates
    ctx.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    log(f"timed {args.steps} rollouts in {wall:.4f} s; last: {stats.as_dict()}")

    # Per-kernel device time with HIP events on the library stream (one more rollout).
    ctx.set_timing(True)
    rollout(grids[-1])
    kern = {}
    for name, cls in [("advect", vr.TIMING_ADVECT), ("crossings", vr.TIMING_CROSSINGS),
                      ("rebuild", vr.TIMING_REBUILD), ("marching_cubes", vr.TIMING_MC)]:
        ms, n, b = ctx.timing(cls)
        if n:
            kern[name] = {"launches": n, "avg_us": 1e3 * ms / n, "algorithmic_bytes_per_launch": b / n,
                          "achieved_GBps": (b / n) / (ms / n * 1e-3) / 1e9}
    ctx.set_timing(False)
    adv = kern["advect"]
    traffic = traffic_from_profiles()
    roofline = {"bound": "hbm", "achieved": round(adv["achieved_GBps"], 1), "peak": PEAK_GBPS, "unit": "GB/s",
                "frac": round(adv["achieved_GBps"] / PEAK_GBPS, 4), "traffic": traffic}
    log(f"kernels: {json.dumps(kern)}")

    extras = {}
    if not args.no_extras:
        try:
            extras["adjoint_sphere256_dense64"] = adjoint_extra(vr, ctx)
            log(f"adjoint: {extras['adjoint_sphere256_dense64']}")
        except Exception as e:  # reported, not fatal for the headline line
            extras["adjoint_sphere256_dense64"] = {"error": str(e)}
        try:
            extras["adjoint_8_noisy_spheres_2048"] = adjoint_noisy_extra(vr, ctx)
            log(f"adjoint noisy: {extras['adjoint_8_noisy_spheres_2048']}")
        except Exception as e:
            extras["adjoint_8_noisy_spheres_2048"] = {"error": str(e)}
        try:
            extras["batched_64_sphere_unions_1024"] = batched_extra(vr, ctx)
            log(f"batched: {extras['batched_64_sphere_unions_1024']}")
        except Exception as e:
            extras["batched_64_sphere_unions_1024"] = {"error": str(e)}

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle) ...")
        cpu = cpu_baseline(args.cpu_seconds)
        cpu["value"] = round(cpu["value"], 1)

    value = updates / wall
    out = {
        "metric": "active-voxel updates/s (advect+MC)",
        "value": round(value, 1),
        "unit": "active-voxel updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (slotted sphere generated on the device)",
        "config": {
            "workload": "Zalesak slotted sphere r=0.15, vs=1/512, band 3, single vortex T=8, "
                        "100 RK3 steps at CFL 0.5, rebuild every 5, marching cubes every 10; one step = one rollout",
            "leaves": leaves0, "active_voxels": active0, "voxel_updates_per_rollout": stats.voxel_updates,
            "final_triangles": stats.last_triangle_count, "parallelism": "single GPU",
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kern,
        "extras": extras,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
