#!/usr/bin/env python3
"""Benchmark of the narrow-band level-set path on one MI355X.

Headline workload (one bench step = one batched step, inputs resident in HBM):
64 level sets, each the csg_union of 1..32 spheres with centres uniform in
[0,1]^3 and radii log-uniform in [0.01, 0.2] (SplitMix64, seed 0), voxel size
1/1024, band +-3 voxels, f32; each grid has its own uniform velocity with
components uniform in [-1, 1] and its own dt = 0.5 vs / max|v_i|.  One step is
RK3 advection of every active voxel of all 64 grids followed by marching cubes
(iso 0) of all 64 results, replayed as one captured hipGraph (leaves per grab
and waves per CU picked by the autotuner).

Metric: active-voxel updates per second (active voxels of all grids times
steps, divided by the wall time of the timed steps).  `roofline` is the
dominant kernel's algorithmic bytes per launch over its average launch time,
measured with HIP events on the library's stream (each launch of an eager
step bracketed by its own events), against the 8 TB/s HBM peak; `roofline_step`
is the whole step's algorithmic bytes (2128 B per leaf read by advection,
2128 written, 2128 read by marching cubes, 24 B per vertex and 12 per
triangle written) over the graph step time.  The CPU baseline is the C++
restatement of the reference's CPU path (oracle/), timed single-threaded on a
bounded sample of the same grids.  Side workloads (the Zalesak rollout, the
differentiable rollouts, the per-grid loop) are reported under `extras`.

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
LEAF = 2128
HW = 3.0

# Zalesak rollout (side workload)
VS_Z = 1.0 / 512
PERIOD = 8.0
STEPS_Z = 100
REBUILD_EVERY = 5
MESH_EVERY = 10
SPHERE = (0.5, 0.75, 0.5, 0.15)
SLOT = (0.475, 0.55, 0.3, 0.525, 0.725, 0.7)

VS_B = 1.0 / 1024  # batched headline


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


def profile_traffic(name):
    """HBM bytes per launch of a kernel from a committed rocprofv3 counter summary, if any."""
    p = os.path.join(ROOT, "profiles", name)
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f).get("bytes_per_launch")
    return None


def host_cpu():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        nproc = os.cpu_count()
    return {"nproc": nproc, "model": model}


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


def kernel_times(vr, ctx):
    """Per-class device time, launches and algorithmic bytes since timing was turned on."""
    out = {}
    for name, cls in [("advect", vr.TIMING_ADVECT), ("crossings", vr.TIMING_CROSSINGS),
                      ("rebuild", vr.TIMING_REBUILD), ("compact", vr.TIMING_COMPACT),
                      ("mc_count", vr.TIMING_MC), ("mc_emit", vr.TIMING_MC_EMIT)]:
        ms, n, b = ctx.timing(cls)
        if n:
            out[name] = {"launches": n, "avg_us": round(1e3 * ms / n, 3),
                         "algorithmic_bytes_per_launch": round(b / n, 1),
                         "achieved_GBps": round((b / n) / (ms / n * 1e-3) / 1e9, 1),
                         "frac": round((b / n) / (ms / n * 1e-3) / 1e9 / PEAK_GBPS, 4)}
    return out


KERNEL_NAMES = {"advect": "k_batch_advect", "mc_count": "k_mc_count", "mc_emit": "k_mc_emit"}


class Batched:
    """The headline workload, built once and reused by the timed steps."""

    def __init__(self, vr, ctx):
        self.vr, self.ctx = vr, ctx
        t0 = time.perf_counter()
        self.spec = batched_sphere_scenes(vr)
        self.grids = [ctx.from_primitives(p, VS_B, HW) for p, _ in self.spec]
        self.vels = [ctx.velocity_analytic(vr.VEL_UNIFORM, v) for _, v in self.spec]
        self.dts = [0.5 * VS_B / max(abs(c) for c in v) for _, v in self.spec]
        counts = [g.counts() for g in self.grids]
        self.leaves = sum(l for l, _ in counts)
        self.active = sum(a for _, a in counts)
        self.per_grid_active = [a for _, a in counts]
        log(f"batched sphere unions: 64 grids, {self.leaves} leaves, {self.active} active voxels "
            f"(min {min(self.per_grid_active)}, max {max(self.per_grid_active)}), built in "
            f"{time.perf_counter() - t0:.1f} s")
        self.b = vr.Batch(ctx, self.grids, self.vels)
        self.b.set_dt(self.dts)
        self.lpg, self.wpc = self.b.autotune(self.dts[0], vr.RK3)
        log(f"autotune: leaves_per_grab {self.lpg}, waves_per_cu {self.wpc}")

    def step(self, mode):
        self.b.step(0.0, 0.0, self.vr.RK3, 0.0, mode)

    def timed(self, mode, steps, warmup):
        for _ in range(max(1, warmup)):
            self.step(mode)
        self.ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            self.step(mode)
        self.ctx.synchronize()
        return time.perf_counter() - t0

    def mesh_totals(self, mode):
        counts = [self.b.mesh_counts(i, mode) for i in range(len(self.grids))]
        return sum(v for v, _ in counts), sum(t for _, t in counts)


def cpu_baseline_batched(bt, seconds_budget):
    """Oracle (literal C++ restatement of the reference CPU path) on a bounded
    sample of the headline workload: grids 0, 1, 2, ... of the 64, one RK3
    advect + marching cubes each, single thread, until the budget is spent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    total_s, updates, used = 0.0, 0, []
    for i, (g, (_, vel)) in enumerate(zip(bt.grids, bt.spec)):
        if total_s >= seconds_budget:
            break
        og = orc.from_nano(g.download())
        a = og.active_count()
        total_s += orc.time_step(og, 2, vel, bt.dts[i], 3, (1.0, 1.0, 1.0), False, HW, True, 0.0)
        updates += a
        used.append(i)
        del og
    return {
        "value": round(updates / total_s, 1),
        "unit": "active-voxel updates/s",
        "cores": 1,
        "kind": "port",
        "sample": f"grids {used[0]}..{used[-1]} of the 64 ({updates} active voxels of {bt.active}), one RK3 "
                  f"advect + marching cubes each, {total_s:.1f} s of single-thread CPU time",
        "host": host_cpu(),
    }


def rollout_extra(vr, ctx, steps=5, warmup=2):
    """Zalesak slotted sphere (r = 0.15 at (0.5, 0.75, 0.5), slot 0.05 wide,
    0.125 deep), 1/512, band 3, single vortex with cos(pi t / 8), 100 RK3
    steps at CFL 0.5 (dt = 0.5 vs), rebuild every 5, marching cubes every 10;
    one step = one rollout."""
    prims = [(vr.PRIM_SPHERE, 0, SPHERE), (vr.PRIM_BOX, vr.CSG_DIFFERENCE, SLOT)]
    base = ctx.from_primitives(prims, VS_Z, HW)
    leaves0, active0 = base.counts()
    vel = ctx.velocity_analytic(vr.VEL_VORTEX, (), PERIOD)
    dt = 0.5 * VS_Z
    mesh = vr.Mesh(ctx)
    grids = [base.clone() for _ in range(warmup + steps + 1)]
    ctx.synchronize()

    def rollout(g):
        return g.rollout(vel, 0.0, dt, vr.RK3, STEPS_Z, REBUILD_EVERY, HW, MESH_EVERY, 0.0, mesh)

    for i in range(warmup):
        rollout(grids[i])
    ctx.synchronize()
    updates, st = 0, None
    t0 = time.perf_counter()
    for i in range(steps):
        st = rollout(grids[warmup + i])
        updates += st.voxel_updates
    ctx.synchronize()
    wall = time.perf_counter() - t0
    ctx.set_timing(True)
    rollout(grids[-1])
    kern = kernel_times(vr, ctx)
    ctx.set_timing(False)
    adv = kern["advect"]
    return {"leaves": leaves0, "active_voxels": active0, "voxel_updates_per_rollout": st.voxel_updates,
            "ms_per_rollout": round(1e3 * wall / steps, 4), "updates_per_s": round(updates / wall, 1),
            "final_leaf_count": st.final_leaf_count, "final_triangles": st.last_triangle_count,
            "advect_roofline": {"achieved": adv["achieved_GBps"], "frac": adv["frac"],
                                "traffic": profile_traffic("advect_vortex_slotted512_traffic.json")},
            "kernels": kern}


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
    All grids in one launch per step (vr_adjoint_rollout_batch)."""
    import numpy as np

    vs, n, steps, ckpt = 1.0 / 2048, 128, 10, 2
    phis, vels, dts, active, leaves = [], [], [], 0, 0
    target = ctx.from_primitives([(vr.PRIM_SPHERE, 0, (0.5, 0.5, 0.5, 0.35))], vs, HW)
    for i in range(count):
        phi = ctx.from_primitives([(vr.PRIM_NOISY_SPHERE, 0, (0.5, 0.5, 0.5, 0.35, 0.02, 8.0, 1000 + i))], vs, HW)
        lc, ac = phi.counts()
        u = np.random.default_rng(100 + i).normal(0.0, 0.5, size=(3, n, n, n)).astype(np.float32)
        phis.append(phi)
        vels.append(ctx.velocity_dense(u, n, 1.0 / n))
        dts.append(0.5 * vs / float(np.abs(u).max()))
        active += ac
        leaves += lc
    log(f"noisy spheres: {count} grids, {leaves} leaves, {active} active voxels")
    targets = [target] * count
    vr.adjoint_rollout_batch(ctx, phis, vels, [n] * count, dts, vr.RK3, steps, ckpt, targets, False)  # warm
    st, _ = vr.adjoint_rollout_batch(ctx, phis, vels, [n] * count, dts, vr.RK3, steps, ckpt, targets, False)
    fwd, adj = st[0].forward_ms, st[0].adjoint_ms
    return {"grids": count, "leaves": leaves, "active_voxels": active, "steps": steps, "checkpoint_every": ckpt,
            "launch": "one launch per step over all grids", "forward_ms": round(fwd, 3), "adjoint_ms": round(adj, 3),
            "adjoint_over_forward": round(adj / fwd, 3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the side workloads")
    ap.add_argument("--only", choices=["batched", "rollout", "adjoint", "adjoint_noisy"],
                    help="run one workload alone (for profiling)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    vr = load_package()
    ctx = vr.Context(0)
    if args.only and args.only != "batched":
        res = {"rollout": rollout_extra, "adjoint": adjoint_extra, "adjoint_noisy": adjoint_noisy_extra}[args.only](
            vr, ctx)
        print(json.dumps({args.only: res}), flush=True)
        ctx.close()
        return

    bt = Batched(vr, ctx)
    wall = bt.timed(vr.BATCH_GRAPH, args.steps, args.warmup)
    ms_step = 1e3 * wall / args.steps
    nv, ntri = bt.mesh_totals(vr.BATCH_GRAPH)
    log(f"timed {args.steps} graph steps in {wall:.4f} s ({ms_step:.3f} ms/step); mesh {nv} vertices, {ntri} triangles")
    if args.only == "batched":
        # A plain copy of known size (a clone of the largest grid: k_copy_u4 over its leaf records) for
        # scaling the FETCH_SIZE / WRITE_SIZE counters of counter runs to bytes.
        big = max(range(len(bt.grids)), key=lambda i: bt.per_grid_active[i])
        copy_bytes = LEAF * bt.grids[big].counts()[0]
        bt.grids[big].clone()
        ctx.synchronize()
        print(json.dumps({"batched_graph_ms": round(ms_step, 4), "calibration_copy_bytes": copy_bytes}), flush=True)
        ctx.close()
        return

    # Per-kernel device time: the same kernels launched eagerly, each launch
    # between its own pair of HIP events on the library stream.
    ctx.set_timing(True)
    for _ in range(3):
        bt.step(vr.BATCH_EAGER)
    kern = kernel_times(vr, ctx)
    ctx.set_timing(False)
    log(f"kernels: {json.dumps(kern)}")
    dom = max(("advect", "mc_count", "mc_emit"), key=lambda k: kern[k]["avg_us"] * kern[k]["launches"])
    dk = kern[dom]
    roofline = {"bound": "hbm", "achieved": dk["achieved_GBps"], "peak": PEAK_GBPS, "unit": "GB/s",
                "frac": dk["frac"], "traffic": profile_traffic(f"{KERNEL_NAMES[dom]}_batched_sphere_unions_64x1024_traffic.json"),
                "kernel": KERNEL_NAMES[dom]}
    step_bytes = 3.0 * LEAF * bt.leaves + 24.0 * nv + 12.0 * ntri
    roofline_step = {"algorithmic_bytes": step_bytes, "ms": round(ms_step, 4),
                     "achieved": round(step_bytes / (ms_step * 1e-3) / 1e9, 1),
                     "frac": round(step_bytes / (ms_step * 1e-3) / 1e9 / PEAK_GBPS, 4)}

    extras = {}
    if not args.no_extras:
        loop_wall = bt.timed(vr.BATCH_LOOP, 3, 1)
        extras["per_grid_loop"] = {"ms_per_step": round(1e3 * loop_wall / 3, 3),
                                   "graph_speedup": round(loop_wall / 3 / (wall / args.steps), 3)}
        log(f"loop: {extras['per_grid_loop']}")
        for name, fn in [("rollout_zalesak512", rollout_extra), ("adjoint_sphere256_dense64", adjoint_extra),
                         ("adjoint_8_noisy_spheres_2048", adjoint_noisy_extra)]:
            try:
                extras[name] = fn(vr, ctx)
            except Exception as e:  # reported, not fatal for the headline line
                extras[name] = {"error": str(e)}
            log(f"{name}: {extras[name]}")

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle) ...")
        cpu = cpu_baseline_batched(bt, args.cpu_seconds)

    value = bt.active * args.steps / wall
    out = {
        "metric": "active-voxel updates/s (advect+MC)",
        "value": round(value, 1),
        "unit": "active-voxel updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (sphere unions generated on the device from SplitMix64 seed 0)",
        "config": {
            "workload": "64 sphere unions (1-32 spheres each, radii log-uniform in [0.01, 0.2]), vs=1/1024, band 3, "
                        "uniform velocity per grid, dt = 0.5 vs / max|v|; one step = RK3 advect + marching cubes of "
                        "all 64 grids as one hipGraph replay",
            "grids": len(bt.grids), "leaves": bt.leaves, "active_voxels": bt.active, "vertices": nv,
            "triangles": ntri, "leaves_per_grab": bt.lpg, "waves_per_cu": bt.wpc, "parallelism": "single GPU",
        },
        "roofline": roofline,
        "roofline_step": roofline_step,
        "cpu_baseline": cpu,
        "kernels": kern,
        "extras": extras,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
