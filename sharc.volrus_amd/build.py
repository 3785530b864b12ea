"""Builds libvolrus_amd.so (gfx950 HIP kernels + C ABI) in-tree with hipcc.

Run as `python sharc.volrus_amd/build.py`.  The library is written next to
this file; nothing is installed.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
SOURCES = ["vr_kernels.hip", "vr_api.hip", "vr_fixtures.hip", "vr_rollout.hip", "vr_adjoint.hip", "vr_mc.hip", "vr_batch.hip", "vr_io.hip", "vr_extend.hip", "vr_adaptive.hip"]
LIB = os.path.join(HERE, "libvolrus_amd.so")

# -ffp-contract=off: no fused multiply-add, so f32/f64 results round like the
# Rust reference.  Packed fp32 instructions are disabled as a codegen choice
# for correctness on this target.
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def build(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objs, cmds = [], []
    for s in SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(HERE, "_build", s + ".o")
        os.makedirs(os.path.dirname(obj), exist_ok=True)
        cmds.append([hipcc] + FLAGS + ["-c", src, "-o", obj])
        objs.append(obj)

    def compile_one(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)

    # one hipcc per source, at most one per host core (at most 16)
    workers = max(1, min(len(cmds), os.cpu_count() or 1, 16))
    with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as pool:
        for f in [pool.submit(compile_one, c) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
