"""Builds the native library, photonoxide_amd/libpxfdtd.so, for the MI355X (gfx950).

hipcc cross-compiles without a GPU. Every HIP and C++ source is compiled without floating-point
contraction (-ffp-contract=off) so that each + - * / is one IEEE 754 rounding, as in the
reference, and without packed FP32 instructions. The sources compile to objects in parallel
(the step kernels of each precision and coefficient mode are units of their own), then link.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpxfdtd.so")
SOURCES = ["resident.hip", "simulation.cpp", "portable.cpp", "batch_launch.hip", "flux.hip", "harmonic_device.hip", "hbm.hip", "coefficients.hip",
           "step_f32_m0.hip", "step_f32_m1.hip", "step_f32_m2.hip",
           "step_f64_m0.hip", "step_f64_m1.hip", "step_f64_m2.hip"]
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]

    def compile_one(s):
        obj = os.path.join(OBJ, s + ".o")
        cmd = [hipcc, *FLAGS, "-c", os.path.join(CSRC, s), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        return s, obj, r

    jobs = max(1, min(len(srcs), os.cpu_count() or 1, 16))
    with ThreadPoolExecutor(jobs) as pool:
        results = list(pool.map(compile_one, srcs))
    failed = [(s, r) for s, _, r in results if r.returncode != 0]
    if failed:
        for s, r in failed:
            sys.stderr.write(f"---- {s}\n" + r.stdout + r.stderr)
        raise RuntimeError("building libpxfdtd.so failed: " + ", ".join(s for s, _ in failed))
    tmp = LIB + ".tmp"
    cmd = [hipcc, *FLAGS, "-shared", *[o for _, o, _ in results], "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("linking libpxfdtd.so failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
