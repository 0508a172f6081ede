"""Builds the HIP/C++ engine into ``milldynamics3_amd/libmillgpu.so`` for gfx950.

Flags that matter for parity with the reference's Rust f32 arithmetic:
  -ffp-contract=off                          never fuse a*b+c (Rust never does)
  -fhip-fp32-correctly-rounded-divide-sqrt   IEEE f32 division and sqrt on the device
  -fno-gpu-flush-denormals-to-zero           keep f32 denormals like the CPU
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed FP32 instructions
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmillgpu.so")
ARCH = os.environ.get("MG_OFFLOAD_ARCH", "gfx950")

SOURCES = ["mg_box.hip", "mg_grid.hip", "mg_metrics.hip", "mg_stress.hip", "mg_dem.hip", "mg_surface.hip", "mg_fluid.hip", "mg_host.cpp", "mg_params.cpp"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math"]
HIP_ONLY = [f"--offload-arch={ARCH}", "-fhip-fp32-correctly-rounded-divide-sqrt",
            "-fno-gpu-flush-denormals-to-zero",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP engine cannot be built")


def build(verbose: bool = False) -> str:
    """Compiles every source (always, even if a library exists) and links the shared library."""
    cc = hipcc()
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(objdir, src + ".o")
        cmd = [cc] + COMMON + (HIP_ONLY if src.endswith(".hip") else ["-x", "c++"]) + ["-c", path, "-o", obj]
        if src.endswith(".cpp"):
            cmd = [cc] + COMMON + ["-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError(f"compilation of {src} failed")
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp] + objs
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("linking libmillgpu.so failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
