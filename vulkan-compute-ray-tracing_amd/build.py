"""Builds librtamd.so (HIP for gfx950 + host C++) in-tree, and the C++
host program rtamd_frame_loop (host/frame_loop_main.cpp over rtamd.hpp)
linked against it.

    python vulkan-compute-ray-tracing_amd/build.py [--jobs N]

Every source is compiled with -ffp-contract=off so no multiply-add is fused
(the render must reproduce the shader's float operations bit for bit), and
with packed FP32 instructions disabled for the device.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "librtamd.so")
HOST_MAIN = os.path.join(PKG, "host", "frame_loop_main.cpp")
HOST_BIN = os.path.join(PKG, "rtamd_frame_loop")
ARCH = os.environ.get("RTAMD_ARCH", "gfx950")

SOURCES = [
    "rt_api.cpp",
    "rt_scene.cpp",
    "rt_bvh.cpp",
    "rt_workloads.cpp",
    "rt_trace.hip",
    "rt_render.hip",
    "rt_batch.hip",
    "rt_build.hip",
    "rt_frame.hip",
    "rt_sphere.hip",
    "rt_image.cpp",
    "rt_obj.cpp",
    "rt_server.cpp",
    "rt_noalloc.hip",
]

COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc() -> str:
    p = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(p):
        raise RuntimeError("hipcc not found: ROCm is required to build librtamd")
    return p


def compile_one(src: str) -> str:
    path = os.path.join(CSRC, src)
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), "-c", path, "-o", out] + COMMON + DEVICE
    if src.endswith(".hip"):
        cmd += ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return out


def build(jobs: int = 8, verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_host_program()
    if verbose:
        print(f"built {LIB} and {HOST_BIN}")
    return LIB


def build_host_program() -> str:
    """g++ over the C-ABI headers only: no HIP header or flag is needed by a
    host program that uses the library."""
    gxx = shutil.which("g++") or shutil.which("c++")
    if not gxx:
        raise RuntimeError("g++ not found: the host program needs a C++17 compiler")
    tmp = HOST_BIN + ".tmp"
    cmd = [gxx, "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
           HOST_MAIN, "-o", tmp, "-L", PKG, "-lrtamd", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"host program build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, HOST_BIN)
    return HOST_BIN


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    a = ap.parse_args()
    try:
        build(a.jobs, verbose=True)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
