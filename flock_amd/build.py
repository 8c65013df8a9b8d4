"""Build the native library (hipcc, gfx950) in-tree: flock_amd/lib/libflock_amd.so.

Compiles on a machine without a GPU (hipcc cross-compiles gfx950 code objects).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libflock_amd.so")
CLI = os.path.join(LIB_DIR, "flock")
SOURCES = ["csrc/runtime.hip", "csrc/server.hip", "csrc/uv_model.hip", "csrc/noalloc.hip"]
HEADERS = ["csrc/device_common.h", "csrc/engine.h", "csrc/runtime_internal.h",
           "../include/flock_amd.h", "../include/flock/structs.h", "../include/flock/uv_model_parallel.h"]

HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
               # packed FP32 forms are kept out of every code object
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-Wno-unused-result", "-Wno-unused-value"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    """Compile every source to an object in parallel (hipcc, gfx950), then link."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(LIB_DIR, exist_ok=True)
    obj_dir = os.path.join(HERE, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    srcs = [os.path.join(HERE, s) for s in SOURCES if os.path.exists(os.path.join(HERE, s))]

    def compile_one(src):
        obj = os.path.join(obj_dir, os.path.basename(src) + ".o")
        cmd = [hipcc(), *HIPCC_FLAGS, "-c", "-I", os.path.join(ROOT, "include"), "-o", obj, src]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        return obj, r

    with ThreadPoolExecutor(max_workers=min(4, len(srcs))) as pool:
        results = list(pool.map(compile_one, srcs))
    for obj, r in results:
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"hipcc failed compiling {obj}")
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *[o for o, _ in results]]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed linking libflock_amd.so")
    os.replace(tmp, LIB)
    cli_src = os.path.join(HERE, "csrc", "main.cpp")
    if os.path.exists(cli_src):
        cmd = [hipcc(), "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), cli_src,
               "-L", LIB_DIR, "-lflock_amd", "-Wl,-rpath,$ORIGIN", "-o", CLI]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("building the flock CLI failed")
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
