"""Build step for libccd_amd.so: hipcc for gfx950, one object per translation unit, in-tree."""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libccd_amd.so")
LOADGEN = os.path.join(PKG, "serve_loadgen")
OBJDIR = os.path.join(PKG, "_build")
SOURCES = ["narrowphase.hip", "workcount.hip", "broadphase.hip", "distance.hip", "devprims.hip", "server.hip",
           "capi.hip", "velocity_filter.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: the reference's arithmetic has no fused multiply-add; every a*b+c must round twice.
# The packed-fp32 target feature is switched off for all device code.
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-function",
         "-Wno-unused-result"]


def _compile(src: str) -> str:
    obj = os.path.join(OBJDIR, src.replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    # The serving load generator: a host program over the C ABI (g++, links the library beside it).
    tmp = LOADGEN + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-pthread", "-o", tmp, os.path.join(PKG, "loadgen", "serve_loadgen.cpp"),
           "-L" + PKG, "-lccd_amd", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"load generator build failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LOADGEN)
    if verbose:
        print(f"built {LIB} and {LOADGEN}", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
