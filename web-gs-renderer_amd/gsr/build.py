"""Builds the renderer library (libgsr.so, gfx950) and the Node N-API addon.

    python -m gsr.build          (from web-gs-renderer_amd/)

Outputs go to web-gs-renderer_amd/build/.  Every HIP source is compiled with
-ffp-contract=off (the FP32 operation order written in the kernels is the one
executed) and with packed FP32 instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(PKG_DIR, "csrc")
NAPI = os.path.join(PKG_DIR, "napi")
BUILD = os.path.join(PKG_DIR, "build")
OBJ = os.path.join(BUILD, "obj")
LIB = os.path.join(BUILD, "libgsr.so")
ADDON = os.path.join(BUILD, "gsr_napi.node")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    "-O3", f"--offload-arch={ARCH}", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-value", "-Wno-unused-result",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}")
    return r.stdout


def build_library(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    sources = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    if not sources:
        raise RuntimeError(f"no HIP sources in {CSRC}")

    def compile_one(src):
        obj = os.path.join(OBJ, os.path.basename(src)[:-4] + ".o")
        out = _run([HIPCC, *HIP_FLAGS, "-c", src, "-o", obj])
        if verbose and out.strip():
            print(out)
        return obj

    workers = min(len(sources), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(workers) as ex:
        objs = list(ex.map(compile_one, sources))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


CALIB_SRC = os.path.join(PKG_DIR, "calib", "valu_probe.hip")
CALIB_LIB = os.path.join(BUILD, "libgsr_calib.so")


def build_calibration() -> str:
    """Counter-calibration kernel for bench.py --valu-probe (not part of the renderer)."""
    os.makedirs(BUILD, exist_ok=True)
    tmp = CALIB_LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", CALIB_SRC, "-o", tmp])
    os.replace(tmp, CALIB_LIB)
    return CALIB_LIB


def build_addon() -> str | None:
    """C N-API addon (Node >= 12) linked against libgsr.so; skipped without Node headers."""
    src = os.path.join(NAPI, "gsr_napi.c")
    inc = "/usr/include/node"
    if not (os.path.exists(src) and os.path.exists(os.path.join(inc, "node_api.h"))):
        return None
    tmp = ADDON + ".tmp"
    _run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-Wall", "-DNODE_GYP_MODULE_NAME=gsr_napi", f"-I{inc}",
          f"-I{os.path.join(os.path.dirname(PKG_DIR), 'include')}", src, "-o", tmp,
          f"-L{BUILD}", "-lgsr", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, ADDON)
    return ADDON


def build(verbose: bool = False):
    lib = build_library(verbose)
    addon = build_addon()
    build_calibration()
    return lib, addon


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
