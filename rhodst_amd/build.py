"""Builds the HIP extension rhodst_amd/_lib/librhodst_amd.so for gfx950.

Every translation unit under rhodst_amd/csrc is compiled with hipcc
(--offload-arch=gfx950, packed-FP32 instructions disabled) in parallel and
linked into one shared library exporting the C-ABI of include/rhodst.h.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
OBJ_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "librhodst_amd.so")
ARCH = os.environ.get("RHODST_OFFLOAD_ARCH", "gfx950")

SOURCES = ["mesh_host.cpp", "capi.cpp", "flux.hip", "jac.hip", "solve.hip", "update.hip", "sa.hip", "adjoint.hip", "wall.hip"]

CFLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-variable",
          "-Wno-unused-function", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I", os.path.join(HERE, "..", "include")]


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


# host-only geometry in plain IEEE arithmetic (no FMA contraction), so the
# fvMesh quantities are reproducible across compilers
EXTRA = {"mesh_host.cpp": ["-ffp-contract=off"]}


def _compile(src: str) -> str:
    obj = os.path.join(OBJ_DIR, os.path.basename(src) + ".o")
    cmd = [_hipcc(), *CFLAGS, *EXTRA.get(src, []), "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    os.makedirs(OBJ_DIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_abi_driver(verbose)
    return LIB


ROOT = os.path.dirname(HERE)
DRIVER_SRC = os.path.join(ROOT, "tests", "abi", "abi_driver.cpp")
DRIVER = os.path.join(ROOT, "tests", "abi", "build", "abi_driver")


def build_abi_driver(verbose: bool = False) -> str | None:
    """The C++ caller of the C-ABI used by the GPU tests: plain g++ against
    include/rhodst.h, linked to the library just built (rpath to _lib)."""
    if not os.path.exists(DRIVER_SRC):
        return None
    os.makedirs(os.path.dirname(DRIVER), exist_ok=True)
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"), DRIVER_SRC,
           "-L", OUT_DIR, "-lrhodst_amd", "-Wl,-rpath,$ORIGIN/../../../rhodst_amd/_lib", "-o", DRIVER]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed on the ABI driver:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print("built", DRIVER)
    return DRIVER


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
