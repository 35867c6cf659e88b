"""Builds libbdstep.so (HIP, gfx950) in-tree with hipcc.

The library is compiled from brownpak_amd/csrc/*.hip into
brownpak_amd/_build/libbdstep.so.  Objects are compiled in parallel and
linked once; every call rebuilds (the build check wants a fresh compile).
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libbdstep.so")
ARCH = os.environ.get("BD_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-variable",
    "-Wno-unused-function",
]


def _compile(src: str) -> str:
    obj = os.path.join(OUT, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def build_library(verbose: bool = False) -> str:
    os.makedirs(OUT, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    with cf.ThreadPoolExecutor(max_workers=min(len(srcs), 8)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_fortran_driver(verbose)
    return LIB


FLANG = os.environ.get("FLANG", "/opt/rocm/bin/amdflang")
FDIR = os.path.join(HERE, "fortran")
DRIVER = os.path.join(OUT, "bdrun")


def build_fortran_driver(verbose: bool = False) -> str:
    """Fortran driver (iso_c_binding) linked against libbdstep.so."""
    mod = os.path.join(OUT, "fmod")
    os.makedirs(mod, exist_ok=True)
    objs = []
    for src in ("bdstep_c.f90", "bdrun_io.f90", "bdrun.f90"):
        obj = os.path.join(mod, src + ".o")
        cmd = [FLANG, "-O2", "-c", os.path.join(FDIR, src), "-o", obj, "-module-dir", mod, "-I", mod]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"amdflang failed for {src}:\n{r.stderr[-4000:]}")
        objs.append(obj)
    cmd = [FLANG, *objs, "-o", DRIVER, "-L", OUT, "-lbdstep", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver link failed:\n{r.stderr[-4000:]}")
    if verbose:
        print("built", DRIVER)
    return DRIVER


if __name__ == "__main__":
    build_library(verbose=True)
