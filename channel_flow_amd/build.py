"""Builds, in-tree, libchf.so (HIP, gfx950) and the Fortran driver that calls
it through ISO_C_BINDING:

    channel_flow_amd/lib/libchf.so
    channel_flow_amd/lib/poiseuilleflow_gpu     (amdflang; rpath $ORIGIN)

Usage: python -m channel_flow_amd.build [--jobs N] [--no-fortran]
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
LIBDIR = os.path.join(PKG, "lib")
OBJDIR = os.path.join(ROOT, "_build", "obj")
LIB = os.path.join(LIBDIR, "libchf.so")

SOURCES = ["api.hip", "fft.hip", "cheb.hip", "solve.hip", "stats.hip", "ydct.hip"]
FORTRAN_DIR = os.path.join(PKG, "fortran")
FORTRAN_SOURCES = ["chf_binding.f90", "paraandcons_gpu.f90", "poiseuilleflow_gpu.f90"]  # in module order
DRIVER = os.path.join(LIBDIR, "poiseuilleflow_gpu")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # packed FP32 instructions stay off: some of their forms have given wrong
    # results on this hardware now and then, with no error
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I" + os.path.join(ROOT, "include"),
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src):
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    cmd = [hipcc()] + HIPCC_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def flang():
    for cand in (os.environ.get("FLANG"), "/opt/rocm/bin/amdflang", shutil.which("amdflang"),
                 shutil.which("flang")):
        if cand and os.path.exists(cand):
            return cand
    return None


def build_fortran(verbose: bool = False) -> str:
    """Compiles the Fortran driver against the freshly built libchf.so."""
    fc = flang()
    if fc is None:
        raise RuntimeError("no Fortran compiler (amdflang) found")
    moddir = os.path.join(ROOT, "_build", "mod")
    os.makedirs(moddir, exist_ok=True)
    objs = []
    for src in FORTRAN_SOURCES:
        obj = os.path.join(OBJDIR, src + ".o")
        cmd = [fc, "-O2", "-J" + moddir, "-c", os.path.join(FORTRAN_DIR, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        objs.append(obj)
    tmp = DRIVER + ".tmp"
    cmd = [fc, "-o", tmp] + objs + ["-L" + LIBDIR, "-lchf", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, DRIVER)
    if verbose:
        print("built", DRIVER)
    return DRIVER


def build(jobs: int = 4, verbose: bool = False, fortran: bool = True) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    if fortran:
        build_fortran(verbose)
    return LIB


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("--no-fortran", action="store_true")
    a = ap.parse_args(argv)
    build(a.jobs, verbose=True, fortran=not a.no_fortran)


if __name__ == "__main__":
    sys.exit(main())
