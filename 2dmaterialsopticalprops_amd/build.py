"""Build recipes: librpa_hip.so (hipcc, gfx950) and the Fortran drivers (flang).

Everything is built in-tree so the shared library travels with the sources.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "librpa_hip.so")
HBM_LIB = os.path.join(PKG, "libhbm_copy.so")  # measurement only: reference copy kernel
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/llvm/bin/flang")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I", os.path.join(ROOT, "include"),
]

SOURCES = ["rpa_api.hip", "k_vertex.hip", "k_vertex_col.hip", "k_accum.hip", "k_kk.hip", "k_dyson.hip", "k_interp.hip", "k_photon.hip", "k_ladder.hip",
           "rpa_host.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: " + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build_lib(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    jobs = []
    for src in SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = [HIPCC] + HIP_FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
        jobs.append(cmd)
        objs.append(obj)
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        for r in ex.map(_run, jobs):
            if verbose and (r.stdout or r.stderr):
                print(r.stdout + r.stderr)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    # the reference copy kernel the benchmark scales HBM counters with
    tmp = HBM_LIB + ".tmp"
    _run([HIPCC] + HIP_FLAGS + ["-shared", "-x", "hip", os.path.join(CSRC, "hbm_copy.hip"), "-o", tmp])
    os.replace(tmp, HBM_LIB)
    return LIB


FORTRAN = os.path.join(PKG, "fortran")
BIN = os.path.join(PKG, "bin")
F_MODULES = ["rpa_hip_f.f90", "rpa_files.f90"]
F_PROGRAMS = {"sloss_hip.x": "sloss_hip.f90", "optabs_hip.x": "optabs_hip.f90",
              "interpolate_hip.x": "interpolate_hip.f90", "photon_hip.x": "photon_hip.f90"}
F_FLAGS = ["-O2"]


def build_fortran(verbose: bool = False) -> dict:
    """The Fortran drivers (ISO_C_BINDING over librpa_hip.so) -> bin/*.x.
    They find the library next to the package through an $ORIGIN rpath."""
    if not os.path.exists(LIB):
        build_lib(verbose)
    fb = os.path.join(BUILD, "fortran")
    os.makedirs(fb, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    objs = []
    for src in F_MODULES:
        obj = os.path.join(fb, os.path.splitext(src)[0] + ".o")
        _run([FLANG] + F_FLAGS + ["-module-dir", fb, "-c", os.path.join(FORTRAN, src), "-o", obj])
        objs.append(obj)
    out = {}
    for exe, src in F_PROGRAMS.items():
        obj = os.path.join(fb, os.path.splitext(src)[0] + ".o")
        _run([FLANG] + F_FLAGS + ["-module-dir", fb, "-c", os.path.join(FORTRAN, src), "-o", obj])
        dst = os.path.join(BIN, exe)
        _run([FLANG, "-o", dst + ".tmp", obj] + objs
             + ["-L", PKG, "-lrpa_hip", "-Wl,-rpath,$ORIGIN/.."])
        os.replace(dst + ".tmp", dst)
        out[exe] = dst
    return out


if __name__ == "__main__":
    build_lib(verbose="-v" in sys.argv)
    print(LIB)
    for path in build_fortran(verbose="-v" in sys.argv).values():
        print(path)
