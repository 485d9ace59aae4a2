"""Compiles the HIP library libvvmi.so for gfx950 in-tree (no JIT cache)."""
from __future__ import annotations

import os
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libvvmi.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["vv_explicit.hip", "vv_mg.hip", "vv3.hip", "vv3_poisson.hip", "vv_band.hip", "vvt.hip", "vvm.hip", "vvs.hip", "vv_api.cpp", "vv_serve.cpp"]
# -ffp-contract=off keeps the explicit-step arithmetic identical to the
# reference's operation order; packed FP32 instructions are disabled.
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument"]


FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
DROPIN_SRC = os.path.join(PKG, "fortran", "vv_dropin.F90")
DROPIN_LIB = os.path.join(PKG, "libvvdropin.so")


DROPIN_ANN_LIB = os.path.join(PKG, "libvvdropin_ann.so")


def build_dropin(annulus: bool = False) -> str:
    """The Fortran drop-in (reference routine names over the C ABI) as a
    shared library next to libvvmi.so: velvor2dperiodicrbconv.f's routines,
    or with annulus=True velvor2dnewsch3.4.f's (same file, -DVV_GEOM=1 and
    that program's array extents)."""
    mod = os.path.join(PKG, "fortran", "_build", "ann" if annulus else "box")
    os.makedirs(mod, exist_ok=True)
    lib = DROPIN_ANN_LIB if annulus else DROPIN_LIB
    defs = ["-DVV_GEOM=1", "-DVV_KR=420", "-DVV_KP=3600"] if annulus else []
    cmd = [FLANG, "-O2", "-fPIC", "-shared", "-module-dir", mod, *defs, "-o", lib + ".tmp",
           DROPIN_SRC, "-L" + PKG, "-lvvmi", "-Wl,-rpath,$ORIGIN"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("flang failed:\n" + res.stdout + res.stderr)
    os.replace(lib + ".tmp", lib)
    return lib


DROPIN3_SRC = os.path.join(PKG, "fortran", "vv_dropin3.F90")
DROPIN3_LIB = os.path.join(PKG, "libvvdropin3.so")


def build_dropin3(hor: bool = False) -> str:
    """The 3D Fortran drop-in (rot3dmg.f routine names over include/vvmi3.h)
    as a shared library next to libvvmi.so; its routine names overlap the
    2D drop-in's, so it is a library of its own.  hor=True builds the
    drop-in of rot3dmghor.f (libvvdropin3h.so: model VV3_MODEL_HOR and
    heat(trans1,trans2))."""
    lib = os.path.join(PKG, "libvvdropin3h.so") if hor else DROPIN3_LIB
    mod = os.path.join(PKG, "fortran", "_build", "d3h" if hor else "d3")
    os.makedirs(mod, exist_ok=True)
    cmd = [FLANG, "-O2", "-fPIC", "-shared", "-cpp", "-DVV3_HOR=%d" % int(hor), "-module-dir", mod,
           "-I" + os.path.join(PKG, "fortran"), "-o", lib + ".tmp", DROPIN3_SRC, "-L" + PKG, "-lvvmi",
           "-Wl,-rpath,$ORIGIN"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("flang failed:\n" + res.stdout + res.stderr)
    os.replace(lib + ".tmp", lib)
    return lib


def build_dropin3_driver(root: str) -> str:
    """Test driver (tests/fortran/vv3_driver.f90) linked against the 3D
    drop-in."""
    src = os.path.join(root, "tests", "fortran", "vv3_driver.f90")
    out_dir = os.path.join(root, "tests", "fortran", "_build")
    os.makedirs(out_dir, exist_ok=True)
    exe = os.path.join(out_dir, "vv3_driver")
    cmd = [FLANG, "-O2", "-module-dir", out_dir, "-o", exe, src, "-L" + PKG, "-lvvdropin3",
           "-lvvmi", "-Wl,-rpath,$ORIGIN/../../../codesthesis_dir_amd"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("flang failed:\n" + res.stdout + res.stderr)
    return exe


def build_dropin_driver(root: str) -> str:
    """Test driver (tests/fortran/vv_driver.f90) linked against the drop-in."""
    src = os.path.join(root, "tests", "fortran", "vv_driver.f90")
    out_dir = os.path.join(root, "tests", "fortran", "_build")
    os.makedirs(out_dir, exist_ok=True)
    exe = os.path.join(out_dir, "vv_driver")
    cmd = [FLANG, "-O2", "-module-dir", out_dir, "-o", exe, src, "-L" + PKG, "-lvvdropin",
           "-lvvmi", "-Wl,-rpath,$ORIGIN/../../../codesthesis_dir_amd"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("flang failed:\n" + res.stdout + res.stderr)
    return exe


def build_library(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, "-o", tmp, *[os.path.join(CSRC, s) for s in SOURCES]]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    if verbose:
        print(res.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_library(verbose=True))
    print(build_dropin())
    print(build_dropin(annulus=True))
    print(build_dropin3())
    print(build_dropin3(hor=True))
