"""Build of the gfx950 library libswm.so (HIP kernels + C ABI) in-tree.

    python sw-mpdec_amd/build.py          # -> sw-mpdec_amd/lib/libswm.so

FP semantics are part of the product: -ffp-contract=off keeps every multiply
and add separately rounded (no fused multiply-add), as in the reference's
contraction-free build; no fast-math flag is ever passed.  Packed FP32 ops are
disabled (the library computes in FP64 only).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
LIB = LIBDIR / "libswm.so"

SOURCES = ["swm_kernels.hip", "swm_initial.hip", "swm_noalloc.hip", "swm_probe.hip",
           "swm_engine.cpp"]

HIPFLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and Path(c).exists():
            return c
    raise RuntimeError("hipcc not found")


def build_library(verbose: bool = False, extra: list[str] | None = None) -> Path:
    LIBDIR.mkdir(exist_ok=True)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPFLAGS, *(extra or []), f"-I{ROOT / 'include'}", f"-I{CSRC}",
           "-shared", "-o", str(tmp), *[str(CSRC / s) for s in SOURCES]]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed ({r.returncode}):\n{r.stdout}\n{r.stderr}")
    if verbose and r.stderr.strip():
        print(r.stderr, file=sys.stderr)
    os.replace(tmp, LIB)
    return LIB


FORTRAN = PKG / "fortran"
FDRIVER = LIBDIR / "swm_fdriver"


def flang() -> str | None:
    for c in (os.environ.get("FC"), "/opt/rocm/bin/amdflang", shutil.which("amdflang"),
              shutil.which("flang")):
        if c and Path(c).exists():
            return c
    return None


def build_fortran_driver(verbose: bool = False) -> Path | None:
    """The Fortran caller (fortran/swm_bridge.f90 + swm_fdriver.f90) linked
    against libswm.so with an $ORIGIN rpath; None when no Fortran compiler."""
    fc = flang()
    if fc is None:
        return None
    moddir = LIBDIR / "fmod"
    moddir.mkdir(parents=True, exist_ok=True)
    cmd = [fc, "-O2", "-ffp-contract=off", "-fdefault-real-8", "-fdefault-double-8",
           f"-J{moddir}", str(FORTRAN / "swm_bridge.f90"), str(FORTRAN / "swm_fdriver.f90"),
           f"-L{LIBDIR}", "-lswm", "-Wl,-rpath,$ORIGIN", "-o", str(FDRIVER)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=str(moddir))
    if r.returncode != 0:
        raise RuntimeError(f"flang failed ({r.returncode}):\n{r.stdout}\n{r.stderr}")
    return FDRIVER


if __name__ == "__main__":
    print(build_library(verbose=True))
    print(build_fortran_driver(verbose=True))
