"""Build librumdeed.so (HIP, gfx950) in-tree.

    python -m rumdeed_amd.build

The library is plain hipcc output with a C ABI (include/rumdeed.h); it does
not link against PyTorch.  Floating-point contraction is disabled so that the
FP64 kernels evaluate the reference expressions operand for operand, and
packed-FP32 instruction selection is disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "librumdeed.so")

SOURCES = ["rd_api.hip", "rd_pair.hip", "rd_engine.hip", "rd_tail.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    return "hipcc"


def build_library(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), *srcs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building librumdeed.so")
    os.replace(tmp, LIB)
    return LIB


FORTRAN_DIR = os.path.join(PKG, "fortran")
FORTRAN_TESTS = os.path.join(FORTRAN_DIR, "rd_fortran_tests")


def flang() -> str | None:
    for cand in (os.environ.get("FLANG"), "/opt/rocm/lib/llvm/bin/flang", "/opt/rocm/bin/amdflang"):
        if cand and os.path.exists(cand):
            return cand
    return None


# The drop-in Fortran modules in dependency order; mod_global.F90 is the
# mirror of the reference's module variables that they use.
FORTRAN_MODULES = ["mod_global.F90", "rumdeed_hip.f90", "rd_device.f90", "mod_verlet.F90",
                   "mod_pair.F90", "mod_field_emission_v2.F90", "mod_field_thermo_emission.F90",
                   "mod_emission_tip.f90"]


def build_fortran(verbose: bool = False) -> str | None:
    """Compile the drop-in modules and their test driver with flang and link
    them against librumdeed.so (found next to the package at run time)."""
    fc = flang()
    if fc is None:
        return None
    bdir = os.path.join(FORTRAN_DIR, "_build")
    os.makedirs(bdir, exist_ok=True)
    srcs = [os.path.join(FORTRAN_DIR, f) for f in FORTRAN_MODULES + ["rd_fortran_tests.f90"]]
    cmd = [fc, "-O2", "-J", bdir, *srcs, "-L", PKG, "-lrumdeed",
           "-Wl,-rpath,$ORIGIN/..", "-o", FORTRAN_TESTS + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True, cwd=bdir)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("flang failed building the Fortran layer")
    os.replace(FORTRAN_TESTS + ".tmp", FORTRAN_TESTS)
    return FORTRAN_TESTS


if __name__ == "__main__":
    print(build_library(verbose=True))
    print(build_fortran(verbose=True))
