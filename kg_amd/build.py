"""Build recipe of the HIP engine: compiles kg_amd/csrc into kg_amd/libkg_hip.so for gfx950."""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libkg_hip.so")

SOURCES = ["kg_capi.hip", "kg_kernels.hip", "kg_spectral.hip", "kg_diag.hip", "kg_initial.hip", "kg_refine.hip",
           "kg_probe.hip", "kg_metric.cpp", "kg_special.cpp"]

# -ffp-contract=off: no fused multiply-add is formed, so every operation
# rounds on its own as in the reference's Fortran expressions.
# -packed-fp32-ops: packed FP32 instructions are not emitted.
HIPFLAGS = [
    "-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    "-I" + os.path.join(ROOT, "include"),
]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


FORTRAN = os.path.join(HERE, "fortran")
FBUILD = os.path.join(FORTRAN, "_build")
FEXE = os.path.join(FORTRAN, "kg_evolve")


def build_fortran(verbose: bool = False) -> str | None:
    """The ISO_C_BINDING module and the example driver (needs amdflang)."""
    fc = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)
    if fc is None:
        return None
    os.makedirs(FBUILD, exist_ok=True)
    cmds = [
        [fc, "-O2", "-module-dir", FBUILD, "-c", os.path.join(FORTRAN, "kg_hip_mod.f90"), "-o",
         os.path.join(FBUILD, "kg_hip_mod.o")],
        [fc, "-O2", "-I", FBUILD, os.path.join(FORTRAN, "kg_evolve.f90"), os.path.join(FBUILD, "kg_hip_mod.o"),
         "-L" + HERE, "-lkg_hip", "-Wl,-rpath,$ORIGIN/..", "-o", FEXE],
    ]
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    return FEXE


def build(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPFLAGS + [os.path.join(CSRC, s) for s in SOURCES] + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
