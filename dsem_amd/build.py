"""Build libdsem_hip.so in-tree with hipcc for gfx950.

    python -m dsem_amd.build

The library is compiled from dsem_amd/csrc (HIP kernels, host set-up and the
C ABI declared in include/dsem_hip.h).  All HIP code is compiled with packed
FP32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libdsem_hip.so")
ARCH = os.environ.get("DSEM_OFFLOAD_ARCH", "gfx950")

SOURCES = ["dsem_hip.hip", "geom.cpp"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    llvm_lib = "/opt/rocm/lib/llvm/lib"
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-fopenmp", f"-Wl,-rpath,{llvm_lib}", "-Wall", "-Wno-unused-function",
           "-I", os.path.join(ROOT, "include")]
    cmd += [os.path.join(CSRC, s) for s in SOURCES]
    cmd += ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    out = subprocess.run(cmd, capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + out.stdout + out.stderr)
    if verbose and out.stderr.strip():
        # the host pass warns about the packed-fp32 feature flag; that is harmless
        lines = [l for l in out.stderr.splitlines() if "packed-fp32-ops" not in l]
        if lines:
            print("\n".join(lines), file=sys.stderr)
    os.replace(tmp, LIB)
    return LIB




def build_fortran(verbose: bool = False) -> str:
    """Compile the iso_c_binding module and the Fortran test driver
    (tests/fortran/tgv_drive) against libdsem_hip.so with flang."""
    flang = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
    if not os.path.exists(flang):
        raise RuntimeError("flang not found")
    fdir = os.path.join(HERE, "fortran")
    tdir = os.path.join(ROOT, "tests", "fortran")
    obj = os.path.join(fdir, "dsem_hip.o")
    exe = os.path.join(tdir, "tgv_drive")
    cmds = [[flang, "-O2", "-c", os.path.join(fdir, "dsem_hip.f90"), "-module-dir", fdir, "-o", obj],
            [flang, "-O2", "-I", fdir, "-module-dir", tdir, os.path.join(tdir, "tgv_drive.f90"), obj,
             "-L", HERE, "-ldsem_hip", "-Wl,-rpath,$ORIGIN/../../dsem_amd", "-o", exe]]
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        out = subprocess.run(cmd, capture_output=True, text=True)
        if out.returncode != 0:
            raise RuntimeError("flang failed:\n" + out.stdout + out.stderr)
    return exe


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_fortran(verbose=True))
