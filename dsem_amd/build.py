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


if __name__ == "__main__":
    print(build(verbose=True))
