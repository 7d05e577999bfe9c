"""Build libkhronos_hip.so for gfx950 in-tree (khronos.jl_amd/lib/).

    python khronos.jl_amd/build_lib.py

hipcc cross-compiles for gfx950 without a GPU present.  Flags:
  -ffp-contract=off    no a*b+c -> FMA contraction: the EXACT mode reproduces
                       the reference CPU path's roundings
  -packed-fp32-ops off no v_pk_* fp32 instructions
  code object v5       loadable by the HIP runtime torch ships (ROCm 7.0) as
                       well as by /opt/rocm (7.2)
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libkhronos_hip.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-mcode-object-version=5",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc") or ""):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, os.path.join(CSRC, "khronos_hip.hip"), "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("libkhronos_hip build failed")
    noisy = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if noisy and verbose:
        sys.stderr.write("\n".join(noisy) + "\n")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
