"""Build-time check of libkhronos_hip's device code (host only; never run on a GPU).

Compiles csrc/khronos_hip.hip to gfx950 assembly and rejects instructions the
device code must not contain: stores and atomics through the scalar data
cache (and that cache's write-back / discard), and packed fp32 arithmetic.
"""
from __future__ import annotations

import os
import re
import subprocess
import sys
import tempfile

from build_lib import CSRC, HIPCC_FLAGS, hipcc

# Instructions the device code must not contain: stores and atomics through
# the scalar data cache (and its write-back), and packed fp32 arithmetic.
FORBIDDEN_ISA = re.compile(r"^\s+(s_store\w*|s_buffer_store\w*|s_scratch_store\w*|s_atomic\w*|"
                           r"s_buffer_atomic\w*|s_dcache_wb\w*|s_dcache_discard\w*|"
                           r"v_pk_add_f32|v_pk_mul_f32|v_pk_fma_f32)\b", re.M)


def check_isa(verbose: bool = False) -> None:
    """Compile the device code to gfx950 assembly and reject forbidden instructions."""
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "dev.s")
        flags = [f for f in HIPCC_FLAGS if f not in ("-shared", "-fPIC")]
        cmd = [hipcc(), *flags, "--cuda-device-only", "-S", os.path.join(CSRC, "khronos_hip.hip"), "-o", out]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("device assembly build failed")
        bad = sorted(set(m.group(1) for m in FORBIDDEN_ISA.finditer(open(out).read())))
        if bad:
            raise RuntimeError("device code contains forbidden instructions: " + ", ".join(bad))
        if verbose:
            print("device ISA check: no scalar-cache stores/atomics, no packed fp32", flush=True)



if __name__ == "__main__":
    check_isa(verbose=True)
