"""Build libot2d.so (gfx950 HIP kernels + C-ABI) in-tree with hipcc.

The library is plain hipcc output; no torch extension machinery is involved,
so the same .so serves the Fortran driver, ctypes and any C caller.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libot2d.so")
SOURCES = ["ot2d_kernels.hip", "ot2d_turb.hip", "ot2d_sst.hip", "ot2d_overset.hip", "ot2d_ihc.cpp",
           "ot2d_api.hip"]
ARCH = os.environ.get("OT2D_ARCH", "gfx950")

# -ffp-contract=off keeps the FP64 path operand-for-operand with the
# reference; packed FP32 instructions are disabled for all device code.
HIPFLAGS = [
    "-O3", f"--offload-arch={ARCH}", "-fPIC", "-std=c++17",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    return "hipcc"


def build(verbose: bool = False) -> str:
    objs = []
    for src in SOURCES:
        s = os.path.join(CSRC, src)
        o = os.path.join(CSRC, src.rsplit(".", 1)[0] + ".o")
        cmd = [hipcc(), *HIPFLAGS, "-c", s, "-o", o]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(o)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
