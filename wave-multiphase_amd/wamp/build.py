"""Compile libwamp.so in-tree with hipcc for gfx950.

    python wave-multiphase_amd/wamp/build.py
"""
from __future__ import annotations

import os
import subprocess
import sys

try:
    from ._lib import CSRC_DIR, LIB_PATH
except ImportError:  # run as a script
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from wamp._lib import CSRC_DIR, LIB_PATH

REPO = os.path.dirname(os.path.dirname(CSRC_DIR))
INCLUDE = os.path.join(REPO, "include")
SOURCES = ["wamp_kernels.hip", "wamp_capi.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # FP64 parity: keep the reference's expression order, no FMA contraction
    "-ffp-contract=off",
    # no packed-FP32 instructions in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC_DIR, s) for s in SOURCES]
    tmp = LIB_PATH + ".tmp"
    cmd = [HIPCC, *FLAGS, "-I", INCLUDE, *srcs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed ({r.returncode})")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
