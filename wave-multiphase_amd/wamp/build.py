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
OBJ_DIR = os.path.join(CSRC_DIR, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # no packed-FP32 instructions in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
# (source, contraction): the FP64 kernels and the C-ABI keep the reference's
# rounding (no FMA contraction, FP64 parity); the FP32-mixed face kernels
# contract a*b+c within an expression.
UNITS = [
    ("wamp_kernels.hip", "-ffp-contract=off"),
    ("wamp_faces_mixed.hip", "-ffp-contract=on"),
    ("wamp3_kernels.hip", "-ffp-contract=off"),
    ("wamp3_faces_mixed.hip", "-ffp-contract=on"),
    ("wamp_capi.cpp", "-ffp-contract=off"),
    ("wamp_noalloc.hip", "-ffp-contract=off"),
    ("wamp_cpu.hip", "-ffp-contract=off"),  # the CPU device: host code only
]


def _run(cmds, verbose):
    procs = []
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                      text=True))
    failed = False
    for p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            failed = True
    if failed:
        raise RuntimeError("hipcc failed")


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ_DIR, exist_ok=True)
    objs, cmds = [], []
    for src, contract in UNITS:
        obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *COMMON, contract, "-I", INCLUDE, "-c",
                     os.path.join(CSRC_DIR, src), "-o", obj])
    _run(cmds, verbose)  # the units compile in parallel
    tmp = LIB_PATH + ".tmp"
    _run([[HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]], verbose)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
