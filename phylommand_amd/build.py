"""Builds libphylo_hip.so (gfx950) in-tree and the host C++ library/tools.

    python -m phylommand_amd.build          # HIP library + host C++ library
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libphylo_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950",
    # Packed FP32 instructions are kept out of every kernel.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_hip(verbose: bool = True) -> str:
    srcs = [os.path.join(CSRC, "phylo_hip.hip")]
    tmp = LIB + ".tmp"
    run([HIPCC, *HIP_FLAGS, "-I" + os.path.join(ROOT, "include"), "-o", tmp, *srcs])
    os.replace(tmp, LIB)
    return LIB


def build_all():
    build_hip()


if __name__ == "__main__":
    build_all()
