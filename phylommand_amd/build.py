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
    srcs = [os.path.join(CSRC, "phylo_hip.hip"), os.path.join(CSRC, "learn.hip")]
    tmp = LIB + ".tmp"
    run([HIPCC, *HIP_FLAGS, "-I" + os.path.join(ROOT, "include"), "-o", tmp, *srcs])
    os.replace(tmp, LIB)
    return LIB


HOST = os.path.join(HERE, "host")
HOST_SRCS = ["sub_model.cpp", "rate_model.cpp", "tree.cpp", "io.cpp", "simpleML.cpp", "optimize.cpp"]
HOST_LIB = os.path.join(HERE, "libphylommand_host.so")
HOST_DRIVER = os.path.join(HOST, "host_driver")
CXX_FLAGS = ["-std=c++17", "-O2", "-ffp-contract=off", "-Wall", "-Wno-unknown-pragmas", "-fPIC"]


def build_host():
    """The simpleML / sub_model / rate_model host classes over the C ABI, and
    their command-line driver."""
    srcs = [os.path.join(HOST, f) for f in HOST_SRCS]
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + HOST]
    link = ["-L" + HERE, "-lphylo_hip", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + HERE]
    tmp = HOST_LIB + ".tmp"
    run(["g++", *CXX_FLAGS, "-shared", *inc, "-o", tmp, *srcs, *link])
    os.replace(tmp, HOST_LIB)
    tmp = HOST_DRIVER + ".tmp"
    run(["g++", *CXX_FLAGS, *inc, "-o", tmp, os.path.join(HOST, "host_driver.cpp"), "-L" + HERE,
         "-lphylommand_host", "-lphylo_hip", "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath," + HERE])
    os.replace(tmp, HOST_DRIVER)


def build_host_sanitized(out_dir):
    """host_driver with the host classes compiled in, under AddressSanitizer
    and UndefinedBehaviorSanitizer (host C++ only, g++; it loads the normal
    libphylo_hip.so), for the CPU sanitizer test.  Returns its path."""
    os.makedirs(out_dir, exist_ok=True)
    exe = os.path.join(out_dir, "host_driver_san")
    srcs = [os.path.join(HOST, f) for f in ["host_driver.cpp", *HOST_SRCS]]
    run(["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
         "-fno-sanitize-recover=all", "-I" + os.path.join(ROOT, "include"), "-I" + HOST, "-o", exe, *srcs,
         "-L" + HERE, "-lphylo_hip", "-Wl,-rpath," + HERE])
    return exe


TREEATOR_DIR = os.path.join(HOST, "_treeator")
REF_SRC = os.environ.get("PHYLOMMAND_SRC", "/root/reference/src")


def build_treeator():
    """phylommand's own treeator (src/treeator.cpp, unmodified) built against
    the host classes by host/treeator.mk, when the phylommand sources are
    present; the binary lands in host/_treeator/."""
    if not os.path.isfile(os.path.join(REF_SRC, "treeator.cpp")):
        print(f"[build] {REF_SRC}/treeator.cpp not found: treeator binding not built", flush=True)
        return None
    run(["make", "-s", "-f", os.path.join(HOST, "treeator.mk"), "REF=" + REF_SRC, "OUT=" + TREEATOR_DIR])
    return os.path.join(TREEATOR_DIR, "treeator")


def build_all():
    build_hip()
    build_host()
    build_treeator()


if __name__ == "__main__":
    build_all()
