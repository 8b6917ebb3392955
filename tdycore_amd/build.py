"""Builds libtdyhip.so (HIP, gfx950) in-tree and the CPU oracle library.

``python -m tdycore_amd.build`` compiles everything; ``build_library()``
compiles only the product library.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtdyhip.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-munsafe-fp-atomics",
    # no packed-FP32 instructions anywhere in this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for p in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if p and os.path.exists(p):
            return p
    raise RuntimeError("hipcc not found")


def build_library(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, "capi.hip"), os.path.join(CSRC, "topology.cpp")]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", os.path.join(ROOT, "include"), "-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libtdyhip.so")
    os.replace(tmp, LIB)
    return LIB


def build_oracle(verbose: bool = False) -> str:
    odir = os.path.join(ROOT, "oracle")
    r = subprocess.run(["make", "-C", odir, "-B", "-s"], capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("oracle build failed")
    return os.path.join(odir, "liboracle.so")


def build_all(verbose: bool = False):
    build_library(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
    print("built", LIB)
