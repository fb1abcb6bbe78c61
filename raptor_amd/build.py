"""Builds the native parts in-tree for gfx950.

    python -m raptor_amd.build          # library + host driver

Outputs (git-ignored, they travel with the tree to the GPU):
    raptor_amd/libraptor_rt.so   the C-ABI library (include/raptor_rt.h)
    raptor_amd/host/raptor       the RAPTOR command-line driver on top of it
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
LIB = os.path.join(PKG, "libraptor_rt.so")
DRIVER = os.path.join(HOST, "raptor")

ARCH = os.environ.get("RAPTOR_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

# Packed FP32 instructions are disabled for all device code.
HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    # a * b + c is fused where one source expression writes it, and nowhere
    # else: the same arithmetic in every kernel that inlines a function, so
    # that launch shapes and frequency counts give the same bits
    "-ffp-contract=on",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _run(cmd, **kw):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, **kw)


def build_lib():
    src = os.path.join(CSRC, "rt_api.hip")
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, src])
    os.replace(tmp, LIB)
    return LIB


def build_driver():
    srcs = [os.path.join(HOST, f) for f in sorted(os.listdir(HOST)) if f.endswith(".c")]
    if not srcs:
        return None
    inc = os.path.join(ROOT, "include")
    _run(["gcc", "-O2", "-std=gnu99", "-Wall", "-Wextra", "-Wno-unused-parameter", "-I", inc,
          "-o", DRIVER, *srcs, "-L", PKG, "-lraptor_rt", "-Wl,-rpath,$ORIGIN/..", "-lm"])
    return DRIVER


def build_all():
    build_lib()
    build_driver()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    if "all" in what:
        build_all()
    else:
        if "lib" in what:
            build_lib()
        if "driver" in what:
            build_driver()
