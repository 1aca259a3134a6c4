"""Builds the gfx950 HIP library libtm_hip.so in-tree with hipcc.

    python -m cistem_downstream_bah_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtm_hip.so")
SOURCES = ["tm_core.hip", "tm_prepare.hip"]
HEADERS = ["tm_fft.hpp", "tm_kernels.hpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 (v_pk_*_f32) stays off for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    cmd = [HIPCC, *FLAGS, "-o", LIB + ".tmp", *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libtm_hip.so")
    os.replace(LIB + ".tmp", LIB)
    # the C++ command-line driver over the library
    drv = os.path.join(HERE, "match_template_hip")
    cmd = ["g++", "-O2", "-std=c++17", "-o", drv + ".tmp", os.path.join(CSRC, "match_template_main.cpp"),
           "-L" + HERE, "-ltm_hip", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed building match_template_hip")
    os.replace(drv + ".tmp", drv)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
