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
# (source, C++ standard): the refinement's descent is written as C++20 coroutines
SOURCES = [("tm_core.hip", "c++17"), ("tm_prepare.hip", "c++17"), ("tm_refine.hip", "c++20"), ("tm_cnn.hip", "c++17")]
# host-only sources (no device code), compiled with g++
HOST_SOURCES = ["tm_results.cpp"]
HEADERS = ["tm_fft.hpp", "tm_kernels.hpp", "tm_host.hpp", "tm_sizes.hpp", "tm_prep_kernels.hpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC",
    # packed FP32 (v_pk_*_f32) stays off for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd, verbose, what):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"{what} failed")


def build(verbose: bool = False) -> str:
    """Compiles each source to an object (in parallel) and links libtm_hip.so."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs, objs = [], []
    for src, std in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [HIPCC, *FLAGS, f"-std={std}", "-c", "-o", obj, os.path.join(CSRC, src)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    for src in HOST_SOURCES:
        obj = os.path.join(objdir, src.replace(".cpp", ".o"))
        objs.append(obj)
        cmd = ["g++", "-O2", "-fPIC", "-std=c++17", "-c", "-o", obj, os.path.join(CSRC, src)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            failed.append(src)
    if failed:
        raise RuntimeError("hipcc failed building " + ", ".join(failed))
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB + ".tmp", *objs], verbose, "linking libtm_hip.so")
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
