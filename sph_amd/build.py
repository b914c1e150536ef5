"""Build recipe for libsph_hip.so (gfx950) and the CPU oracle.

``python -m sph_amd.build`` compiles, in tree:

* sph_amd/libsph_hip.so - HIP kernels + context + reference-named entry
  points + scene builder, for gfx950 only;
* sph_amd/sph_main - the C driver (a start_simulation-shaped loop);
* oracle/*.so (test infrastructure) via oracle/Makefile, and oracle/_ref
  when the reference sources are present.

Floating point: -ffp-contract=off everywhere (no fused multiply-add, so
every expression rounds exactly as the reference's C does), no fast-math,
IEEE-correct division and sqrt (HIP's default).  Packed FP32 instructions
are disabled for device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "sph_amd")
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libsph_hip.so")
DRIVER = os.path.join(PKG, "sph_main")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
C_FLAGS = ["-O2", "-fPIC", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-Wall"]

HIP_SRCS = ["sph_kernels.hip", "sph_ctx.hip", "sph_mlp.hip"]
C_SRCS = ["scene.c", "sph_shim.c"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")
    return r.stdout


def build_library(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in HIP_SRCS:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in C_SRCS:
        obj = os.path.join(BUILD, src.replace(".c", ".o"))
        _run(["gcc", *C_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_driver(verbose=False):
    src = os.path.join(CSRC, "sph_main.c")
    if not os.path.exists(src):
        return None
    _run(["gcc", *C_FLAGS, src, "-o", DRIVER, f"-L{PKG}", "-lsph_hip",
          "-Wl,-rpath,$ORIGIN", "-lm"])
    if verbose:
        print("built", DRIVER)
    return DRIVER


def build_oracle(verbose=False):
    """Test infrastructure: the CPU oracle, and oracle/_ref where the
    reference sources are available (this container only)."""
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-C", odir, "-B", "all"])
    if os.path.isdir("/root/reference/src"):
        _run(["make", "-s", "-C", odir, "-B", "ref"])
    if verbose:
        print("built oracle")


def build(verbose=False):
    build_library(verbose)
    build_driver(verbose)
    build_oracle(verbose)
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    build(verbose=True)
