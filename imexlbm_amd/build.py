"""Build the in-tree artefacts.

  lib/libimexlbm.so   the HIP kernels and the C ABI (hipcc, gfx950)
  bin/imexlbm         the C++ driver over the C ABI (System, LBM, main)

Everything is built in place so the files travel with the source tree.
Run as ``python -m imexlbm_amd.build``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib", "libimexlbm.so")
EXE = os.path.join(HERE, "bin", "imexlbm")
ARCH = os.environ.get("IMEXLBM_ARCH", "gfx950")

HIP_SOURCES = [os.path.join(HERE, "csrc", "imexlbm.hip")]
HIP_DEPS = HIP_SOURCES + [os.path.join(HERE, "csrc", f) for f in ("d3q27.hpp", "kernels.hpp")] + [
    os.path.join(ROOT, "include", "imexlbm.h")]
HOST_SOURCES = [os.path.join(HERE, "host", f) for f in ("System.cpp", "lbm.cpp", "output.cpp", "main.cpp")]

# -ffp-contract=off: the FP64 path reproduces the reference's separately
# rounded multiplies and adds; the FP32 path writes its fused multiply-adds
# explicitly.  Packed FP32 instructions are switched off for every kernel.
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return out


def build_lib() -> str:
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    tmp = LIB + ".tmp"
    _run([_hipcc(), *HIPCC_FLAGS, "-o", tmp, *HIP_SOURCES])
    os.replace(tmp, LIB)
    return LIB


def build_exe() -> str:
    os.makedirs(os.path.dirname(EXE), exist_ok=True)
    cxx = os.environ.get("CXX", "g++")
    _run([cxx, "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
          *HOST_SOURCES, "-o", EXE, "-L", os.path.dirname(LIB), "-limexlbm",
          "-Wl,-rpath,$ORIGIN/../lib"])
    return EXE


def build_all() -> None:
    build_lib()
    build_exe()


if __name__ == "__main__":
    build_all()
    print(LIB)
    print(EXE)
    sys.exit(0)
