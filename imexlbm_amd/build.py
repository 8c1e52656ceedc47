"""Build the in-tree artefacts.

  lib/libimexlbm.so     the HIP kernels and the C ABI (hipcc, gfx950)
  bin/imexlbm           the C++ driver over the C ABI (System, LBM, main),
                        scaling build (-DLBM_NO_PERSTEP_TIMING) against the
                        one-rank mpi.h of host/mpi_stub
  bin/imexlbm_perstep   the same driver as the reference's default build:
                        the per-step "collision | ... GFLOPS" line
  bin/imexlbm_mpi       the scaling build against a real MPI (mpicxx), when
                        one is installed

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
EXE_PERSTEP = os.path.join(HERE, "bin", "imexlbm_perstep")
EXE_MPI = os.path.join(HERE, "bin", "imexlbm_mpi")
MPI_STUB = os.path.join(HERE, "host", "mpi_stub")
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


def _run(cmd, env=None):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env)
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


def mpicxx():
    """The MPI C++ compiler wrapper, or None."""
    for cand in (os.environ.get("MPICXX"), shutil.which("mpicxx"), "/opt/conda/bin/mpicxx"):
        if cand and os.path.exists(cand):
            return cand
    return None


def mpi_flags():
    """Include and library options of the installed MPI, from its wrapper's
    -show output: only the C interface is used.  The MPI library directory
    goes last in the run path so that the host's C++ runtime, which the HIP
    runtime needs, is found before an older one shipped beside the MPI."""
    cxx = mpicxx()
    if cxx is None:
        return None
    out = subprocess.run([cxx, "-show"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True).stdout
    inc, libdirs, libs = [], [], []
    for tok in out.split():
        if tok.startswith("-I") and tok not in inc:
            inc.append(tok)
        elif tok.startswith("-L") and tok[2:] not in libdirs:
            libdirs.append(tok[2:])
        elif tok.startswith("-l") and tok != "-lmpicxx":
            libs.append(tok)
    if not libs:
        return None
    return inc, libdirs, libs


def build_exe(out=EXE, perstep=False, mpi=False) -> str:
    os.makedirs(os.path.dirname(out), exist_ok=True)
    cxx = os.environ.get("CXX", "g++")
    rpath = ["$ORIGIN/../lib"]
    if mpi:
        flags = mpi_flags()
        if flags is None:
            raise RuntimeError("no MPI compiler wrapper found")
        inc, libdirs, libs = flags
        sysdir = "/usr/lib/x86_64-linux-gnu"
        rpath += [sysdir] + libdirs
        # the MPI libraries by path: a -L for their directory would also
        # hand the linker the older C++ runtime kept there
        link = []
        for lib in libs:
            name = "lib" + lib[2:] + ".so"
            hit = next((os.path.join(d, name) for d in libdirs if os.path.exists(os.path.join(d, name))), None)
            link.append(hit or lib)
        link += ["-Wl,-rpath-link," + ":".join([sysdir] + libdirs)]
    else:
        inc, link = ["-I", MPI_STUB], []
    defs = [] if perstep else ["-DLBM_NO_PERSTEP_TIMING"]
    _run([cxx, "-O2", "-std=c++17", "-Wall", *inc, "-I", os.path.join(ROOT, "include"), *defs,
          *HOST_SOURCES, "-o", out, "-L", os.path.dirname(LIB), "-limexlbm", *link,
          "-Wl,--disable-new-dtags", "-Wl,-rpath," + ":".join(rpath)])
    return out


def build_all() -> None:
    build_lib()
    build_exe(EXE)
    build_exe(EXE_PERSTEP, perstep=True)
    if mpicxx():
        build_exe(EXE_MPI, mpi=True)


if __name__ == "__main__":
    build_all()
    print(LIB)
    print(EXE)
    print(EXE_PERSTEP)
    sys.exit(0)
