"""The C++ boundary: a driver written against the reference's interface
(MPI_Init, LBM(MPI_COMM_WORLD, sx, sy, sz, tau, rho0, u0), Initialize,
Collision, Streaming, the collective Conserved, the lattice names q, dim and
ghost; src/main.cpp:6-107, src/lbm.hpp:22-24,179-203) compiles and links
against the drop-in headers and libimexlbm.so, both with the one-rank mpi.h
and with the installed MPI, and runs up to the device.  Without a GPU the
library's own error ends the run; with one the run completes.  More than one
rank is refused with a message before any device call."""
import os
import shutil
import subprocess

import pytest

from tests.conftest import ROOT

HOST = os.path.join(ROOT, "imexlbm_amd", "host")
LIBDIR = os.path.join(ROOT, "imexlbm_amd", "lib")
DRIVER = os.path.join(ROOT, "tests", "drivers", "api_driver.cpp")
SOURCES = [os.path.join(HOST, f) for f in ("System.cpp", "lbm.cpp", "output.cpp")] + [DRIVER]
DECK = "1.0  0.1  1600\n0.05  4  1\n16  16  16\n"


def _need_lib():
    if not os.path.exists(os.path.join(LIBDIR, "libimexlbm.so")):
        pytest.fail("libimexlbm.so is not built: run __graft_entry__.build() first")


def _compile(out, mpi):
    from imexlbm_amd import build
    rpath = ["$ORIGIN"]
    if mpi:
        flags = build.mpi_flags()
        if flags is None:
            pytest.skip("no MPI installed")
        inc, libdirs, libs = flags
        link = [os.path.join(d, "lib" + l[2:] + ".so") for l in libs for d in libdirs
                if os.path.exists(os.path.join(d, "lib" + l[2:] + ".so"))]
        link += ["-Wl,-rpath-link,/usr/lib/x86_64-linux-gnu:" + ":".join(libdirs)]
        rpath += ["/usr/lib/x86_64-linux-gnu"] + libdirs
    else:
        inc, link = ["-I", os.path.join(HOST, "mpi_stub")], []
    cmd = ["g++", "-O1", "-std=c++17", "-Wall", "-Werror", *inc, "-I", HOST, "-I",
           os.path.join(ROOT, "include"), *SOURCES, "-o", str(out), "-L", LIBDIR, "-limexlbm", *link,
           "-Wl,--disable-new-dtags", "-Wl,-rpath," + ":".join(rpath + [LIBDIR])]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr


def _check_run(r):
    out = r.stdout + r.stderr
    assert "3D Taylor-Green Vortex" in out and "nx,ny,nz = 16 16 16" in out
    # the geometry fields of the reference: lx = glx + 2*ghost, interior [3, 19)
    assert "driver: ranks 1 lx 22 ly 22 lz 22 l_s 3 l_e 19 cs2 0.33333333333333331" in out
    if r.returncode == 0:
        assert "cons  step      4 | mass " in out
    else:
        assert "FATAL: imex_create:" in out, out[-2000:]


def test_reference_style_driver_against_one_rank_mpi_stub(tmp_path):
    _need_lib()
    exe = tmp_path / "driver_stub"
    _compile(exe, mpi=False)
    (tmp_path / "input.in").write_text(DECK)
    r = subprocess.run([str(exe)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    _check_run(r)


def _mpirun():
    for cand in (shutil.which("mpirun"), "/opt/conda/bin/mpirun"):
        if cand and os.path.exists(cand):
            return cand
    pytest.skip("no mpirun")


def test_reference_style_driver_against_installed_mpi(tmp_path):
    _need_lib()
    exe = tmp_path / "driver_mpi"
    _compile(exe, mpi=True)
    (tmp_path / "input.in").write_text(DECK)
    r = subprocess.run([_mpirun(), "-n", "1", str(exe)], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    _check_run(r)


def test_more_than_one_rank_is_refused_before_the_device(tmp_path):
    _need_lib()
    exe = tmp_path / "driver_mpi2"
    _compile(exe, mpi=True)
    (tmp_path / "input.in").write_text(DECK)
    r = subprocess.run([_mpirun(), "-n", "2", str(exe)], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0
    assert "run it with one MPI rank (got 2)" in r.stdout + r.stderr
    assert "imex_create" not in r.stdout + r.stderr


def test_mpi_stub_one_rank_semantics(tmp_path):
    src = tmp_path / "stub.c"
    src.write_text(r'''
#include <mpi.h>
#include <stdio.h>
int main(int argc, char **argv) {
    int r = -1, n = -1, init = 0;
    double a[5] = {1, 2, 3, 4, 5}, b[5] = {0};
    MPI_Init(&argc, &argv);
    MPI_Initialized(&init);
    MPI_Comm_rank(MPI_COMM_WORLD, &r);
    MPI_Comm_size(MPI_COMM_WORLD, &n);
    MPI_Allreduce(a, b, 5, MPI_DOUBLE, MPI_SUM, MPI_COMM_WORLD);
    MPI_Allreduce(MPI_IN_PLACE, a, 5, MPI_DOUBLE, MPI_SUM, MPI_COMM_WORLD);
    double t0 = MPI_Wtime(), t1 = MPI_Wtime();
    MPI_Barrier(MPI_COMM_WORLD);
    MPI_Finalize();
    printf("%d %d %d %g %g %g %d\n", init, r, n, b[0], b[4], a[2], t1 >= t0);
    return 0;
}
''')
    exe = tmp_path / "stub"
    r = subprocess.run(["gcc", "-std=c99", "-D_POSIX_C_SOURCE=199309L", "-Wall", "-Werror", "-I",
                        os.path.join(HOST, "mpi_stub"), str(src), "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=30).stdout.split()
    assert out == ["1", "0", "1", "1", "5", "3", "1"]
