"""ctypes wrapper of the FP64 CPU oracle (oracle/imex_oracle.c).

Test infrastructure only: imported by tests/, by __graft_entry__.smoke() and by
the CPU-baseline leg of bench.py, always as the checker or the timed CPU
baseline, never as part of the product path.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

Q = 27


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE, "liboracle.so"],
                       check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.oracle_tau.restype = d
        L.oracle_tau.argtypes = [d, i, i]
        L.oracle_tables.argtypes = [dp, ip, ip]
        L.oracle_decompose.argtypes = [i, i, i, i, ip]
        L.oracle_local_len.argtypes = [i, i, i]
        L.oracle_block_lo.argtypes = [i, i, i]
        L.oracle_create.restype = vp
        L.oracle_create.argtypes = [i, i, i, d, d, d, i]
        L.oracle_destroy.argtypes = [vp]
        for name in ("oracle_initialize", "oracle_collision", "oracle_streaming"):
            getattr(L, name).argtypes = [vp]
        L.oracle_step.argtypes = [vp, i]
        L.oracle_conserved.argtypes = [vp, dp]
        L.oracle_conserved_compensated.argtypes = [vp, dp]
        L.oracle_get_f.argtypes = [vp, dp]
        L.oracle_set_f.argtypes = [vp, dp]
        L.oracle_rank_geometry.argtypes = [vp, i, ip]
        L.oracle_grid.argtypes = [vp, ip]
        L.oracle_set_threads.argtypes = [i]
        L.oracle_max_threads.restype = i
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def tau(u0: float, sx: int, re: int) -> float:
    return lib().oracle_tau(float(u0), int(sx), int(re))


def tables():
    w = np.zeros(Q, np.float64)
    e = np.zeros((Q, 3), np.int32)
    opp = np.zeros(Q, np.int32)
    lib().oracle_tables(_dp(w), _ip(e), _ip(opp))
    return w, e.astype(np.int64), opp.astype(np.int64)


def decompose(nranks: int, glx: int, gly: int, glz: int):
    out = np.zeros(3, np.int32)
    lib().oracle_decompose(nranks, glx, gly, glz, _ip(out))
    return tuple(int(v) for v in out)


def set_threads(n: int) -> None:
    lib().oracle_set_threads(int(n))


def max_threads() -> int:
    return int(lib().oracle_max_threads())


class Oracle:
    """The reference's LBM object on the CPU: Initialize / Collision /
    Streaming / Conserved with the reference's semantics (f holds the
    post-streaming state S_n after n steps)."""

    def __init__(self, nx, ny, nz, tau0, rho0, u0, nranks=1):
        self.shape = (int(nx), int(ny), int(nz))
        self.tau0, self.rho0, self.u0 = float(tau0), float(rho0), float(u0)
        self.nranks = int(nranks)
        self._h = lib().oracle_create(self.shape[0], self.shape[1], self.shape[2],
                                      self.tau0, self.rho0, self.u0, self.nranks)
        if not self._h:
            raise MemoryError("oracle_create failed")

    def close(self):
        if self._h:
            lib().oracle_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def initialize(self):
        lib().oracle_initialize(self._h)

    def collision(self):
        lib().oracle_collision(self._h)

    def streaming(self):
        lib().oracle_streaming(self._h)

    def step(self, n=1):
        lib().oracle_step(self._h, int(n))

    def conserved(self):
        out = np.zeros(5, np.float64)
        lib().oracle_conserved(self._h, _dp(out))
        return out

    def conserved_compensated(self):
        out = np.zeros(5, np.float64)
        lib().oracle_conserved_compensated(self._h, _dp(out))
        return out

    def get_f(self):
        """S_n as [27][nx][ny][nz] (the reference's (ii, i, j, k) order)."""
        out = np.empty((Q,) + self.shape, np.float64)
        lib().oracle_get_f(self._h, _dp(out))
        return out

    def set_f(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        assert f.shape == (Q,) + self.shape
        lib().oracle_set_f(self._h, _dp(f))

    def grid(self):
        out = np.zeros(3, np.int32)
        lib().oracle_grid(self._h, _ip(out))
        return tuple(int(v) for v in out)

    def rank_geometry(self, me):
        out = np.zeros(9, np.int32)
        lib().oracle_rank_geometry(self._h, int(me), _ip(out))
        return {"p": tuple(out[0:3]), "l": tuple(out[3:6]), "lo": tuple(out[6:9])}
