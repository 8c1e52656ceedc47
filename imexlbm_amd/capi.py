"""ctypes binding of the C ABI declared in include/imexlbm.h.

The shared library is the product: every call here goes straight to
libimexlbm.so, and loading fails loudly when it has not been built.  There is
no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libimexlbm.so")

F64, F32, F16, BF16 = 0, 1, 2, 3
WRAP, TILED = 0, 1
PRECISIONS = {"fp64": F64, "fp32": F32, "fp16": F16, "bf16": BF16}
STORAGE_BYTES = {F64: 8, F32: 4, F16: 2, BF16: 2}

EXPORTS = (
    "imex_config_default", "imex_create", "imex_initialize", "imex_collision",
    "imex_streaming", "imex_step", "imex_conserved", "imex_get_f", "imex_set_f",
    "imex_synchronize", "imex_timed_steps", "imex_autotune", "imex_query",
    "imex_destroy", "imex_device_count", "imex_strerror", "imex_last_error",
    "imex_copy_probe", "imex_get_macro", "imex_profile_phase_name", "imex_profile_reset",
    "imex_profile_read", "imex_conserved_noalloc",
)


class Domain(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("tau0", ctypes.c_double), ("rho0", ctypes.c_double), ("u0", ctypes.c_double)]


class Config(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "precision", "mode", "ntiles", "halo_depth", "block", "cells_per_thread",
        "persistent", "nontemporal", "dir_pad", "use_graph", "sequential", "device",
        "xcd_groups", "contiguous", "io16", "profile", "xcd_run", "two_step", "two_step_tile",
        "two_step_zseg", "two_step_xcds", "two_step_prefetch", "two_step_io", "halo_direct", "whole_x")]


class Info(ctypes.Structure):
    _fields_ = [("cells", ctypes.c_longlong), ("bytes_per_update", ctypes.c_longlong),
                ("device_bytes", ctypes.c_longlong)] + [(n, ctypes.c_int) for n in (
                    "storage_bytes", "ndomains", "ntiles", "work_tiles", "block",
                    "cells_per_thread", "persistent", "nontemporal", "dir_pad", "use_graph",
                    "grid", "rx", "ry", "rz", "xcd_groups", "contiguous", "io16", "xcd_run",
                    "two_step", "two_step_tile", "two_step_zseg", "two_step_xcds", "two_step_prefetch",
                    "two_step_io", "halo_direct", "halo_depth", "whole_x")] + [
                        ("halo_send_bytes", ctypes.c_longlong)]


class ImexError(RuntimeError):
    def __init__(self, status, detail):
        super().__init__(f"imexlbm error {status}: {detail}")
        self.status = status


_lib = None


def load():
    """Load libimexlbm.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = ctypes.CDLL(LIB_PATH)
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        dp = ctypes.POINTER(ctypes.c_double)
        L.imex_config_default.argtypes = [ctypes.POINTER(Config)]
        L.imex_config_default.restype = None
        L.imex_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(Config), ctypes.POINTER(Domain), i]
        for n in ("imex_initialize", "imex_collision", "imex_streaming", "imex_synchronize"):
            getattr(L, n).argtypes = [vp]
        L.imex_step.argtypes = [vp, i]
        L.imex_conserved.argtypes = [vp, i, dp]
        L.imex_get_f.argtypes = [vp, i, dp]
        L.imex_set_f.argtypes = [vp, i, dp]
        L.imex_timed_steps.argtypes = [vp, i, dp, dp]
        L.imex_autotune.argtypes = [vp, i, dp]
        L.imex_query.argtypes = [vp, ctypes.POINTER(Info)]
        L.imex_destroy.argtypes = [vp]
        L.imex_destroy.restype = None
        L.imex_device_count.restype = i
        L.imex_copy_probe.argtypes = [i, ctypes.c_longlong, i, dp]
        L.imex_get_macro.argtypes = [vp, i, dp, dp, dp, dp]
        L.imex_profile_phase_name.argtypes = [i]
        L.imex_profile_phase_name.restype = ctypes.c_char_p
        L.imex_profile_reset.argtypes = [vp]
        L.imex_profile_read.argtypes = [vp, i, dp, ctypes.POINTER(ctypes.c_longlong)]
        L.imex_conserved_noalloc.argtypes = [vp, i, i, i, i, ctypes.c_longlong, d, vp, dp]
        L.imex_conserved_noalloc.restype = i
        L.imex_strerror.restype = ctypes.c_char_p
        L.imex_strerror.argtypes = [i]
        L.imex_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def device_count() -> int:
    return int(load().imex_device_count())


def copy_probe(nbytes=8 << 30, reps=10, device=0) -> float:
    """Read+write GB/s of a plain device copy of nbytes."""
    out = np.zeros(1, np.float64)
    _check(load().imex_copy_probe(int(device), int(nbytes), int(reps), _dptr(out)))
    return float(out[0])


def _check(status):
    if status != 0:
        L = load()
        raise ImexError(status, f"{L.imex_strerror(status).decode()}: {L.imex_last_error().decode()}")


def default_config() -> Config:
    c = Config()
    load().imex_config_default(ctypes.byref(c))
    return c


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Context:
    """One imex_ctx: a single domain, a tiled domain or a batch of domains."""

    def __init__(self, domains, precision=F64, mode=WRAP, **cfg):
        if isinstance(precision, str):
            precision = PRECISIONS[precision]
        L = load()
        c = default_config()
        c.precision = int(precision)
        c.mode = int(mode)
        for k, v in cfg.items():
            if not hasattr(c, k):
                raise TypeError(f"unknown config field {k}")
            setattr(c, k, int(v))
        if isinstance(domains, dict):
            domains = [domains]
        arr = (Domain * len(domains))()
        for a, dm in zip(arr, domains):
            a.nx, a.ny, a.nz = int(dm["nx"]), int(dm["ny"]), int(dm["nz"])
            a.tau0, a.rho0, a.u0 = float(dm["tau0"]), float(dm["rho0"]), float(dm["u0"])
        self.domains = [dict(d) for d in domains]
        self.precision = int(precision)
        self._h = ctypes.c_void_p()
        _check(L.imex_create(ctypes.byref(self._h), ctypes.byref(c), arr, len(domains)))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            load().imex_destroy(self._h)
            self._h = ctypes.c_void_p()

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
        _check(load().imex_initialize(self._h))

    def collision(self):
        _check(load().imex_collision(self._h))

    def streaming(self):
        _check(load().imex_streaming(self._h))

    def step(self, n=1):
        _check(load().imex_step(self._h, int(n)))

    def synchronize(self):
        _check(load().imex_synchronize(self._h))

    def conserved(self, domain=0):
        out = np.zeros(5, np.float64)
        _check(load().imex_conserved(self._h, int(domain), _dptr(out)))
        return out

    def get_f(self, domain=0):
        d = self.domains[domain]
        out = np.empty((27, d["nx"], d["ny"], d["nz"]), np.float64)
        _check(load().imex_get_f(self._h, int(domain), _dptr(out)))
        return out

    def get_macro(self, domain=0):
        """ComputeMacroscopic of S_n: (p, u, v, w), each [nx][ny][nz]."""
        d = self.domains[domain]
        shape = (d["nx"], d["ny"], d["nz"])
        out = [np.empty(shape, np.float64) for _ in range(4)]
        _check(load().imex_get_macro(self._h, int(domain), *(_dptr(a) for a in out)))
        return tuple(out)

    def set_f(self, f, domain=0):
        d = self.domains[domain]
        f = np.ascontiguousarray(f, dtype=np.float64)
        if f.shape != (27, d["nx"], d["ny"], d["nz"]):
            raise ValueError("population array has the wrong shape")
        _check(load().imex_set_f(self._h, int(domain), _dptr(f)))

    def timed_steps(self, n):
        w = np.zeros(1, np.float64)
        k = np.zeros(1, np.float64)
        _check(load().imex_timed_steps(self._h, int(n), _dptr(w), _dptr(k)))
        return float(w[0]), float(k[0])

    def autotune(self, reps=3):
        b = np.zeros(1, np.float64)
        _check(load().imex_autotune(self._h, int(reps), _dptr(b)))
        return float(b[0])

    def profile(self):
        """{group name: (calls, total ms, fastest ms, slowest ms, bytes, flops)}"""
        L = load()
        out = {}
        k = 0
        while True:
            name = L.imex_profile_phase_name(k)
            if name is None:
                return out
            v = np.zeros(5, np.float64)
            n = ctypes.c_longlong()
            _check(L.imex_profile_read(self._h, k, _dptr(v), ctypes.byref(n)))
            out[name.decode()] = (n.value, *map(float, v))
            k += 1

    def profile_reset(self):
        _check(load().imex_profile_reset(self._h))

    def info(self):
        i = Info()
        _check(load().imex_query(self._h, ctypes.byref(i)))
        return {n: getattr(i, n) for n, _ in Info._fields_}
