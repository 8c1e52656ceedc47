"""imex_conserved_noalloc against imex_conserved, through the C ABI.

The allocation-free entry point sums {mass, px, py, pz, ke} over populations
the caller holds in device memory, with its partial sums in the library's
fixed scratch array.  For every storage format it must give the same bits as
imex_conserved on a default single-domain context holding the same state, on
two calls in a row, the second with a smaller domain (the scratch keeps the
first call's partials), and with a padded direction stride.  The device
buffers here come from the HIP runtime the library itself links.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D = 1   # hipMemcpyHostToDevice


def _hip():
    from imexlbm_amd import capi
    capi.load()                                   # the library's own HIP runtime
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipDeviceSynchronize.argtypes = []
    return h


def _entry():
    from imexlbm_amd import capi
    L = capi.load()
    fn = L.imex_conserved_noalloc
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_longlong,
                   ctypes.c_double, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
    fn.restype = ctypes.c_int
    return fn


def _bf16_bits(x32):
    u = x32.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def _state(prec, nx, ny, nz, rho0, seed, dstride):
    """Stored populations in the library's layout f[z][ii][y][x] (direction
    planes `dstride` apart) and the same state as doubles f[z][ii][y][x]."""
    from oracle import oracle
    w, _, _ = oracle.tables()
    rng = np.random.default_rng(seed)
    shape = (nz, 27, ny, nx)
    wb = w[None, :, None, None]
    if prec == "fp64":
        f = wb * rho0 * (1.0 + 0.1 * rng.standard_normal(shape))
        stored = f
    else:
        g32 = (wb * 0.1 * rng.standard_normal(shape)).astype(np.float32)
        if prec == "fp32":
            stored, gd = g32, g32.astype(np.float64)
        elif prec == "fp16":
            stored = g32.astype(np.float16)
            gd = stored.astype(np.float64)
        else:
            stored = _bf16_bits(g32)
            gd = (stored.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
        f = gd + wb * rho0
    if dstride != nx * ny:
        pad = np.zeros((nz, 27, dstride), dtype=stored.dtype)
        pad[:, :, :nx * ny] = stored.reshape(nz, 27, nx * ny)
        stored = pad
    return np.ascontiguousarray(stored), f


def _reference(prec, nx, ny, nz, rho0, f):
    from imexlbm_amd import capi
    d = {"nx": nx, "ny": ny, "nz": nz, "tau0": 0.01, "rho0": rho0, "u0": 0.05}
    with capi.Context(d, precision=prec, block=256, cells_per_thread=2) as c:
        c.initialize()
        c.set_f(f.transpose(1, 3, 2, 0))          # -> (ii, i, j, k)
        return c.conserved()


def _noalloc(hip, fn, prec, stored, nx, ny, nz, dstride, rho0):
    from imexlbm_amd import capi
    ptr = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(ptr), stored.nbytes) == 0
    try:
        assert hip.hipMemcpy(ptr, stored.ctypes.data, stored.nbytes, H2D) == 0
        out = np.full(5, np.nan)
        st = fn(ptr, capi.PRECISIONS[prec], nx, ny, nz, dstride, rho0, None,
                out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        assert st == 0, capi.load().imex_last_error().decode()
        return out
    finally:
        hip.hipDeviceSynchronize()
        hip.hipFree(ptr)


@pytest.mark.parametrize("prec", ["fp64", "fp32", "fp16", "bf16"])
def test_noalloc_conserved_equals_context_conserved(prec):
    hip, fn = _hip(), _entry()
    rho0 = 1.0
    # two calls in a row, the second smaller; then a padded direction stride
    for k, (nx, ny, nz, pad) in enumerate(((40, 24, 10, 0), (17, 12, 5, 0), (33, 7, 6, 64))):
        print(f"case: noalloc {prec} {nx}x{ny}x{nz} pad {pad}", flush=True)
        dstride = nx * ny + pad
        stored, f = _state(prec, nx, ny, nz, rho0, 100 + k, dstride)
        got = _noalloc(hip, fn, prec, stored, nx, ny, nz, dstride if pad else 0, rho0)
        want = _reference(prec, nx, ny, nz, rho0, f)
        assert np.array_equal(got, want), (got, want)
        # and the sums themselves, against a plain FP64 sum of the state
        from oracle import oracle
        _, e, _ = oracle.tables()
        pl = f.sum(axis=1)
        mom = np.einsum("ia,zixy->azxy", e.astype(float), f)
        assert abs(got[0] - pl.sum()) <= 1e-12 * abs(pl.sum())
        assert np.abs(got[1:4] - mom.reshape(3, -1).sum(axis=1)).max() <= 1e-12 * abs(pl.sum())
        assert abs(got[4] - 0.5 * (mom ** 2).sum()) <= 1e-12 * abs(got[4])


def test_noalloc_rejects_planes_beyond_the_kernels_addressing():
    from imexlbm_amd import capi
    fn = _entry()
    out = np.zeros(5)
    # 27 direction planes of 10^8 FP64 values exceed the 2 GiB plane window;
    # the host check refuses before any device access
    st = fn(ctypes.c_void_p(16), capi.F64, 10000, 10000, 1, 0, 1.0, None,
            out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert st == -6   # IMEX_ERANGE
    assert "32-bit addressing" in capi.load().imex_last_error().decode()
    st = fn(ctypes.c_void_p(16), capi.F64, 8, 8, 8, 10, 1.0, None,
            out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert st == -1   # IMEX_EINVAL: direction stride below nx*ny
