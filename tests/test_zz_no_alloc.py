"""paradmm_square_quality_noalloc (the scratch-free variant of
paradmm_gpu_square_quality) against the entry point it varies, through the
C ABI, on two calls in a row, the second with a smaller packing.  The
caller (this test) owns the device buffers: it allocates them through the
HIP runtime itself."""
import ctypes as C

import numpy as np
import pytest

from engine_amd import paradmm as P

pytestmark = pytest.mark.gpu


class _Hip:
    """The few HIP runtime calls a caller of the scratch-free entry point makes."""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipDeviceSynchronize.argtypes = []

    def malloc(self, n):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), n) == 0
        return p

    def h2d(self, dst, a):
        assert self.h.hipMemcpy(dst, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0

    def d2h(self, a, src):
        assert self.h.hipMemcpy(a.ctypes.data_as(C.c_void_p), src, a.nbytes, 2) == 0


def _noalloc(lib, hip, zdev, n, outdev):
    rc = lib.paradmm_square_quality_noalloc(zdev, n, outdev, None)
    assert rc == 0, lib.paradmm_last_error()
    assert hip.h.hipDeviceSynchronize() == 0
    out = np.zeros(3)
    hip.d2h(out, outdev)
    return out


def test_noalloc_quality_equals_allocating_entry_point():
    lib = P.load()
    lib.paradmm_square_quality_noalloc.restype = C.c_int
    lib.paradmm_square_quality_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    hip = _Hip()
    nmax = 700
    zdev = hip.malloc(4 * 3 * nmax)
    outdev = hip.malloc(8 * 3)
    try:
        for n, seed, iters in ((700, 3, 40), (150, 8, 25)):
            g = P.packing_square(n, seed, rho=1.0)
            g.iterate(iters)
            want = np.array(g.square_quality())
            z32 = g.read("Z").astype(np.float32)  # the device's fp32 Z, exactly
            hip.h2d(zdev, np.ascontiguousarray(z32))
            got = _noalloc(lib, hip, zdev, n, outdev)
            assert np.array_equal(got, want), (n, got, want)
            assert want[0] > 0.0  # the packing still overlaps: the pair maximum is exercised
            g.close()
        # empty input: zeros; beyond the limit: an error, nothing launched
        assert np.array_equal(_noalloc(lib, hip, zdev, 0, outdev), np.zeros(3))
        rc = lib.paradmm_square_quality_noalloc(zdev, 65537, outdev, None)
        assert rc == 1 and b"scratch limit" in lib.paradmm_last_error()
    finally:
        hip.h.hipFree(zdev)
        hip.h.hipFree(outdev)
