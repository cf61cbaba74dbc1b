"""GPU check of lbmx_totals_noalloc, the allocation-free variant of lbmx_totals.

Two calls in a row through the C ABI, the second on a smaller lattice, each on
device arrays the test allocates itself (hipMalloc from the HIP runtime the
library already loaded): the totals equal lbmx_totals' on the same field bit for
bit.  The first lattice has more than 1,024 rows (several rows per block of the
first pass) and solid nodes (a mask); the second is all fluid (no mask).  Sizes
outside 1 .. LBMX_TOTALS_NOALLOC_MAX_NODES and null arrays are refused before
any launch.
"""
import ctypes as C
import math

import numpy as np
import pytest

from lbc_amd import cases, lib, native

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime the library is linked to."""

    def __init__(self):
        native._bind()                          # the library loads libamdhip64.so.7 first
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipDeviceSynchronize.argtypes = []
        self.held = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), a.nbytes) == 0
        self.held.append(p)
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def empty(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), nbytes) == 0
        self.held.append(p)
        return p

    def get(self, p, shape, dtype=np.float64):
        out = np.empty(shape, dtype=dtype)
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.held:
            self.h.hipFree(p)
        self.held = []


def noalloc_totals(hip, rho, u, v, mask):
    L = lib()
    L.lbmx_totals_noalloc.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p]
    Ny, Nx = rho.shape
    out = hip.empty(16)
    d_mask = None if mask is None else hip.put(mask.astype(np.int32))
    native.check(L.lbmx_totals_noalloc(Nx, Ny, hip.put(rho), hip.put(u), hip.put(v), d_mask,
                                       out, None))
    assert hip.h.hipDeviceSynchronize() == 0
    return tuple(hip.get(out, (2,)))


def test_noalloc_totals_equal_lbmx_totals():
    hip = Hip()
    try:
        # 1: channel with a cylinder, 300 x 1100 (rows dealt several to a block)
        lat, _ = cases.cylinder_channel(300, 1100, 20.0, 150.0, 550.0, 0.05, 100.0, native.F64)
        lat.step(120)
        want = lat.totals()
        rho, u, v = lat.moments()
        lat.close()
        mask = np.isnan(rho)
        got = noalloc_totals(hip, rho, u, v, mask)
        assert got == want, (got, want)
        assert want[0] > 0.9 * (~mask).sum() and want[1] > 0.0
        # 2: smaller, all fluid (periodic Taylor-Green vortex, FP32 storage), no mask
        lat = cases.taylor_green(96, native.F32_SHIFTED, Ny=80)
        lat.step(33)
        want2 = lat.totals()
        rho2, u2, v2 = lat.moments()
        lat.close()
        got2 = noalloc_totals(hip, rho2, u2, v2, None)
        assert got2 == want2, (got2, want2)
        assert abs(want2[0] - 96 * 80) < 1e-9 * 96 * 80 and want2[1] > 0.0
        assert got2 != got
    finally:
        hip.free()


def test_noalloc_rejects_bad_arguments():
    hip = Hip()
    try:
        L = lib()
        L.lbmx_totals_noalloc.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        a = hip.put(np.ones(64))
        out = hip.empty(16)
        assert L.lbmx_totals_noalloc(0, 8, a, a, a, None, out, None) == -1
        assert L.lbmx_totals_noalloc(8, -1, a, a, a, None, out, None) == -1
        assert L.lbmx_totals_noalloc(1 << 16, 1 << 16, a, a, a, None, out, None) == -1
        assert "nodes" in L.lbmx_last_error().decode()
        assert L.lbmx_totals_noalloc(8, 8, None, a, a, None, out, None) == -1
        assert L.lbmx_totals_noalloc(8, 8, a, a, a, None, None, None) == -1
        # a valid call still works after the refusals
        assert L.lbmx_totals_noalloc(8, 8, a, a, a, None, out, None) == 0
        assert hip.h.hipDeviceSynchronize() == 0
        m, e = hip.get(out, (2,))
        assert m == 64.0 and e == 64.0
    finally:
        hip.free()
    assert math.isfinite(m)
