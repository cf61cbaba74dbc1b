"""ctypes binding of the reference's C API as ``liblbm_hip.so`` exports it.

This is the same ABI the reference's C drivers link against (include/lbm_core.h,
lbm_bc.h, lbm_types.h): ``main_lbm`` with a ``Boundary`` table whose ``apply``
members are this library's own BC functions.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

CX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1], dtype=np.int32)
CY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1], dtype=np.int32)
OPP = np.array([0, 3, 4, 1, 2, 7, 8, 5, 6], dtype=np.int32)
W = np.array([4.0 / 9.0] + [1.0 / 9.0] * 4 + [1.0 / 36.0] * 4, dtype=np.float64)


class Boundary(C.Structure):
    """include/lbm_types.h:22-27."""
    _fields_ = [("index", C.c_int), ("val1", _dp), ("val2", _dp), ("apply", C.c_void_p)]


def _d(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _i(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


class DropIn:
    """The drop-in library through its reference-compatible entry points."""

    def __init__(self):
        L = lib()
        i, d = C.c_int, C.c_double
        L.main_lbm.argtypes = [i, i, i, _dp, _dp, _dp, _dp, _dp, _ip, _ip, _dp, _ip, _dp, _ip,
                               C.POINTER(Boundary), i, i, i, d, d, _dp, _dp, _dp, i]
        L.main_lbm.restype = None
        L.mei_initialization.argtypes = [i, i, i, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _ip,
                                         _ip, _dp, _dp, i, i, i, d, _ip]
        L.mei_initialization.restype = None
        L.collision.argtypes = [i, i, i, _dp, _dp, _dp, _dp, _dp, _ip, _ip, _ip, _dp, d, d]
        L.streaming.argtypes = [i, i, i, _dp, _dp, _ip, _ip, _ip, i, i]
        L.bounce_back.argtypes = [i, i, i, _dp, _dp, _ip, _ip, _ip, _ip, i, i]
        L.interpolated_bounce_back.argtypes = [i, i, i, _dp, _dp, _ip, _dp, _ip, _ip, _ip, i, i]
        L.compute_macroscopic_fields.argtypes = [i, i, i, _dp, _ip, _ip, _ip, _dp, _dp, _dp,
                                                 d, d]
        L.compute_force_mem.argtypes = [i, i, i, _dp, _dp, _ip, i, i, _ip, _ip, _ip, _dp, _dp]
        L.sponge_layer.argtypes = [i, i] + [d] * 8 + [_dp]
        L.lbm_hip_sync.restype = None
        L.lbm_hip_release.restype = None
        self.lib = L
        self.cx, self.cy, self.opp, self.w = CX.copy(), CY.copy(), OPP.copy(), W.copy()

    def fn(self, name):
        return C.cast(getattr(self.lib, name), C.c_void_p).value

    def table(self, bcs):
        """bcs: list of (name, index, val1|None, val2|None) -> Boundary array."""
        arr = (Boundary * max(1, len(bcs)))()
        for n, (name, index, v1, v2) in enumerate(bcs):
            arr[n].index = int(index)
            arr[n].val1 = _d(v1)
            arr[n].val2 = _d(v2)
            arr[n].apply = self.fn(name)
        return arr

    def main_lbm(self, f, f_new, rho, u, v, omega, mask, bcs, per_x, per_y, Fx, Fy, phi,
                 use_ibb, table=None):
        Ny, Nx = rho.shape
        Lf, Df = C.c_double(0.0), C.c_double(0.0)
        tbl = table if table is not None else self.table(bcs)
        self.lib.main_lbm(Nx, Ny, 9, _d(f), _d(f_new), _d(rho), _d(u), _d(v), _i(self.cx),
                          _i(self.cy), _d(self.w), _i(self.opp), _d(omega), _i(mask), tbl,
                          len(bcs), per_x, per_y, Fx, Fy, C.byref(Lf), C.byref(Df), _d(phi),
                          use_ibb)
        return Lf.value, Df.value

    def mei_initialization(self, f, f_new, rho_in, U_in, V_in, rho, u, v, omega, per_x, per_y,
                           iters, tol, mask):
        Ny, Nx = rho.shape
        self.lib.mei_initialization(Nx, Ny, 9, _d(f), _d(f_new), _d(rho_in), _d(U_in),
                                    _d(V_in), _d(rho), _d(u), _d(v), _i(self.cx), _i(self.cy),
                                    _d(self.w), _d(omega), per_x, per_y, iters, tol, _i(mask))

    def sync(self):
        self.lib.lbm_hip_sync()

    def release(self):
        self.lib.lbm_hip_release()
