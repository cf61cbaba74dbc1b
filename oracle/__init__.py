"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may
import this module.  It exposes two libraries with the same array conventions
(the reference's AoS ``f[j][i][k]`` and row-major ``a[j][i]``):

* :data:`ORACLE` -- ``liblbm_oracle.so``, this project's C restatement of the
  lbc step (``oracle/lbm_oracle.c``).
* :func:`reference` -- ``_ref/liblbm_ref.so``, the real lbc library compiled from
  the reference's own sources by ``oracle/Makefile``, or ``None`` when it was
  not built.  It is called through its native C API (``main_lbm`` with a
  ``Boundary`` table of its own function pointers, lbm_core.h:6-31).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liblbm_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "liblbm_ref.so")

# D2Q9 tables of the reference drivers (cases/nit/nit.c:195-202).
CX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1], dtype=np.int32)
CY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1], dtype=np.int32)
OPP = np.array([0, 3, 4, 1, 2, 7, 8, 5, 6], dtype=np.int32)
W = np.array([4.0 / 9.0] + [1.0 / 9.0] * 4 + [1.0 / 36.0] * 4, dtype=np.float64)

# Boundary-condition codes (lbm_oracle.h; same numbering as include/lbmx.h).
BC_CODES = {
    "convective_inlet": 1, "convective_outlet": 2, "pressure_inlet": 3,
    "pressure_outlet": 4, "periodic_x": 5, "periodic_y": 6, "velocity_inlet": 7,
    "velocity_outlet": 8, "velocity_top": 9, "velocity_bottom": 10,
    "velocity_outlet_regularized": 11, "velocity_inlet_regularized": 12,
    "velocity_top_regularized": 13, "velocity_bottom_regularized": 14,
    "pressure_outlet_regularized": 15, "pressure_inlet_regularized": 16,
    "pressure_top_regularized": 17, "pressure_bottom_regularized": 18,
    "convective_top": 19, "convective_bottom": 20,
}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _p(a):
    """Pointer to a contiguous numpy array (float64 -> double*, int32 -> int*)."""
    if a is None:
        return None
    assert a.flags.c_contiguous
    if a.dtype == np.float64:
        return a.ctypes.data_as(_dp)
    if a.dtype == np.int32:
        return a.ctypes.data_as(_ip)
    raise TypeError(a.dtype)


class OrcBC(C.Structure):
    _fields_ = [("type", C.c_int), ("index", C.c_int), ("val1", _dp), ("val2", _dp)]


class OrcTdibc(C.Structure):
    _fields_ = [(n, C.c_double * 2 if n != "Zinf" else C.c_double) for n in
                ("Zinf", "A", "lambdas", "B", "C", "alpha", "beta", "phi", "psi", "chi")]


class OrcNitCfg(C.Structure):
    _fields_ = [("Nx", C.c_int), ("Ny", C.c_int), ("nt", C.c_int), ("init_iter_max", C.c_int)] + \
               [(n, C.c_double) for n in ("dx", "dt", "cu", "cp", "cs", "rho0", "tau", "Ly",
                                          "y1", "y2", "d", "tol_rho")] + \
              [("n_orifices", C.c_int), ("variant", C.c_int)]


class OrcLinerCfg(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("Nx", "Ny", "liner_x0", "liner_x1", "sponge", "probe_x",
                                       "probe_y")] + \
               [(n, C.c_double) for n in ("tau", "tau_sponge", "rho0", "amp", "x0",
                                          "halfwidth", "gain")] + [("imp", OrcTdibc)]


def _load_oracle():
    if not os.path.exists(ORACLE_SO):
        raise RuntimeError(f"{ORACLE_SO} is not built (run `make -C oracle`)")
    lib = C.CDLL(ORACLE_SO)
    i, d = C.c_int, C.c_double
    lib.orc_step.argtypes = [i, i, _dp, _dp, _dp, _dp, _dp, _dp, _ip, C.POINTER(OrcBC), i,
                             i, i, d, d, _dp, _dp, _dp, i]
    lib.orc_run.argtypes = [i, i, i, _dp, _dp, _dp, _dp, _dp, _dp, _ip, C.POINTER(OrcBC), i,
                            i, i, d, d, _dp, _dp, _dp, i, _ip]
    lib.orc_mei.argtypes = [i, i, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, i, i, i, d,
                            _ip, _dp]
    lib.orc_mei.restype = i
    lib.orc_collision.argtypes = [i, i, _dp, _dp, _dp, _dp, _dp, _ip, d, d]
    lib.orc_streaming.argtypes = [i, i, _dp, _dp, _ip, i, i]
    lib.orc_bounce_back.argtypes = [i, i, _dp, _dp, _ip, i, i]
    lib.orc_interpolated_bounce_back.argtypes = [i, i, _dp, _dp, _ip, _dp, i, i]
    lib.orc_force_mem.argtypes = [i, i, _dp, _dp, _ip, i, i, _dp, _dp]
    lib.orc_moments.argtypes = [i, i, _dp, _ip, _dp, _dp, _dp, d, d]
    lib.orc_apply_bc.argtypes = [i, i, C.POINTER(OrcBC), _dp, _dp, _dp, _dp, _dp, _ip]
    lib.orc_sponge_layer.argtypes = [i, i] + [d] * 8 + [_dp]
    lib.orc_tdibc_velocity.argtypes = [C.POINTER(OrcTdibc), d, d]
    lib.orc_tdibc_velocity.restype = d
    lib.orc_nit_config.argtypes = [d, d, C.POINTER(OrcNitCfg)]
    lib.orc_nit_run.argtypes = [C.POINTER(OrcNitCfg), d, d, C.POINTER(OrcTdibc), i, _dp,
                                _dp, _dp]
    lib.orc_nit_run.restype = i
    lib.orc_liner_setup.argtypes = [C.POINTER(OrcLinerCfg), _dp, _dp, _dp, _dp, _dp, _ip]
    lib.orc_liner_run.argtypes = [C.POINTER(OrcLinerCfg), i, _dp, _dp, _dp, _dp]
    return lib


ORACLE = _load_oracle()


def make_bcs(bcs):
    """bcs: list of (name, index, val1 ndarray|None, val2 ndarray|None)."""
    arr = (OrcBC * max(1, len(bcs)))()
    for n, (name, index, v1, v2) in enumerate(bcs):
        arr[n].type = BC_CODES[name]
        arr[n].index = int(index)
        arr[n].val1 = _p(v1)
        arr[n].val2 = _p(v2)
    return arr


class Lattice:
    """Host state of one lattice in the reference layout (numpy-owned)."""

    def __init__(self, Nx, Ny):
        self.Nx, self.Ny = Nx, Ny
        self.f = np.zeros((Ny, Nx, 9))
        self.f_new = np.zeros((Ny, Nx, 9))
        self.rho = np.ones((Ny, Nx))
        self.u = np.zeros((Ny, Nx))
        self.v = np.zeros((Ny, Nx))
        self.omega = np.ones((Ny, Nx))
        self.mask = np.zeros((Ny, Nx), dtype=np.int32)
        self.phi = np.zeros((Ny, Nx))

    def copy(self):
        o = Lattice(self.Nx, self.Ny)
        for k in ("f", "f_new", "rho", "u", "v", "omega", "mask", "phi"):
            setattr(o, k, getattr(self, k).copy())
        return o


def oracle_step(lat, bcs, per_x, per_y, Fx=0.0, Fy=0.0, use_ibb=0):
    """One orc_step then the driver swap; returns (L, D)."""
    L, D = C.c_double(0.0), C.c_double(0.0)
    tbl = make_bcs(bcs)
    ORACLE.orc_step(lat.Nx, lat.Ny, _p(lat.f), _p(lat.f_new), _p(lat.rho), _p(lat.u),
                    _p(lat.v), _p(lat.omega), _p(lat.mask), tbl, len(bcs), per_x, per_y,
                    Fx, Fy, C.byref(L), C.byref(D), _p(lat.phi), use_ibb)
    lat.f, lat.f_new = lat.f_new, lat.f
    return L.value, D.value


def oracle_run(lat, nsteps, bcs, per_x, per_y, Fx=0.0, Fy=0.0, use_ibb=0):
    """nsteps of orc_step with swaps; returns per-step (L, D) arrays."""
    L = np.zeros(max(1, nsteps))
    D = np.zeros(max(1, nsteps))
    sw = C.c_int(0)
    tbl = make_bcs(bcs)
    ORACLE.orc_run(lat.Nx, lat.Ny, nsteps, _p(lat.f), _p(lat.f_new), _p(lat.rho), _p(lat.u),
                   _p(lat.v), _p(lat.omega), _p(lat.mask), tbl, len(bcs), per_x, per_y, Fx, Fy,
                   _p(L), _p(D), _p(lat.phi), use_ibb, C.byref(sw))
    if sw.value:
        lat.f, lat.f_new = lat.f_new, lat.f
    return L[:nsteps], D[:nsteps]


def oracle_mei(lat, rho_in, U_in, V_in, per_x, per_y, iter_max, tol):
    md = C.c_double(0.0)
    n = ORACLE.orc_mei(lat.Nx, lat.Ny, _p(lat.f), _p(lat.f_new), _p(rho_in), _p(U_in),
                       _p(V_in), _p(lat.rho), _p(lat.u), _p(lat.v), _p(lat.omega), per_x,
                       per_y, iter_max, tol, _p(lat.mask), C.byref(md))
    return n, md.value


# ----------------------------------------------------------------------------
# The real reference library (built from /root/reference by oracle/Makefile).
# ----------------------------------------------------------------------------
class RefBoundary(C.Structure):
    """Boundary of include/lbm_types.h:22-27."""
    _fields_ = [("index", C.c_int), ("val1", _dp), ("val2", _dp), ("apply", C.c_void_p)]


class Reference:
    def __init__(self, path=REF_SO):
        self.lib = C.CDLL(path)
        i, d = C.c_int, C.c_double
        self.lib.main_lbm.argtypes = [i, i, i, _dp, _dp, _dp, _dp, _dp, _ip, _ip, _dp, _ip,
                                      _dp, _ip, C.POINTER(RefBoundary), i, i, i, d, d, _dp,
                                      _dp, _dp, i]
        self.lib.mei_initialization.argtypes = [i, i, i, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                                _dp, _ip, _ip, _dp, _dp, i, i, i, d, _ip]
        self.lib.sponge_layer.argtypes = [i, i] + [d] * 8 + [_dp]
        self.cx, self.cy, self.opp, self.w = CX.copy(), CY.copy(), OPP.copy(), W.copy()

    def fn(self, name):
        return C.cast(getattr(self.lib, name), C.c_void_p).value

    def table(self, bcs):
        arr = (RefBoundary * max(1, len(bcs)))()
        for n, (name, index, v1, v2) in enumerate(bcs):
            arr[n].index = int(index)
            arr[n].val1 = _p(v1)
            arr[n].val2 = _p(v2)
            arr[n].apply = self.fn(name)
        return arr

    def step(self, lat, bcs, per_x, per_y, Fx=0.0, Fy=0.0, use_ibb=0):
        L, D = C.c_double(0.0), C.c_double(0.0)
        tbl = self.table(bcs)
        self.lib.main_lbm(lat.Nx, lat.Ny, 9, _p(lat.f), _p(lat.f_new), _p(lat.rho),
                          _p(lat.u), _p(lat.v), _p(self.cx), _p(self.cy), _p(self.w),
                          _p(self.opp), _p(lat.omega), _p(lat.mask), tbl, len(bcs), per_x,
                          per_y, Fx, Fy, C.byref(L), C.byref(D), _p(lat.phi), use_ibb)
        lat.f, lat.f_new = lat.f_new, lat.f
        return L.value, D.value

    def mei(self, lat, rho_in, U_in, V_in, per_x, per_y, iter_max, tol):
        self.lib.mei_initialization(lat.Nx, lat.Ny, 9, _p(lat.f), _p(lat.f_new), _p(rho_in),
                                    _p(U_in), _p(V_in), _p(lat.rho), _p(lat.u), _p(lat.v),
                                    _p(self.cx), _p(self.cy), _p(self.w), _p(lat.omega),
                                    per_x, per_y, iter_max, tol, _p(lat.mask))


def reference():
    """The compiled reference library, or None when oracle/_ref was not built."""
    if not os.path.exists(REF_SO):
        return None
    return Reference()
