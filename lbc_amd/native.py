"""ctypes view of the native device-resident C ABI (include/lbmx.h).

Arrays follow the reference's conventions: populations ``[Ny, Nx, 9]`` float64
(AoS, k fastest), fields ``[Ny, Nx]`` float64, solid mask ``[Ny, Nx]`` int32.
"""
from __future__ import annotations

import ctypes
import ctypes as C
import os

import numpy as np

from . import lib

F64 = 0
F32_SHIFTED = 1      # FP32 storage of f - w*rho0, FP32 collision
F32_WIDE = 2         # the same storage, collision in FP64
# launch grouping of multi-step runs on periodic all-fluid lattices (lbmx_set_launch)
LAUNCH_AUTO, LAUNCH_SINGLE, LAUNCH_TILES, LAUNCH_STREAM = 0, 1, 2, 3
EXACT_STATE = 1

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


class Tdibc(C.Structure):
    """lbmx_tdibc (include/lbmx.h)."""
    _fields_ = [(n, C.c_int) for n in ("bc", "which", "probe_is_column", "probe_index",
                                       "per_node", "record_node", "capacity")] + \
               [(n, C.c_double) for n in ("rho0", "cs", "cp", "dt", "cu", "gain", "Zinf")] + \
               [(n, C.c_double * 2) for n in ("A", "lambdas", "B", "C", "alpha", "beta")] + \
               [("active", C.POINTER(C.c_int))]


class Config(C.Structure):
    _fields_ = [("Nx", C.c_int), ("Ny", C.c_int), ("precision", C.c_int),
                ("per_x", C.c_int), ("per_y", C.c_int), ("use_ibb", C.c_int),
                ("Fx", C.c_double), ("Fy", C.c_double), ("rho0", C.c_double),
                ("flags", C.c_uint)]


def _d(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(_dp)


def _i(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


_bound = False


def _bind():
    global _bound
    if _bound:
        return lib()
    L = lib()
    i, d, vp = C.c_int, C.c_double, C.c_void_p
    L.lbmx_last_error.restype = C.c_char_p
    L.lbmx_device_name.restype = C.c_char_p
    L.lbmx_create.argtypes = [C.POINTER(Config), C.POINTER(vp)]
    L.lbmx_destroy.argtypes = [vp]
    L.lbmx_destroy.restype = None
    L.lbmx_set_geometry.argtypes = [vp, _ip, _dp, _dp]
    L.lbmx_set_boundaries.argtypes = [vp, _ip, _ip, i]
    L.lbmx_set_bc_values.argtypes = [vp, i, _dp, _dp]
    L.lbmx_upload.argtypes = [vp, _dp, _dp, _dp, _dp, _dp]
    L.lbmx_step.argtypes = [vp, i]
    L.lbmx_init_taylor_green.argtypes = [vp, d, d, d]
    L.lbmx_set_uniform.argtypes = [vp, d]
    L.lbmx_set_force.argtypes = [vp, d, d]
    L.lbmx_time_steps.argtypes = [vp, i, _dp]
    L.lbmx_set_launch.argtypes = [vp, i]
    L.lbmx_launch_depth.argtypes = [vp]
    L.lbmx_launch_plan.argtypes = [i, i, i, i, _ip, _ip, _ip]
    L.lbmx_set_tdibc.argtypes = [vp, C.POINTER(Tdibc)]
    L.lbmx_tdibc_history.argtypes = [vp, _dp, i]
    L.lbmx_set_probes.argtypes = [vp, _ip, _ip, i, i]
    L.lbmx_probe_history.argtypes = [vp, _dp, i]
    L.lbmx_mei_initialization.argtypes = [i, i, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                          i, i, i, d, _ip]
    L.lbmx_forces.argtypes = [vp, _dp, _dp]
    L.lbmx_totals.argtypes = [vp, _dp, _dp]
    L.lbmx_set_force_record.argtypes = [vp, d, d, i]
    L.lbmx_force_record.argtypes = [vp, _dp, i]
    L.lbmx_set_force_body.argtypes = [vp, _ip]
    L.lbmx_body_forces.argtypes = [vp, _dp, _dp]
    L.lbmx_body_force_record.argtypes = [vp, _dp, i]
    L.lbmx_moments.argtypes = [vp, _dp, _dp, _dp]
    L.lbmx_download.argtypes = [vp, _dp, _dp]
    L.lbmx_synchronize.argtypes = [vp]
    L.lbmx_steps_done.argtypes = [vp]
    L.lbmx_steps_done.restype = C.c_int64
    L.lbmx_device_bytes.argtypes = [vp]
    L.lbmx_device_bytes.restype = C.c_int64
    L.lbmx_tail_plan.argtypes = [i, i, i, i, _ip, _ip, _ip, i, _ip, _ip]
    L.lbmx_copy_rate.argtypes = [d, _dp]
    L.lbmx_write_vtk.argtypes = [vp, C.c_char_p, d, d, d]
    L.write_vtk_binary_2D.argtypes = [C.c_char_p, i, i, d, _dp, _dp, _dp, d, d]
    L.write_vtk_binary_2D.restype = None
    _bound = True
    return L


class LbmxError(RuntimeError):
    pass


def check(rc):
    if rc != 0:
        raise LbmxError(f"lbmx error {rc}: {lib().lbmx_last_error().decode()}")
    return rc


def device_name():
    return _bind().lbmx_device_name().decode()


def mei_initialization(f, f_new, rho_in, U_in, V_in, rho, u, v, omega, per_x, per_y, iter_max,
                       tol, mask):
    """Mei's consistent initialization on the device (lbmx_mei_initialization);
    arrays are updated in place as the reference function updates them.
    Returns the number of iterations run."""
    L = _bind()
    Ny, Nx = rho.shape
    n = L.lbmx_mei_initialization(Nx, Ny, _d(f), _d(f_new), _d(rho_in), _d(U_in), _d(V_in),
                                  _d(rho), _d(u), _d(v), _d(omega), int(per_x), int(per_y),
                                  int(iter_max), float(tol), _i(mask))
    if n < 0:
        check(n)
    return n


def write_vtk_host(filename, dx, u, v, rho, cu=1.0, crho=1.0):
    """The drop-in host writer write_vtk_binary_2D (lbm_utils.h) on numpy arrays."""
    Ny, Nx = rho.shape
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (u, v, rho)]
    _bind().write_vtk_binary_2D(os.fsencode(filename), Nx, Ny, float(dx), *(_d(a) for a in arrs),
                                float(cu), float(crho))


def copy_rate(gib=8.0):
    """GB/s (read + written) of a plain device copy of `gib` GiB (lbmx_copy_rate)."""
    out = C.c_double()
    check(_bind().lbmx_copy_rate(float(gib), C.byref(out)))
    return out.value


def launch_plan(Nx, Ny, precision=F64, mode=LAUNCH_AUTO):
    """(steps per launch, strip columns, rows per workgroup) of a long multi-step
    run of a periodic all-fluid lattice (lbmx_launch_plan; host only)."""
    L = _bind()
    d, c, r = C.c_int(), C.c_int(), C.c_int()
    check(L.lbmx_launch_plan(int(Nx), int(Ny), int(precision), int(mode), C.byref(d),
                             C.byref(c), C.byref(r)))
    return d.value, c.value, r.value


def tail_plan(Nx, Ny, per_x, per_y, mask, bcs):
    """How lbmx_step groups the boundary table and the stale-edge bookkeeping into
    launches (lbmx_tail_plan; host only).  bcs: [(name, index), ...].  Returns
    (nlaunch, launch_of, edge_launch): launch_of[b] is the launch of entry b (-1:
    left out), edge_launch that of the edge job (-1: none)."""
    L = _bind()
    t = np.array([BC_CODES[b[0]] for b in bcs] or [0], dtype=np.int32)
    ix = np.array([int(b[1]) for b in bcs] or [0], dtype=np.int32)
    out = np.zeros(max(1, len(bcs)), dtype=np.int32)
    edge = np.zeros(1, dtype=np.int32)
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.int32)
    n = L.lbmx_tail_plan(int(Nx), int(Ny), int(per_x), int(per_y), _i(m), _i(t), _i(ix),
                         len(bcs), _i(out), _i(edge))
    if n < 0:
        check(n)
    return n, out[:len(bcs)].tolist(), int(edge[0])


class Lattice:
    """One lattice resident on the GPU (lbmx_lattice)."""

    def __init__(self, Nx, Ny, precision=F64, per_x=0, per_y=0, use_ibb=0, Fx=0.0, Fy=0.0,
                 rho0=1.0, exact=False):
        self._L = _bind()
        self.Nx, self.Ny = int(Nx), int(Ny)
        cfg = Config(self.Nx, self.Ny, precision, int(per_x), int(per_y), int(use_ibb),
                     float(Fx), float(Fy), float(rho0), EXACT_STATE if exact else 0)
        h = C.c_void_p()
        check(self._L.lbmx_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.bc_names = []

    def close(self):
        if self.h:
            self._L.lbmx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_geometry(self, mask, omega, phi=None):
        self._keep = (np.ascontiguousarray(mask, dtype=np.int32),
                      np.ascontiguousarray(omega, dtype=np.float64),
                      None if phi is None else np.ascontiguousarray(phi, dtype=np.float64))
        check(self._L.lbmx_set_geometry(self.h, _i(self._keep[0]), _d(self._keep[1]),
                                        _d(self._keep[2])))

    def set_boundaries(self, bcs):
        """bcs: list of (name, index)."""
        self.bc_names = [b[0] for b in bcs]
        t = np.array([BC_CODES[b[0]] for b in bcs] or [0], dtype=np.int32)
        ix = np.array([int(b[1]) for b in bcs] or [0], dtype=np.int32)
        check(self._L.lbmx_set_boundaries(self.h, _i(t), _i(ix), len(bcs)))

    def set_bc_values(self, b, val1=None, val2=None):
        v1 = None if val1 is None else np.ascontiguousarray(val1, dtype=np.float64)
        v2 = None if val2 is None else np.ascontiguousarray(val2, dtype=np.float64)
        check(self._L.lbmx_set_bc_values(self.h, int(b), _d(v1), _d(v2)))

    def upload(self, f, f_new, rho, u, v):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (f, f_new, rho, u, v)]
        check(self._L.lbmx_upload(self.h, *[_d(a) for a in arrs]))

    def set_uniform(self, omega):
        check(self._L.lbmx_set_uniform(self.h, float(omega)))

    def set_force(self, Fx, Fy):
        check(self._L.lbmx_set_force(self.h, float(Fx), float(Fy)))

    def set_launch(self, mode):
        """Launch grouping of multi-step runs (LAUNCH_AUTO, _SINGLE, _TILES, _STREAM)."""
        check(self._L.lbmx_set_launch(self.h, int(mode)))

    def launch_depth(self):
        """Steps per launch of a long run of this lattice under its launch mode."""
        r = self._L.lbmx_launch_depth(self.h)
        check(min(r, 0))
        return r

    def time_steps(self, n):
        """Device milliseconds (HIP events on the lattice's stream) for n steps."""
        ms = C.c_double()
        check(self._L.lbmx_time_steps(self.h, int(n), C.byref(ms)))
        return ms.value

    def set_tdibc(self, bc, which, probe_is_column, probe_index, per_node, record_node,
                  capacity, rho0, cs, cp, dt, cu, gain, Zinf, A, lambdas, B, C, alpha, beta,
                  active=None):
        t = Tdibc()
        t.bc, t.which, t.probe_is_column, t.probe_index = bc, which, probe_is_column, probe_index
        t.per_node, t.record_node, t.capacity = per_node, record_node, capacity
        t.rho0, t.cs, t.cp, t.dt, t.cu, t.gain, t.Zinf = rho0, cs, cp, dt, cu, gain, Zinf
        for name, val in (("A", A), ("lambdas", lambdas), ("B", B), ("C", C),
                          ("alpha", alpha), ("beta", beta)):
            arr = getattr(t, name)
            arr[0], arr[1] = val
        self._tdibc_active = None
        if active is not None:
            self._tdibc_active = np.ascontiguousarray(active, dtype=np.int32)
            t.active = _i(self._tdibc_active)
        check(self._L.lbmx_set_tdibc(self.h, ctypes.byref(t)))
        self._tdibc_capacity = capacity

    def tdibc_history(self):
        rec = np.zeros((max(1, self._tdibc_capacity), 3))
        n = self._L.lbmx_tdibc_history(self.h, _d(rec), rec.shape[0])
        if n < 0:
            check(n)
        return rec[:n]

    def set_probes(self, xs, ys, capacity):
        x = np.ascontiguousarray(xs, dtype=np.int32)
        y = np.ascontiguousarray(ys, dtype=np.int32)
        check(self._L.lbmx_set_probes(self.h, _i(x), _i(y), len(x), int(capacity)))
        self._probe_shape = (int(capacity), len(x))

    def probe_history(self):
        rec = np.zeros(self._probe_shape)
        n = self._L.lbmx_probe_history(self.h, _d(rec), rec.shape[0])
        if n < 0:
            check(n)
        return rec[:n]

    def init_taylor_green(self, u0, kx, ky):
        check(self._L.lbmx_init_taylor_green(self.h, float(u0), float(kx), float(ky)))

    def step(self, n=1):
        check(self._L.lbmx_step(self.h, int(n)))

    def set_force_record(self, capacity, dt=1.0, cf=1.0):
        check(self._L.lbmx_set_force_record(self.h, float(dt), float(cf), int(capacity)))
        self._frec_cap = int(capacity)

    def force_record(self):
        rows = np.zeros((max(1, self._frec_cap), 3))
        n = self._L.lbmx_force_record(self.h, _d(rows), rows.shape[0])
        if n < 0:
            check(n)
        return rows[:n]

    def forces(self):
        Lf, Df = C.c_double(), C.c_double()
        check(self._L.lbmx_forces(self.h, C.byref(Lf), C.byref(Df)))
        return Lf.value, Df.value

    def totals(self):
        """FP64 (mass, kinetic energy) of the current moments over the fluid nodes,
        fixed summation order (lbmx_totals)."""
        m, e = C.c_double(), C.c_double()
        check(self._L.lbmx_totals(self.h, C.byref(m), C.byref(e)))
        return m.value, e.value

    def set_force_body(self, body):
        """Solid nodes (nonzero in the Ny x Nx array) whose links form the body
        force; None clears it."""
        b = None if body is None else np.ascontiguousarray(body, dtype=np.int32)
        check(self._L.lbmx_set_force_body(self.h, _i(b)))

    def body_forces(self):
        Lf, Df = C.c_double(), C.c_double()
        check(self._L.lbmx_body_forces(self.h, C.byref(Lf), C.byref(Df)))
        return Lf.value, Df.value

    def body_force_record(self):
        rows = np.zeros((max(1, self._frec_cap), 3))
        n = self._L.lbmx_body_force_record(self.h, _d(rows), rows.shape[0])
        if n < 0:
            check(n)
        return rows[:n]

    def moments(self):
        rho = np.empty((self.Ny, self.Nx))
        u = np.empty((self.Ny, self.Nx))
        v = np.empty((self.Ny, self.Nx))
        check(self._L.lbmx_moments(self.h, _d(rho), _d(u), _d(v)))
        return rho, u, v

    def write_vtk(self, filename, dx=1.0, cu=1.0, crho=1.0):
        """write_vtk_binary_2D of the current moments, packed on the device."""
        check(self._L.lbmx_write_vtk(self.h, os.fsencode(filename), float(dx), float(cu),
                                     float(crho)))

    def download(self, both=False):
        f = np.empty((self.Ny, self.Nx, 9))
        fn = np.empty((self.Ny, self.Nx, 9)) if both else None
        check(self._L.lbmx_download(self.h, _d(f), _d(fn)))
        return (f, fn) if both else f

    def synchronize(self):
        check(self._L.lbmx_synchronize(self.h))

    @property
    def device_bytes(self):
        return int(self._L.lbmx_device_bytes(self.h))
