"""ctypes view of the batched engine and the request server (include/lbmx_batch.h).

:class:`Case` describes one lattice program; its methods ``solid_mask``,
``omega_column``, ``boundary_table`` and ``initial_fields`` state the same
program in the reference's own terms (solid mask, per-node omega, boundary table
in application order, initial moment arrays), which is what the oracle and the
single-lattice API (``single_lattice``) run for comparison.  :class:`Batch` and
:class:`Server` wrap the C objects.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import lib
from . import native
from .native import check

WALL_BOTTOM, WALL_TOP, WALL_LEFT, WALL_RIGHT = 1, 2, 4, 8
GUARD_NAN, GUARD_RHO, GUARD_SPEED = 1, 2, 4
ERR_DIVERGED = -5
INIT_PULSE, INIT_TGV, INIT_FIELDS = 0, 1, 2
CS = 1.0 / math.sqrt(3.0)

_dp = C.POINTER(C.c_double)


class Impedance(C.Structure):
    _fields_ = [("Zinf", C.c_double)] + [(n, C.c_double * 2) for n in
                                         ("A", "lambdas", "B", "C", "alpha", "beta")]


class CaseStruct(C.Structure):
    """lbmx_case."""
    _fields_ = [("Nx", C.c_int), ("Ny", C.c_int), ("per_x", C.c_int), ("per_y", C.c_int),
                ("walls", C.c_uint), ("gap_x0", C.c_int), ("gap_x1", C.c_int),
                ("tau", C.c_double), ("sponge", C.c_int), ("tau_sponge", C.c_double),
                ("Fx", C.c_double), ("Fy", C.c_double), ("top_velocity", C.c_int),
                ("U_top", C.c_double), ("V_top", C.c_double), ("convective_ends", C.c_int),
                ("liner", C.c_int), ("gain", C.c_double), ("imp", Impedance),
                ("init", C.c_int), ("rho0", C.c_double), ("amp", C.c_double),
                ("x0", C.c_double), ("halfwidth", C.c_double), ("u0", C.c_double),
                ("nprobe", C.c_int), ("probe_x", C.c_int * 4), ("probe_y", C.c_int * 4),
                ("steps", C.c_int), ("keep_moments", C.c_int)]


class Tuning(C.Structure):
    _fields_ = [("threads", C.c_int), ("tile", C.c_int), ("groups_per_cu", C.c_int),
                ("order", C.c_int)]

    def as_tuple(self):
        return (self.threads, self.tile, self.groups_per_cu, self.order)


class TlSpec(C.Structure):
    _fields_ = [("nfreq", C.c_int), ("f_lo", C.c_double), ("f_hi", C.c_double),
                ("up_probe", C.c_int), ("down_probe", C.c_int), ("up_t0", C.c_int),
                ("up_t1", C.c_int), ("down_t0", C.c_int), ("down_t1", C.c_int)]


class Result(C.Structure):
    _fields_ = [("tag", C.c_longlong), ("status", C.c_int), ("steps", C.c_int),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
                ("nfreq", C.c_int), ("tl_db", C.c_double * 64), ("up_abs", C.c_double * 64),
                ("down_abs", C.c_double * 64), ("guard", C.c_uint), ("mass", C.c_double),
                ("energy", C.c_double)]


_bound = False


def _bind():
    global _bound
    L = lib()
    if _bound:
        return L
    native._bind()
    vp, i, d, sz = C.c_void_p, C.c_int, C.c_double, C.c_size_t
    cp = C.POINTER(CaseStruct)
    L.lbmx_case_bytes.argtypes = [i, cp]
    L.lbmx_case_bytes.restype = sz
    L.lbmx_case_check.argtypes = [cp]
    L.lbmx_batch_create.argtypes = [i, sz, C.POINTER(vp)]
    L.lbmx_batch_destroy.argtypes = [vp]
    L.lbmx_batch_destroy.restype = None
    L.lbmx_batch_add.argtypes = [vp, cp, C.POINTER(i)]
    L.lbmx_batch_set_fields.argtypes = [vp, i, _dp, _dp, _dp]
    L.lbmx_batch_remove.argtypes = [vp, i]
    L.lbmx_batch_run.argtypes = [vp, i]
    L.lbmx_batch_time.argtypes = [vp, i, C.POINTER(d)]
    L.lbmx_batch_synchronize.argtypes = [vp]
    L.lbmx_batch_steps_done.argtypes = [vp, i, C.POINTER(i)]
    L.lbmx_batch_moments.argtypes = [vp, i, _dp, _dp, _dp]
    L.lbmx_batch_probes.argtypes = [vp, i, _dp, i]
    L.lbmx_batch_tune.argtypes = [vp, i, C.POINTER(Tuning)]
    L.lbmx_batch_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
    L.lbmx_batch_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
    L.lbmx_batch_set_chunk.argtypes = [vp, i]
    L.lbmx_server_create.argtypes = [i, sz, i, C.POINTER(Tuning), C.POINTER(vp)]
    L.lbmx_server_submit.argtypes = [vp, cp, C.POINTER(TlSpec), C.c_longlong]
    L.lbmx_server_poll.argtypes = [vp, C.POINTER(Result), i]
    L.lbmx_server_clock.argtypes = [vp]
    L.lbmx_server_clock.restype = d
    L.lbmx_server_stats.argtypes = [vp, C.POINTER(i), C.POINTER(i), C.POINTER(i),
                                    C.POINTER(sz)]
    L.lbmx_server_destroy.argtypes = [vp]
    L.lbmx_server_destroy.restype = None
    L.lbmx_batch_spectrum.argtypes = [vp, i, C.POINTER(TlSpec), C.POINTER(Result)]
    L.lbmx_batch_status.argtypes = [vp, i, C.POINTER(C.c_uint)]
    L.lbmx_batch_set_pairs.argtypes = [vp, i]
    L.lbmx_batch_totals.argtypes = [vp, i, C.POINTER(d), C.POINTER(d)]
    _bound = True
    return L


@dataclass
class Case:
    """One lattice program (the fields of lbmx_case)."""
    Nx: int
    Ny: int
    steps: int
    tau: float = 0.8
    per_x: int = 0
    per_y: int = 0
    walls: int = 0
    gap: tuple = (0, -1)
    sponge: int = 0
    tau_sponge: float = 0.9
    Fx: float = 0.0
    Fy: float = 0.0
    top_velocity: tuple | None = None      # (U_top, V_top)
    convective_ends: bool = False
    liner: dict | None = None              # impedance (Zinf, A, lambdas, B, C, alpha, beta)
    gain: float = -1.0
    init: int = INIT_PULSE
    rho0: float = 1.0
    amp: float = 0.0
    x0: float = 0.0
    halfwidth: float = 1.0
    u0: float = 0.0
    probes: list = field(default_factory=list)   # [(x, y)]
    keep_moments: bool = True
    kind: str = ""

    def struct(self) -> CaseStruct:
        s = CaseStruct()
        s.Nx, s.Ny, s.per_x, s.per_y = self.Nx, self.Ny, int(self.per_x), int(self.per_y)
        s.walls = self.walls
        s.gap_x0, s.gap_x1 = self.gap
        s.tau, s.sponge, s.tau_sponge = self.tau, self.sponge, self.tau_sponge
        s.Fx, s.Fy = self.Fx, self.Fy
        if self.top_velocity is not None:
            s.top_velocity = 1
            s.U_top, s.V_top = self.top_velocity
        s.convective_ends = int(self.convective_ends)
        if self.liner is not None:
            s.liner = 1
            s.gain = self.gain
            s.imp.Zinf = self.liner["Zinf"]
            for n in ("A", "lambdas", "B", "C", "alpha", "beta"):
                arr = getattr(s.imp, n)
                arr[0], arr[1] = self.liner[n]
        s.init, s.rho0, s.amp, s.x0, s.halfwidth, s.u0 = (self.init, self.rho0, self.amp,
                                                           self.x0, self.halfwidth, self.u0)
        s.nprobe = len(self.probes)
        for i, (x, y) in enumerate(self.probes):
            s.probe_x[i], s.probe_y[i] = x, y
        s.steps = self.steps
        s.keep_moments = int(self.keep_moments)
        return s

    # ---- the same program in the reference's terms ----------------------------------
    def solid_mask(self):
        j, i = np.mgrid[0:self.Ny, 0:self.Nx]
        m = np.zeros((self.Ny, self.Nx), dtype=bool)
        if self.walls & WALL_BOTTOM:
            m |= (j == 0) & ((i < self.gap[0]) | (i > self.gap[1]))
        if self.walls & WALL_TOP:
            m |= j == self.Ny - 1
        if self.walls & WALL_LEFT:
            m |= i == 0
        if self.walls & WALL_RIGHT:
            m |= i == self.Nx - 1
        return m.astype(np.int32)

    def omega_column(self):
        om = np.empty(self.Nx)
        for x in range(self.Nx):
            tau = self.tau
            d = min(x, self.Nx - 1 - x)
            if self.sponge > 0 and d < self.sponge:
                tau = self.tau_sponge + (self.tau - self.tau_sponge) * (d / self.sponge)
            om[x] = 1.0 / tau
        return om

    def initial_fields(self, fields=None):
        Nx, Ny = self.Nx, self.Ny
        if self.init == INIT_PULSE:
            if self.amp == 0.0:
                prof = np.full(Nx, self.rho0)
            else:
                prof = np.array([self.rho0 + self.amp * math.exp(
                    -math.log(2.0) * ((x - self.x0) / self.halfwidth) *
                    ((x - self.x0) / self.halfwidth)) for x in range(Nx)])
            rho = np.broadcast_to(prof, (Ny, Nx)).copy()
            return rho, np.zeros((Ny, Nx)), np.zeros((Ny, Nx))
        if self.init == INIT_TGV:
            kx, ky = 2.0 * math.pi / Nx, 2.0 * math.pi / Ny
            r = kx / ky
            sx = np.array([math.sin(kx * x) for x in range(Nx)])
            cx = np.array([math.cos(kx * x) for x in range(Nx)])
            c2x = np.array([math.cos(2.0 * kx * x) for x in range(Nx)])
            sy = np.array([math.sin(ky * y) for y in range(Ny)])[:, None]
            cy = np.array([math.cos(ky * y) for y in range(Ny)])[:, None]
            c2y = np.array([math.cos(2.0 * ky * y) for y in range(Ny)])[:, None]
            u = self.u0 * sx * cy
            v = -self.u0 * r * cx * sy
            rho = 1.0 - 0.75 * self.u0 * self.u0 * (c2x + r * r * c2y)
            return (np.ascontiguousarray(rho), np.ascontiguousarray(u),
                    np.ascontiguousarray(v))
        assert fields is not None
        return tuple(np.ascontiguousarray(a, dtype=np.float64) for a in fields)

    def boundary_table(self):
        """[(name, index, val1, val2)] in the order the batched kernel applies them;
        the impedance wall's val2 is filled in every step by its controller."""
        t = []
        if self.convective_ends:
            t.append(("convective_inlet", 0, None, None))
            t.append(("convective_outlet", self.Nx - 1, None, None))
        if self.liner is not None:
            t.append(("velocity_bottom_regularized", 0, np.zeros(self.Nx), np.zeros(self.Nx)))
        if self.top_velocity is not None:
            U, V = self.top_velocity
            t.append(("velocity_top", self.Ny - 1, np.full(self.Nx, float(U)),
                      np.full(self.Nx, float(V))))
        return t

    def single_lattice(self, precision, fields=None, record_steps=0):
        """The same program on the single-lattice API (lbmx.h), ready to step."""
        from .cases import equilibrium
        rho, u, v = self.initial_fields(fields)
        f = equilibrium(rho, u, v)
        lat = native.Lattice(self.Nx, self.Ny, precision=precision, per_x=self.per_x,
                             per_y=self.per_y, Fx=self.Fx, Fy=self.Fy, rho0=self.rho0)
        omega = np.ascontiguousarray(np.broadcast_to(self.omega_column(), (self.Ny, self.Nx)))
        lat.set_geometry(self.solid_mask(), omega)
        table = self.boundary_table()
        if table:
            lat.set_boundaries([(n, i) for n, i, _, _ in table])
            for b, (_, _, v1, v2) in enumerate(table):
                if v1 is not None:
                    lat.set_bc_values(b, v1, v2)
        if self.liner is not None:
            b = [n for n, _, _, _ in table].index("velocity_bottom_regularized")
            active = np.zeros(self.Nx, dtype=np.int32)
            active[self.gap[0]:self.gap[1] + 1] = 1
            lat.set_tdibc(bc=b, which=2, probe_is_column=0, probe_index=0, per_node=1,
                          record_node=self.gap[0], capacity=record_steps, rho0=self.rho0,
                          cs=CS, cp=1.0, dt=1.0, cu=1.0, gain=self.gain, active=active,
                          **self.liner)
        if self.probes and record_steps:
            lat.set_probes([p[0] for p in self.probes], [p[1] for p in self.probes],
                           record_steps)
        lat.upload(f, f, rho, u, v)
        return lat


def case_bytes(case: Case, precision=native.F32_SHIFTED) -> int:
    s = case.struct()
    return int(_bind().lbmx_case_bytes(precision, C.byref(s)))


def case_check(case: Case):
    s = case.struct()
    check(_bind().lbmx_case_check(C.byref(s)))


# ---- the four families of the batched workload ---------------------------------------
def poiseuille(Nx, Ny, tau, steps, Fx=1e-6):
    """Channel between solid rows driven by a body force, periodic in x."""
    return Case(Nx, Ny, steps, tau=tau, per_x=1, walls=WALL_BOTTOM | WALL_TOP, Fx=Fx,
                kind="poiseuille")


def couette(Nx, Ny, tau, steps, U=0.02):
    """Moving top row (velocity_top) over a solid bottom row, periodic in x."""
    return Case(Nx, Ny, steps, tau=tau, per_x=1, walls=WALL_BOTTOM, top_velocity=(U, 0.0),
                kind="couette")


def cavity(Nx, Ny, Re, steps, U=0.1):
    """Lid-driven cavity: solid bottom and sides, velocity_top lid; nu = U (Nx-2) / Re."""
    nu = U * (Nx - 2) / Re
    return Case(Nx, Ny, steps, tau=3.0 * nu + 0.5, walls=WALL_BOTTOM | WALL_LEFT | WALL_RIGHT,
                top_velocity=(U, 0.0), kind="cavity")


def taylor_green(Nx, Ny, steps, tau=0.8, u0=0.01):
    return Case(Nx, Ny, steps, tau=tau, per_x=1, per_y=1, init=INIT_TGV, u0=u0, kind="tgv")


def liner_duct(Nx, Ny, steps, impedance, tau=0.51, amp=1e-3, x0=None, halfwidth=20.0,
               liner=None, sponge=None, tau_sponge=0.9, probes=None, gain=-1.0):
    """The acoustic liner duct: rigid top wall, bottom wall with an impedance opening,
    convective ends with absorbing sponges, a Gaussian pressure pulse at rest."""
    x0 = Nx / 8.0 if x0 is None else x0
    liner = (3 * Nx // 8, 5 * Nx // 8) if liner is None else liner
    sponge = max(8, (3 * Nx) // 128) if sponge is None else sponge
    probes = [(Nx // 4, Ny // 2), (3 * Nx // 4, Ny // 2)] if probes is None else probes
    return Case(Nx, Ny, steps, tau=tau, walls=WALL_BOTTOM | WALL_TOP, gap=liner,
                sponge=sponge, tau_sponge=tau_sponge, convective_ends=True, liner=impedance,
                gain=gain, amp=amp, x0=x0, halfwidth=halfwidth, probes=probes,
                kind="liner")


def impedance_tube(Nx, Ny, steps, impedance, y0, halfwidth, probe_y, amp=1e-4, tau=0.51,
                   gain=-1.0):
    """Normal incidence on an impedance wall: periodic in x, the whole bottom row is
    the impedance wall (impedance None: a solid row, i.e. a rigid wall), rigid top
    row, and a Gaussian plane pulse centred at y0 travelling toward the bottom
    (rho' = amp g(y), v = -cs rho' / rho0).  Returns (case, fields) for
    Batch.add(case, fields)."""
    liner = impedance is not None
    c = Case(Nx, Ny, steps, tau=tau, per_x=1, walls=WALL_BOTTOM | WALL_TOP,
             gap=(0, Nx - 1) if liner else (0, -1), liner=impedance, gain=gain,
             init=INIT_FIELDS, probes=[(Nx // 2, probe_y)], kind="tube")
    y = np.arange(Ny, dtype=np.float64)
    g = amp * np.exp(-math.log(2.0) * ((y - y0) / halfwidth) ** 2)
    rho = np.ascontiguousarray(np.broadcast_to((1.0 + g)[:, None], (Ny, Nx)))
    v = np.ascontiguousarray(np.broadcast_to((-CS * g)[:, None], (Ny, Nx)))
    return c, (rho, np.zeros((Ny, Nx)), v)


def batched_workload(n=1024, steps=5000, seed=2024):
    """The batched benchmark set: sizes uniform in {64..1024} x {64..512} (multiples
    of 32), family uniform over Poiseuille (tau in [0.6, 1.2]), Couette, lid-driven
    cavity (Re in [100, 1000]) and Taylor-Green (u0 = 0.01)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        Nx = int(rng.integers(2, 33)) * 32
        Ny = int(rng.integers(2, 17)) * 32
        fam = int(rng.integers(0, 4))
        tau = float(rng.uniform(0.6, 1.2))
        if fam == 0:
            c = poiseuille(Nx, Ny, tau, steps)
        elif fam == 1:
            c = couette(Nx, Ny, tau, steps)
        elif fam == 2:
            c = cavity(Nx, Ny, float(rng.uniform(100.0, 1000.0)), steps)
        else:
            c = taylor_green(Nx, Ny, steps)
        out.append(c)
    return out


class Batch:
    def __init__(self, precision=native.F32_SHIFTED, budget=1 << 30, chunk=None):
        L = _bind()
        self._L = L
        self.h = C.c_void_p()
        check(L.lbmx_batch_create(precision, int(budget), C.byref(self.h)))
        self.precision = precision
        self.cases = {}
        if chunk is not None:
            self.set_chunk(chunk)

    def close(self):
        if self.h:
            self._L.lbmx_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, case: Case, fields=None) -> int:
        s = case.struct()
        i = C.c_int()
        check(self._L.lbmx_batch_add(self.h, C.byref(s), C.byref(i)))
        self.cases[i.value] = case
        if case.init == INIT_FIELDS:
            rho, u, v = (np.ascontiguousarray(a, dtype=np.float64) for a in fields)
            check(self._L.lbmx_batch_set_fields(self.h, i.value, rho.ctypes.data_as(_dp),
                                                u.ctypes.data_as(_dp), v.ctypes.data_as(_dp)))
        return i.value

    def remove(self, i):
        check(self._L.lbmx_batch_remove(self.h, i))
        self.cases.pop(i, None)

    def run(self, n):
        check(self._L.lbmx_batch_run(self.h, int(n)))

    def time(self, n):
        ms = C.c_double()
        check(self._L.lbmx_batch_time(self.h, int(n), C.byref(ms)))
        return ms.value

    def synchronize(self):
        check(self._L.lbmx_batch_synchronize(self.h))

    def steps_done(self, i):
        s = C.c_int()
        check(self._L.lbmx_batch_steps_done(self.h, i, C.byref(s)))
        return s.value

    def moments(self, i):
        c = self.cases[i]
        rho, u, v = (np.empty((c.Ny, c.Nx)) for _ in range(3))
        check(self._L.lbmx_batch_moments(self.h, i, rho.ctypes.data_as(_dp),
                                         u.ctypes.data_as(_dp), v.ctypes.data_as(_dp)))
        return rho, u, v

    def probes(self, i):
        c = self.cases[i]
        h = np.zeros((c.steps + 1, max(1, len(c.probes))))
        n = self._L.lbmx_batch_probes(self.h, i, h.ctypes.data_as(_dp), h.shape[0])
        if n < 0:
            check(n)
        return h[:n, :len(c.probes)]

    def status(self, i):
        """LBMX_GUARD_* bits lattice i has tripped (0: healthy)."""
        g = C.c_uint()
        check(self._L.lbmx_batch_status(self.h, i, C.byref(g)))
        return g.value

    def totals(self, i):
        """FP64 (mass, kinetic energy) of completed keep_moments lattice i."""
        m, e = C.c_double(), C.c_double()
        check(self._L.lbmx_batch_totals(self.h, i, C.byref(m), C.byref(e)))
        return m.value, e.value

    def spectrum(self, i, tl: TlSpec) -> Result:
        """The server's device spectrum and TL of completed lattice i."""
        r = Result()
        check(self._L.lbmx_batch_spectrum(self.h, i, C.byref(tl), C.byref(r)))
        return r

    def tune(self, trial_steps=4):
        t = Tuning()
        check(self._L.lbmx_batch_tune(self.h, int(trial_steps), C.byref(t)))
        return t.as_tuple()

    def set_tuning(self, threads, tile, groups_per_cu, order):
        t = Tuning(threads, tile, groups_per_cu, order)
        check(self._L.lbmx_batch_set_tuning(self.h, C.byref(t)))

    def tuning(self):
        t = Tuning()
        check(self._L.lbmx_batch_get_tuning(self.h, C.byref(t)))
        return t.as_tuple()

    def set_chunk(self, steps):
        check(self._L.lbmx_batch_set_chunk(self.h, int(steps)))

    def set_pairs(self, on=True):
        """Two steps per pass (lbmx_batch_set_pairs); runs must then be even."""
        check(self._L.lbmx_batch_set_pairs(self.h, int(bool(on))))


class Server:
    def __init__(self, precision=native.F32_SHIFTED, budget=200 << 30, chunk=64, tuning=None):
        L = _bind()
        self._L = L
        self.h = C.c_void_p()
        t = Tuning(*tuning) if tuning is not None else None
        check(L.lbmx_server_create(precision, int(budget), int(chunk),
                                   C.byref(t) if t is not None else None, C.byref(self.h)))

    def submit(self, case: Case, tl: TlSpec | None, tag: int):
        """Queues a request; tl None: no spectrum (a channel solve), the result
        carries its status, timing, mass and energy only."""
        s = case.struct()
        check(self._L.lbmx_server_submit(self.h, C.byref(s), None if tl is None else C.byref(tl),
                                         int(tag)))

    def poll(self, max_results=256):
        buf = (Result * max_results)()
        n = self._L.lbmx_server_poll(self.h, buf, max_results)
        if n < 0:
            check(n)
        return [buf[i] for i in range(n)]

    def clock(self):
        return self._L.lbmx_server_clock(self.h)

    def stats(self):
        w, r, f = C.c_int(), C.c_int(), C.c_int()
        p = C.c_size_t()
        check(self._L.lbmx_server_stats(self.h, C.byref(w), C.byref(r), C.byref(f),
                                        C.byref(p)))
        return dict(waiting=w.value, running=r.value, finished=f.value, peak_bytes=p.value)

    def close(self):
        if self.h:
            self._L.lbmx_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
