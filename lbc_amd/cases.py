"""Builders of the workloads the benchmarks and the request server run.

These play the role of the reference's case drivers (cases/*/*.c): they set up
geometry, boundary table, initial state and, for liners, the impedance model,
then hand a resident lattice to the native library.  No arithmetic of the time
step happens here.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import native

CS = 1.0 / math.sqrt(3.0)
W9 = np.array([4.0 / 9.0] + [1.0 / 9.0] * 4 + [1.0 / 36.0] * 4)
CX9 = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
CY9 = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])


def equilibrium(rho, u, v):
    """Second-order D2Q9 equilibrium [Ny, Nx, 9] (the form of collision.c:31-32)."""
    cu = CX9 * u[..., None] + CY9 * v[..., None]
    usq = (u * u + v * v)[..., None]
    return W9 * rho[..., None] * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq)


# ---------------------------------------------------------------------------
# Taylor-Green vortex (periodic), generated on the device
# ---------------------------------------------------------------------------
def taylor_green(N, precision=native.F64, tau=0.8, u0=0.01, Ny=None):
    Ny = N if Ny is None else Ny
    lat = native.Lattice(N, Ny, precision=precision, per_x=1, per_y=1)
    lat.set_uniform(1.0 / tau)
    lat.set_boundaries([("periodic_x", 0), ("periodic_y", 0)])
    lat.init_taylor_green(u0, 2.0 * math.pi / N, 2.0 * math.pi / Ny)
    return lat


# ---------------------------------------------------------------------------
# Channel past a cylinder (interpolated bounce-back, regularized inlet/outlet)
# ---------------------------------------------------------------------------
def cylinder_channel(Nx=2048, Ny=512, R=32.0, xc=512.0, yc=256.0, umax=0.05, Re=100.0,
                     precision=native.F64):
    nu = umax * (2 * R) / Re
    tau = 3.0 * nu + 0.5
    j, i = np.mgrid[0:Ny, 0:Nx]
    walls = (j == 0) | (j == Ny - 1)
    d = np.sqrt((i - xc) ** 2 + (j - yc) ** 2) - R
    mask = (walls | (d <= 0)).astype(np.int32)
    phi = np.where(walls, -1e30, d)          # walls: q clamps to 1e-12 (plain BB)
    lat = native.Lattice(Nx, Ny, precision=precision, use_ibb=1)
    lat.set_geometry(mask, np.full((Ny, Nx), 1.0 / tau), phi)
    lat.set_boundaries([("velocity_inlet_regularized", 0),
                        ("pressure_outlet_regularized", Nx - 1)])
    y = np.arange(Ny) - 0.5                  # half-way walls at y = 0.5 and Ny - 1.5
    H = Ny - 2.0
    uin = np.clip(4 * umax * y * (H - y) / H ** 2, 0.0, None)
    lat.set_bc_values(0, uin, np.zeros(Ny))
    lat.set_bc_values(1, np.ones(Ny), None)
    rho = np.ones((Ny, Nx))
    zero = np.zeros((Ny, Nx))
    f = equilibrium(rho, zero, zero)
    lat.upload(f, f, rho, zero, zero)
    return lat, dict(tau=tau, nu=nu)


# ---------------------------------------------------------------------------
# Normal-incidence impedance tube with a TDIBC end wall (cases/nit/nit.c)
# ---------------------------------------------------------------------------
@dataclass
class NitSpec:
    p_a: float = 89.0          # pressure amplitude [Pa]
    f_exc: float = 2000.0      # excitation frequency [Hz]
    variant: int = 0           # 0: orifice rows get the area-scaled velocity (nit.c:286-300)
                               # 1: every row gets v_ac / cu ("uniform" run)
    n_orifices: int = 2
    Ly: float = 0.009906
    y1: float = 0.002528
    y2: float = 0.007478
    d: float = 0.00117
    c0: float = 340.0
    tau: float = 0.5005
    nu: float = 1.5e-5
    rho0: float = 1.184

    def derived(self):
        """nit.c:104-130, with the same floating-point expression order."""
        cs = 1.0 / math.sqrt(3.0)
        dx = 3 * self.nu * cs / (self.c0 * (self.tau - 0.5))
        dt = dx * dx * (self.tau - 0.5) / (3 * self.nu)
        cu = dx / dt
        cp = 1.0 * cu * cu
        tot_cycles = 1 + 10 + 2
        lam = self.c0 / self.f_exc
        Lx = tot_cycles * lam
        kx = 2 * math.pi / lam
        rho_a = (self.p_a / cp) / (cs * cs)
        x_end = Lx - 1 * lam
        x_start = x_end - 10 * lam
        Nx = int(math.ceil(Lx / dx))
        Ny = int(math.ceil(self.Ly / dx))
        Nt = int(math.ceil(tot_cycles / (self.f_exc * dt)))
        return dict(cs=cs, dx=dx, dt=dt, cu=cu, cp=cp, kx=kx, rho_a=rho_a, x_start=x_start,
                    x_end=x_end, Nx=Nx, Ny=Ny, Nt=Nt)


def nit_initial_fields(s: NitSpec):
    """rho_in, U_in, V_in of nit.c:234-247 (libm sin per column, as the C code)."""
    D = s.derived()
    Nx, Ny, dx = D["Nx"], D["Ny"], D["dx"]
    x = np.arange(Nx) * dx
    env = ((x >= D["x_start"]) & (x <= D["x_end"])).astype(np.float64)
    carrier = np.array([math.sin(D["kx"] * xi) for xi in x])
    rho_p = D["rho_a"] * env * carrier
    rho_in = np.broadcast_to(s.rho0 + rho_p, (Ny, Nx)).copy()
    U_in = np.broadcast_to(D["cs"] * rho_p / s.rho0, (Ny, Nx)).copy()
    V_in = np.zeros((Ny, Nx))
    return rho_in, U_in, V_in


def nit_tube(s: NitSpec, impedance, record_steps, precision=native.F64):
    """The NIT tube on the device: Mei's initialization on the device, then a lattice
    with periodic_y, pressure_inlet(0) and velocity_outlet_regularized(Nx-1) whose
    velocity is driven by the lumped TDIBC controller.  impedance: dict with Zinf, A,
    lambdas, B, C, alpha, beta (the impedance_<p>Pa.txt fit)."""
    D = s.derived()
    Nx, Ny = D["Nx"], D["Ny"]
    rho_in, U_in, V_in = nit_initial_fields(s)
    mask = np.zeros((Ny, Nx), dtype=np.int32)
    omega = np.full((Ny, Nx), 1.0 / s.tau)
    f = np.zeros((Ny, Nx, 9))
    f_new = np.zeros((Ny, Nx, 9))
    rho = np.empty((Ny, Nx))
    u = np.empty((Ny, Nx))
    v = np.empty((Ny, Nx))
    native.mei_initialization(f, f_new, rho_in, U_in, V_in, rho, u, v, omega, 0, 1, 100, 1e-10,
                              mask)
    lat = native.Lattice(Nx, Ny, precision=precision, per_x=0, per_y=1)
    lat.set_geometry(mask, omega)
    lat.set_boundaries([("periodic_y", 0), ("pressure_inlet", 0),
                        ("velocity_outlet_regularized", Nx - 1)])
    lat.set_bc_values(1, np.full(Ny, s.rho0), None)
    lat.set_bc_values(2, np.zeros(Ny), np.zeros(Ny))
    y = np.arange(Ny) * D["dx"]
    d2 = s.d / 2
    holes = ((y >= s.y1 - d2) & (y <= s.y1 + d2)) | ((y >= s.y2 - d2) & (y <= s.y2 + d2))
    if s.variant == 0:
        gain = (4 * s.Ly * s.Ly) / (s.n_orifices * s.d * s.d)
        active = holes.astype(np.int32)
    else:
        gain = 1.0
        active = np.ones(Ny, dtype=np.int32)
    lat.set_tdibc(bc=2, which=1, probe_is_column=1, probe_index=Nx - 1, per_node=0,
                  record_node=0, capacity=record_steps, rho0=s.rho0, cs=D["cs"], cp=D["cp"],
                  dt=D["dt"], cu=D["cu"], gain=gain, active=active, **impedance)
    lat.upload(f, f_new, rho, u, v)
    return lat


# ---------------------------------------------------------------------------
# Acoustic-liner duct with a time-domain impedance wall
# ---------------------------------------------------------------------------
def msd_ade(M, Rr, K):
    """Mass-spring-damper admittance Y(s) = s / (M s^2 + Rr s + K) as the ODE
    coefficients of cases/nit/nit.c:13-47 (one complex pole pair, no real pole,
    Zinf = 0).  With a = alpha + i beta the upper root of M s^2 + Rr s + K, the
    residue r = a / (2 i beta M) of Y at a gives B = -Re r, C = Im r, since the
    pair contributes -(B - iC)/(s - a) - (B + iC)/(s - conj a)."""
    alpha = -Rr / (2.0 * M)
    disc = K / M - alpha * alpha
    if disc > 1e-12 * (K / M):
        beta = math.sqrt(disc)
        return dict(Zinf=0.0, A=(0.0, 0.0), lambdas=(-1.0, -1.0),
                    B=(-1.0 / (2.0 * M), 0.0), C=(-alpha / (2.0 * beta * M), 0.0),
                    alpha=(alpha, -1.0), beta=(beta, 0.0))
    # over-damped (or critically damped, split slightly): two real poles r1, r2;
    # Y = a1/(s - r1) + a2/(s - r2), a_i = r_i / (M (r_i - r_j)), and a real pole
    # enters the ODE as -A/(s - lambda), so A_i = -a_i, lambda_i = r_i
    root = math.sqrt(max(-disc, 1e-6 * (K / M)))
    r1, r2 = alpha + root, alpha - root
    a1 = r1 / (M * (r1 - r2))
    a2 = r2 / (M * (r2 - r1))
    return dict(Zinf=0.0, A=(-a1, -a2), lambdas=(r1, r2), B=(0.0, 0.0), C=(0.0, 0.0),
                alpha=(-1.0, -1.0), beta=(0.0, 0.0))


def ade_stable(imp, margin=1e-9):
    """Whether the impedance ODE decays under the reference's step (nit.c:19-35,
    dt = 1): a real pole updates phi <- (1 + lambda) phi, so it needs
    |1 + lambda| < 1; a complex pair updates psi first and chi with the new psi,
    [[1+a, -b], [b(1+a), 1+a-b^2]], stable when its determinant (1+a)^2 < 1 and
    |trace| = |2(1+a) - b^2| < 1 + (1+a)^2 (Jury).  An over-damped mass-spring-
    damper has a fast real pole near -R/M: small masses with large resistances
    fail the first rule."""
    for i in range(2):
        if imp["A"][i] != 0.0 and not abs(1.0 + imp["lambdas"][i]) < 1.0 - margin:
            return False
        if imp["B"][i] != 0.0 or imp["C"][i] != 0.0:
            a, b = imp["alpha"][i], imp["beta"][i]
            det = (1.0 + a) ** 2
            tr = 2.0 * (1.0 + a) - b * b
            if not (det < 1.0 - margin and abs(tr) < 1.0 + det - margin):
                return False
    return True


def ade_admittance(imp, w):
    """Y(i w) of the impedance ODE in continuous time (nit.c:19-45 with dt -> 0):
    v = Zinf p - sum A phi - 2 sum (B psi + C chi), phi' = p + lambda phi,
    psi' = p + alpha psi - beta chi, chi' = beta psi + alpha chi.  For p = e^{st}:
    Phi = 1/(s - lambda), Psi = (s - alpha)/D, Chi = beta/D, D = (s - alpha)^2 + beta^2."""
    s = 1j * np.asarray(w, dtype=np.float64)
    Y = np.full(s.shape, complex(imp["Zinf"]))
    for i in range(2):
        if imp["A"][i] != 0.0:
            Y = Y - imp["A"][i] / (s - imp["lambdas"][i])
        if imp["B"][i] != 0.0 or imp["C"][i] != 0.0:
            a, b = imp["alpha"][i], imp["beta"][i]
            D = (s - a) ** 2 + b * b
            Y = Y - 2.0 * (imp["B"][i] * (s - a) + imp["C"][i] * b) / D
    return Y


def normal_reflection(imp, w, rho0=1.0, gain=-1.0):
    """Plane-wave reflection coefficient of an impedance wall at normal incidence,
    R = (Z - rho0 c0) / (Z + rho0 c0) = (1 - rho0 c0 Y) / (1 + rho0 c0 Y), where Y
    relates the velocity INTO the wall to p'.  The nit tube's outlet takes the ODE
    velocity as u, toward its wall (nit.c:302-303); the liner's bottom wall imposes
    v_y = gain * v_ODE, so the velocity into it is -gain * v_ODE."""
    Y = -gain * ade_admittance(imp, w)
    return (1.0 - rho0 * CS * Y) / (1.0 + rho0 * CS * Y)


@dataclass
class LinerSpec:
    Nx: int = 8192
    Ny: int = 256
    tau: float = 0.51
    rho0: float = 1.0
    amp: float = 1e-3            # pulse amplitude (x rho0)
    x0: float = 1024.0           # pulse centre
    halfwidth: float = 20.0      # half width at half maximum (cells)
    liner_x0: int = 3072
    liner_x1: int = 5120
    R: float = 1.0               # resistance (x rho0 c0)
    f_res: float = 1.0 / 40.0    # resonance (cycles per step)
    M: float = 8.0               # mass per area (x rho0 c0 steps): reactance slope / 2
    sponge: int = 192            # absorbing-layer width at each duct end
    tau_sponge: float = 0.9
    probe_x: int = 6144          # transmitted-pulse probe
    probe_y: int = 128

    def impedance(self):
        Rr = self.R * self.rho0 * CS
        M = self.M * self.rho0 * CS
        w = 2.0 * math.pi * self.f_res
        return msd_ade(M, Rr, M * w * w)


def liner_arrays(s: LinerSpec):
    """(f, rho, u, v, omega, mask) of the liner duct's initial state."""
    j, i = np.mgrid[0:s.Ny, 0:s.Nx]
    # libm exp per column (the pulse depends on x only), as a C driver computes it
    pulse = np.array([s.rho0 + s.amp * math.exp(-math.log(2.0) * ((x - s.x0) / s.halfwidth)
                                                * ((x - s.x0) / s.halfwidth))
                      for x in range(s.Nx)])
    rho = np.broadcast_to(pulse, (s.Ny, s.Nx)).copy()
    u = np.zeros((s.Ny, s.Nx))
    v = np.zeros((s.Ny, s.Nx))
    wall = (j == s.Ny - 1) | ((j == 0) & ((i < s.liner_x0) | (i > s.liner_x1)))
    mask = wall.astype(np.int32)
    d = np.minimum(i, s.Nx - 1 - i).astype(np.float64)
    tau = np.full((s.Ny, s.Nx), s.tau)
    if s.sponge > 0:
        ramp = s.tau_sponge + (s.tau - s.tau_sponge) * (d / s.sponge)
        tau = np.where(d < s.sponge, ramp, tau)
    omega = 1.0 / tau
    f = equilibrium(rho, u, v)
    return f, rho, u, v, omega, mask


def liner_duct(s: LinerSpec, precision=native.F64, record_steps=0):
    """A resident liner-duct lattice; the impedance wall is driven on the device."""
    f, rho, u, v, omega, mask = liner_arrays(s)
    lat = native.Lattice(s.Nx, s.Ny, precision=precision, rho0=s.rho0)
    lat.set_geometry(mask, omega)
    lat.set_boundaries([("convective_inlet", 0), ("convective_outlet", s.Nx - 1),
                        ("velocity_bottom_regularized", 0)])
    lat.set_bc_values(2, np.zeros(s.Nx), np.zeros(s.Nx))
    imp = s.impedance()
    active = np.zeros(s.Nx, dtype=np.int32)
    active[s.liner_x0:s.liner_x1 + 1] = 1
    lat.set_tdibc(bc=2, which=2, probe_is_column=0, probe_index=0, per_node=1,
                  record_node=(s.liner_x0 + s.liner_x1) // 2, capacity=record_steps,
                  rho0=s.rho0, cs=CS, cp=1.0, dt=1.0, cu=1.0, gain=-1.0, active=active, **imp)
    if record_steps:
        lat.set_probes([s.probe_x], [s.probe_y], record_steps)
    lat.upload(f, f, rho, u, v)
    return lat
