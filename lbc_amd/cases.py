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
    if disc <= 0:
        raise ValueError("over-damped impedance: no complex pole pair")
    beta = math.sqrt(disc)
    return dict(Zinf=0.0, A=(0.0, 0.0), lambdas=(-1.0, -1.0),
                B=(-1.0 / (2.0 * M), 0.0), C=(-alpha / (2.0 * beta * M), 0.0),
                alpha=(alpha, -1.0), beta=(beta, 0.0))


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
    rho = s.rho0 + s.amp * np.exp(-math.log(2.0) * ((i - s.x0) / s.halfwidth) ** 2)
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
