"""Normal incidence on the liner's impedance wall: the oracle's run of the tube and
the reflection coefficient read from probe histories.

The tube (lbc_amd.batch.impedance_tube): periodic in x, the whole bottom row is the
impedance wall (velocity_bottom_regularized driven per node by the impedance ODE
with the liner's gain), a rigid top row, and a plane Gaussian pulse travelling
toward the wall.  The same lattice with a solid bottom row is the rigid tube.

|R(f)| = |P_wall(f)| / |P_rigid(f)| over the window that holds the reflected pulse at
the probe: the rigid wall reflects fully, and the two runs share the incident pulse,
the probe and the viscous attenuation of the path, which the ratio cancels.
"""
import ctypes as C
import math

import numpy as np

import oracle as O
from lbc_amd import batch as B
from lbc_amd import cases

CS = 1.0 / math.sqrt(3.0)


def tube(scale, impedance, Nx=4, tau=0.51, amp=1e-4, gain=-1.0):
    """The tube at length scale `scale`: probe 40 s above the wall, pulse centre
    160 s, half width 6 s, height 400 s, 700 s steps (the reflected pulse has
    passed the probe long before anything the top wall returns arrives)."""
    s = scale
    return B.impedance_tube(Nx, 400 * s, 700 * s, impedance, y0=160.0 * s,
                            halfwidth=6.0 * s, probe_y=40 * s, amp=amp, tau=tau, gain=gain)


def reflected_window(scale):
    """[t0, t1) of the probe samples that hold the reflected pulse: from half way
    between the incident and the reflected arrival, 400 s samples."""
    t_inc = (160.0 - 40.0) * scale / CS
    t_ref = (160.0 + 40.0) * scale / CS
    t0 = int(0.5 * (t_inc + t_ref))
    return t0, t0 + 400 * scale


def oracle_tube(case, fields, nsteps):
    """The tube on the oracle: each step, as orc_liner_run (oracle/lbm_oracle.c:702-712)
    -- the probe's lagged density, then per wall node p' = cs^2 (rho - rho0) of its
    lagged density through the impedance ODE (orc_tdibc_velocity), v_y = gain * v,
    then orc_step with the boundary table -- here with a periodic x axis.
    Returns (probe history [nsteps], final rho)."""
    lat = O.Lattice(case.Nx, case.Ny)
    rho, u, v = (np.array(a) for a in fields)
    lat.f = cases.equilibrium(rho, u, v)
    lat.f_new = lat.f.copy()
    lat.rho, lat.u, lat.v = rho, u, v
    lat.omega = np.ascontiguousarray(np.broadcast_to(case.omega_column(), (case.Ny, case.Nx)))
    lat.mask = case.solid_mask()
    table = case.boundary_table()
    states = []
    vbc = None
    if case.liner is not None:
        vbc = table[0][3]
        for _ in range(case.gap[0], case.gap[1] + 1):
            st = O.OrcTdibc()
            st.Zinf = case.liner["Zinf"]
            for n in ("A", "lambdas", "B", "C", "alpha", "beta"):
                arr = getattr(st, n)
                arr[0], arr[1] = case.liner[n]
            states.append(st)
    px, py = case.probes[0]
    hist = np.zeros(nsteps)
    for it in range(nsteps):
        hist[it] = lat.rho[py, px]
        for e, st in enumerate(states):
            i = case.gap[0] + e
            p = CS * CS * (lat.rho[0, i] - case.rho0) * 1.0
            vbc[i] = case.gain * (O.ORACLE.orc_tdibc_velocity(C.byref(st), p, 1.0) / 1.0)
        O.oracle_step(lat, table, case.per_x, case.per_y)
    return hist, lat.rho


def reflection(h_wall, h_rigid, scale, freqs, rho0=1.0):
    """|R(f)| from the two probe histories (densities) over the reflected window."""
    t0, t1 = reflected_window(scale)
    t = np.arange(t0, t1)
    pw = h_wall[t0:t1] - rho0
    pr = h_rigid[t0:t1] - rho0
    E = np.exp(-2j * math.pi * np.outer(freqs, t))
    return np.abs(E @ pw) / np.abs(E @ pr)


# Largest departure of the measured |R| from (Z - rho0 c0) / (Z + rho0 c0) over
# 0.5 .. 1.6 times the resonance, and |R| at the resonance, per tube scale (the
# oracle measures departures 0.204 / 0.091 / 0.042 / 0.020 and notches
# 0.023 / 0.010 / 0.005 / 0.002 at scales 1 / 2 / 4 / 8).
DEPARTURE = {1: 0.22, 2: 0.10, 4: 0.05, 8: 0.025}
NOTCH = {1: 0.03, 2: 0.015, 4: 0.0075, 8: 0.004}


def reflection_checks(scale, h_wall, h_rigid, imp_theory):
    """(largest |R| departure over 0.5 .. 1.6 f_res, |R(f_res)|, the frequency of the
    smallest |R| over 0.8 .. 1.2 f_res in units of f_res)."""
    f_res = 1.0 / (40.0 * scale)
    fs = f_res * np.linspace(0.5, 1.6, 23)
    Rm = reflection(h_wall, h_rigid, scale, fs)
    Rt = np.abs(cases.normal_reflection(imp_theory, 2.0 * math.pi * fs))
    fine = f_res * np.linspace(0.8, 1.2, 161)
    Rf = reflection(h_wall, h_rigid, scale, fine)
    dep = float(np.max(np.abs(Rm - Rt))) if np.all(np.isfinite(Rm)) else math.inf
    notch = float(reflection(h_wall, h_rigid, scale, [f_res])[0])
    return dep, notch, float(fine[np.argmin(Rf)] / f_res)


def passes(results):
    """results {scale: reflection_checks(...)}: every scale within its bounds, the
    notch within 1 % of the resonance, and the departure at least 1.8 times smaller
    at each doubling of the scale (first-order convergence)."""
    ok = all(results[s][0] <= DEPARTURE[s] and results[s][1] <= NOTCH[s] and
             abs(results[s][2] - 1.0) <= 0.01 for s in results)
    sc = sorted(results)
    return ok and all(results[a][0] >= 1.8 * results[b][0] for a, b in zip(sc, sc[1:]))


def msd(scale, R=1.0, f_res=1.0 / 40.0, M=8.0):
    """The served liners' mass-spring-damper wall (x rho0 c0), its resonance and
    reactance slope scaled with the tube so the pulse spectrum covers the notch."""
    return cases.LinerSpec(R=R, f_res=f_res / scale, M=M * scale).impedance()
