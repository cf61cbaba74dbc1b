"""Request generation and the client side of the serving loop (lbmx_server).

A request is an acoustic liner duct (:func:`lbc_amd.batch.liner_duct`) with seeded
size, impedance, pulse amplitude and length, run in FP32 storage on the device
while other requests run beside it; its answer is a transmission-loss spectrum
from two centre-line probes.

Transmission loss, as this project defines it (the reference's own TL
post-processing lives in an unvendored toolkit, so there is nothing to match):

    TL(f) = 20 log10( |P_up(f)| / |P_down(f)| )

with P(f) = sum_t p'(t) exp(-2 pi i f t), p' = cs^2 (rho - rho0) the lagged
density of the probe before step t, summed over windows of equal length centred on
the pulse's arrival at each probe: the upstream probe (x = Nx/4) sees the incident
pulse at t_u = (Nx/8)/cs and the liner's first reflection only after 2 t_u, so its
window is [0, 2 t_u); the downstream probe (x = 3 Nx/4) sees the transmitted pulse
at t_d = (5 Nx/8)/cs and its window is [t_d - t_u, t_d + t_u), which ends before
any wave reflected at either duct end can arrive.  A rigid duct gives TL near 0:
only viscous and dispersive loss between the probes (a few tenths of a dB at the
low end of the band at tau = 0.51, rising with frequency).
"""
from __future__ import annotations

import math
import threading
import time
from dataclasses import dataclass

import numpy as np

from . import batch as B
from .cases import ade_stable, msd_ade

CS = B.CS
NFREQ = 32
F_LO, F_HI = 1.0 / 160.0, 1.0 / 20.0
PULSE_HALFWIDTH = 6.0


@dataclass
class LinerRequest:
    Nx: int
    Ny: int
    R: float          # resistance, x rho0 c0
    f_res: float      # resonance, cycles per step
    M: float          # mass, x rho0 c0 steps (the reactance slope)
    amp: float        # pulse amplitude, x rho0
    steps: int
    rigid: bool = False

    def impedance(self, rho0=1.0):
        Rr = self.R * rho0 * CS
        M = self.M * rho0 * CS
        w = 2.0 * math.pi * self.f_res
        return msd_ade(M, Rr, M * w * w)

    def case(self, keep_moments=False) -> B.Case:
        Nx, Ny = self.Nx, self.Ny
        probes = [(Nx // 4, Ny // 2), (3 * Nx // 4, Ny // 2)]
        if self.rigid:
            c = B.Case(Nx, Ny, self.steps, tau=0.51, walls=B.WALL_BOTTOM | B.WALL_TOP,
                       sponge=max(8, (3 * Nx) // 128), tau_sponge=0.9, convective_ends=True,
                       amp=self.amp, x0=Nx / 8.0, halfwidth=PULSE_HALFWIDTH, probes=probes,
                       kind="rigid")
        else:
            c = B.liner_duct(Nx, Ny, self.steps, self.impedance(), tau=0.51, amp=self.amp,
                             x0=Nx / 8.0, halfwidth=PULSE_HALFWIDTH, probes=probes)
        c.keep_moments = keep_moments
        return c

    def windows(self):
        """[(t0, t1) upstream, (t0, t1) downstream] sample windows of the spectra.
        A request shorter than min_steps(Nx) ends before the transmitted pulse has
        passed the downstream probe, so it has no TL: ValueError."""
        if self.steps < min_steps(self.Nx):
            raise ValueError(f"a {self.Nx}-wide duct needs >= {min_steps(self.Nx)} steps for "
                             f"its transmitted pulse (asked for {self.steps})")
        t_u = (self.Nx / 8.0) / CS
        t_d = (5.0 * self.Nx / 8.0) / CS
        up = (0, int(round(2.0 * t_u)))
        dn = (int(round(t_d - t_u)), int(round(t_d + t_u)))
        return up, dn

    def tl_spec(self) -> B.TlSpec:
        (u0, u1), (d0, d1) = self.windows()
        return B.TlSpec(NFREQ, F_LO, F_HI, 0, 1, u0, u1, d0, d1)


def min_steps(Nx):
    """Shortest run whose transmitted pulse (from x = Nx/8) has passed the
    downstream probe's window (x = 3 Nx/4, half-width Nx/8 / cs): 3 Nx / (4 cs)."""
    return int(math.ceil(0.75 * Nx / CS))


def draw_requests(n, seed=7, Nx=(1024, 8192), Ny=(64, 256), R=(0.5, 3.0),
                  f_res=(1.0 / 80.0, 1.0 / 20.0), amp=(1e-4, 1e-3), steps=(2000, 20000),
                  M=(1.0, 16.0)):
    """n seeded requests: Nx, Ny uniform (multiples of 32); R, f_res, mass and
    amplitude uniform in the given ranges; the length uniform in the given range
    above min_steps(Nx) (a shorter run has no transmitted pulse to measure, so
    such a draw is drawn again).  The mass is drawn again until the impedance ODE
    is stable under the reference's explicit step (cases.ade_stable): an
    over-damped draw with a small mass has a real pole below -2, which that step
    amplifies (|1 + lambda| > 1) until the request diverges; a complex pair is
    updated in place (nit.c:31), amplification [[1+a, -b], [b(1+a), 1+a-b^2]] with
    determinant (1+a)^2 < 1 for the damped poles a +- ib."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nx = int(rng.integers(Nx[0] // 32, Nx[1] // 32 + 1)) * 32
        ny = int(rng.integers(Ny[0] // 32, Ny[1] // 32 + 1)) * 32
        r = float(rng.uniform(*R))
        fr = float(rng.uniform(*f_res))
        m = float(rng.uniform(*M))
        while not ade_stable(LinerRequest(nx, ny, r, fr, m, 0.0, 0).impedance()):
            m = float(rng.uniform(*M))
        a = float(rng.uniform(*amp))
        if min_steps(nx) > steps[1]:
            raise ValueError(f"steps up to {steps[1]} cannot cover a {nx}-wide duct")
        st = int(rng.integers(steps[0], steps[1] + 1))
        while st < min_steps(nx):
            st = int(rng.integers(steps[0], steps[1] + 1))
        out.append(LinerRequest(nx, ny, r, fr, m, a, st))
    return out


def poisson_arrivals(n, rate, seed=11):
    """Arrival times (s) of n requests of a Poisson stream of the given rate."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


def spectrum(hist_col, window, f, rho0=1.0):
    """P(f) of one probe column over samples [t0, t1) on the host (the server's
    definition)."""
    t0, t1 = window
    t = np.arange(t0, t1, dtype=np.float64)
    p = CS * CS * (hist_col[t0:t1] - rho0)
    return np.array([np.sum(p * np.exp(-2j * np.pi * fk * t)) for fk in f])


def tl_from_history(hist, req: LinerRequest):
    f = np.linspace(F_LO, F_HI, NFREQ)
    wu, wd = req.windows()
    up = spectrum(hist[:, 0], wu, f)
    dn = spectrum(hist[:, 1], wd, f)
    return 20.0 * np.log10(np.abs(up) / np.abs(dn)), np.abs(up), np.abs(dn)


def serve(server: B.Server, reqs, arrivals, poll=0.002):
    """Submits reqs at the given arrival times (seconds from now) and collects
    every result.  Returns {tag: Result} and the wall time of the whole run."""
    results = {}
    t0 = time.perf_counter()
    done = threading.Event()

    def submitter():
        for i, (r, ta) in enumerate(zip(reqs, arrivals)):
            dt = ta - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            server.submit(r.case(), r.tl_spec(), i)
        done.set()

    th = threading.Thread(target=submitter)
    th.start()
    while len(results) < len(reqs):
        for res in server.poll():
            results[res.tag] = res
        if len(results) < len(reqs):
            time.sleep(poll)
    th.join()
    return results, time.perf_counter() - t0


def latency_summary(results):
    lat = np.array([r.t_done - r.t_submit for r in results.values()])
    return {"p50_s": float(np.percentile(lat, 50)), "p95_s": float(np.percentile(lat, 95)),
            "p99_s": float(np.percentile(lat, 99)), "mean_s": float(lat.mean())}
