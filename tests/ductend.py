"""Reflection of a duct end, read from a probe history.

A rigid duct (walls top and bottom) holds a Gaussian pulse at rest (half width 3)
150 cells before a probe that stands `gap` cells before the end's absorbing
layer; the pulse's right-going half passes the probe, meets the end and comes
back.  |R(f)| = |P_end(f)| / |P_rigid(f)| over the window holding the returned
pulse, where the rigid duct closes that end with a solid column (full
reflection): the two runs share the incident pulse, the path and its viscous
attenuation.  The left-going half returns only after the window.
"""
import math

import numpy as np

from lbc_amd import batch as B

CS = 1.0 / math.sqrt(3.0)


def layout(Nx, sponge, gap=60):
    """(probe x, pulse x, first sample of the returned pulse, steps)."""
    xp = Nx - 1 - sponge - gap
    x0 = xp - 150
    t_inc = (xp - x0) / CS
    t_ret = ((Nx - 1 - x0) + (Nx - 1 - xp)) / CS
    mid = int(0.5 * (t_inc + t_ret))
    steps = mid + 1600
    assert (x0 + xp) / CS > steps      # the left-going half is still away
    return xp, x0, mid, steps


def duct(Nx, Ny, sponge, end, gap=60):
    """end: 'rigid' (solid end columns), 'convective' (convective_inlet/outlet),
    'sponge' (convective ends inside a tau ramp `sponge` cells wide), 'open'
    (solid left column only; the caller adds its own right-end condition)."""
    xp, x0, mid, steps = layout(Nx, sponge, gap)
    kw = dict(tau=0.51, walls=B.WALL_BOTTOM | B.WALL_TOP, amp=1e-4, x0=float(x0),
              halfwidth=3.0, probes=[(xp, Ny // 2)])
    if end == "rigid":
        kw["walls"] |= B.WALL_LEFT | B.WALL_RIGHT
    elif end == "convective":
        kw["convective_ends"] = True
    elif end == "sponge":
        kw.update(convective_ends=True, sponge=sponge, tau_sponge=0.9)
    elif end == "open":
        kw["walls"] |= B.WALL_LEFT
    return B.Case(Nx, Ny, steps, **kw), mid


def reflection(h_end, h_rigid, mid, freqs):
    t = np.arange(mid, len(h_rigid))
    E = np.exp(-2j * math.pi * np.outer(freqs, t))
    return np.abs(E @ (h_end[mid:] - 1.0)) / np.abs(E @ (h_rigid[mid:] - 1.0))
