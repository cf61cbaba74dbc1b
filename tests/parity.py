"""Helpers shared by the parity tests: run a scenario on the oracle and on the
HIP library, compare state bit for bit (FP64) or within a stated tolerance."""
from __future__ import annotations

import numpy as np

import oracle as O
from cases import Scenario

# Required FP64 agreement with the reference (relative, in rho and u).  The FP64
# HIP path is expected to be bit-identical; the tolerance is what the tests enforce.
FP64_RTOL = 1e-12


def oracle_lattice(sc: Scenario, state):
    f, fn, rho, u, v = state
    lat = O.Lattice(sc.Nx, sc.Ny)
    lat.f[:] = f
    lat.f_new[:] = fn
    lat.rho[:] = rho
    lat.u[:] = u
    lat.v[:] = v
    lat.omega[:] = sc.omega
    lat.mask[:] = sc.mask
    lat.phi[:] = sc.phi
    return lat


def rel_err(a, b, fluid):
    """max |a-b| / max |b| over fluid nodes (norm-relative: u passes through 0)."""
    a = a[fluid]
    b = b[fluid]
    den = np.max(np.abs(b)) if b.size else 1.0
    if den == 0:
        den = 1.0
    return float(np.max(np.abs(a - b)) / den) if b.size else 0.0


def max_abs(a, b):
    d = np.abs(a - b)
    d = d[~np.isnan(d)]
    return float(d.max()) if d.size else 0.0


def same_bits(a, b):
    return np.array_equal(a, b, equal_nan=True)
