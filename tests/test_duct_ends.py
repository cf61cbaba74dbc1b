"""The liner duct's ends on the oracle: why convective ends inside a sponge.

The reference has no non-reflecting boundary condition: its regularized inlet
and outlet hold a fixed velocity or density, its convective outlet copies the
inner neighbour's populations (src/bc/convective_outlet.c:25-27), and
sponge_layer ramps tau toward the ends (src/utils).  Measured here (tests/
ductend.py) on the oracle, which is bit-identical to the reference's routines,
over f = 1/160 .. 1/25 cycles per step:
* pressure_outlet_regularized reflects almost fully at low frequency
  (|R| = 0.96 at 1/160, falling to 0.53 at 1/25);
* a convective outlet alone reflects a flat |R| of 0.23 .. 0.30;
* convective ends inside a 46-cell sponge (tau 0.51 -> 0.9): |R| 0.21 .. 0.24
  below 1/60, at most 0.12 from 1/50 up -- the served ducts use this, with
  sponges 3 Nx / 128 wide (tests/test_duct_ends_gpu.py measures those).
The transmission-loss windows end before any end reflection reaches a probe, so
the reflection shapes the run after the pulse, not the served spectrum.
"""
import numpy as np

import oracle as O
from lbc_amd import cases

import ductend as DE

NX, NY, SPONGE = 900, 48, 46
FREQS = np.linspace(1.0 / 160.0, 1.0 / 25.0, 8)


def probe_history(case, bcs=None):
    lat = O.Lattice(case.Nx, case.Ny)
    rho, u, v = case.initial_fields()
    lat.f = cases.equilibrium(rho, u, v)
    lat.f_new = lat.f.copy()
    lat.rho, lat.u, lat.v = rho.copy(), u.copy(), v.copy()
    lat.omega = np.ascontiguousarray(np.broadcast_to(case.omega_column(), (case.Ny, case.Nx)))
    lat.mask = case.solid_mask()
    table = case.boundary_table() if bcs is None else bcs
    px, py = case.probes[0]
    h = np.zeros(case.steps)
    for it in range(case.steps):
        h[it] = lat.rho[py, px]
        O.oracle_step(lat, table, case.per_x, case.per_y)
    return h


def test_duct_end_reflections():
    rigid, mid = DE.duct(NX, NY, SPONGE, "rigid")
    h_r = probe_history(rigid)
    R = {}
    for end in ("convective", "sponge"):
        c, _ = DE.duct(NX, NY, SPONGE, end)
        R[end] = DE.reflection(probe_history(c), h_r, mid, FREQS)
    c, _ = DE.duct(NX, NY, SPONGE, "open")
    out = [("pressure_outlet_regularized", NX - 1, np.full(NY, 1.0), None)]
    R["pressure"] = DE.reflection(probe_history(c, out), h_r, mid, FREQS)
    print({k: np.round(v, 3).tolist() for k, v in R.items()})
    assert R["pressure"][0] > 0.9 and np.all(R["pressure"] > 0.45)
    assert np.all((R["convective"] > 0.18) & (R["convective"] < 0.36))
    assert np.all(R["sponge"] < 0.3) and np.all(R["sponge"][FREQS >= 1.0 / 50.0] < 0.15)
    assert np.all(R["sponge"] < R["convective"])
