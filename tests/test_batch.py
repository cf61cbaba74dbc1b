"""CPU checks of the batched engine's host side (no GPU calls).

* The case families, stated in the reference's terms (mask, omega, boundary
  table, initial fields), are the programs they claim to be: run by the oracle
  they reproduce the Poiseuille parabola, the Couette line and the Taylor-Green
  decay, and the liner duct builds exactly the arrays of the oracle's liner set-up.
* The batched workload generator draws what it states (sizes, families, seeds).
* lbmx_case_check enforces the rules under which the batched kernel's results
  equal the reference's (stale entries never copied, impedance opening inside
  the bottom wall, ...), and lbmx_case_bytes sizes the arena blocks.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O
from lbc_amd import batch as B
from lbc_amd import cases, native


def run_oracle(case, nsteps):
    lat = O.Lattice(case.Nx, case.Ny)
    rho, u, v = case.initial_fields()
    lat.f = cases.equilibrium(rho, u, v)
    lat.f_new = lat.f.copy()
    lat.rho, lat.u, lat.v = rho.copy(), u.copy(), v.copy()
    lat.omega = np.ascontiguousarray(np.broadcast_to(case.omega_column(), (case.Ny, case.Nx)))
    lat.mask = case.solid_mask()
    O.oracle_run(lat, nsteps, case.boundary_table(), case.per_x, case.per_y, case.Fx, case.Fy)
    return lat


def test_poiseuille_case_reaches_parabola():
    # half-way bounce-back walls sit half a cell outside the first fluid rows:
    # u(y) = F/(2 nu) (y - 0.5)(H + 0.5 - y), H = Ny - 2 fluid rows
    c = B.poiseuille(8, 18, 1.0, 4000, Fx=1e-6)
    lat = run_oracle(c, c.steps)
    nu = (c.tau - 0.5) / 3.0
    y = np.arange(1, c.Ny - 1, dtype=float)
    exact = c.Fx / (2 * nu) * (y - 0.5) * (c.Ny - 1.5 - y)
    got = lat.u[1:-1, 3]
    assert np.abs(got - exact).max() / exact.max() < 2e-3


def test_couette_case_reaches_line():
    c = B.couette(8, 17, 0.9, 6000, U=0.02)
    lat = run_oracle(c, c.steps)
    y = np.arange(1, c.Ny, dtype=float)
    got = lat.u[1:, 2]
    exact = 0.02 * (y - 0.5) / (c.Ny - 1 - 0.5)
    assert np.abs(got - exact).max() < 2e-3 * 0.02 * 10


def test_taylor_green_case_decays():
    c = B.taylor_green(32, 48, 200, tau=0.8, u0=0.01)
    lat = run_oracle(c, c.steps)
    nu = (0.8 - 0.5) / 3.0
    kx, ky = 2 * math.pi / 32, 2 * math.pi / 48
    rho0, u0, v0 = c.initial_fields()
    decay = math.exp(-nu * (kx * kx + ky * ky) * c.steps)
    ratio = np.abs(lat.u).max() / np.abs(u0).max()
    assert abs(ratio / decay - 1.0) < 0.02


def test_liner_case_matches_oracle_setup():
    imp = cases.LinerSpec().impedance()
    c = B.liner_duct(384, 40, 10, imp, amp=1e-3, x0=64.0, halfwidth=8.0, liner=(128, 256),
                     sponge=24)
    cfg = O.OrcLinerCfg()
    cfg.Nx, cfg.Ny, cfg.liner_x0, cfg.liner_x1, cfg.sponge = 384, 40, 128, 256, 24
    cfg.tau, cfg.tau_sponge, cfg.rho0, cfg.amp, cfg.x0, cfg.halfwidth = 0.51, 0.9, 1.0, 1e-3, 64.0, 8.0
    nn = 384 * 40
    f, rho, u, v, om = (np.zeros(nn * 9), np.zeros(nn), np.zeros(nn), np.zeros(nn), np.zeros(nn))
    mask = np.zeros(nn, dtype=np.int32)
    dp = C.POINTER(C.c_double)
    O.ORACLE.orc_liner_setup(C.byref(cfg), f.ctypes.data_as(dp), rho.ctypes.data_as(dp),
                             u.ctypes.data_as(dp), v.ctypes.data_as(dp), om.ctypes.data_as(dp),
                             mask.ctypes.data_as(C.POINTER(C.c_int)))
    r, uu, vv = c.initial_fields()
    assert np.array_equal(r.ravel(), rho)
    assert np.array_equal(c.solid_mask().ravel(), mask)
    assert np.array_equal(np.broadcast_to(c.omega_column(), (40, 384)).ravel(), om)
    assert np.array_equal(cases.equilibrium(r, uu, vv).ravel(), f)
    assert [t[0] for t in c.boundary_table()] == ["convective_inlet", "convective_outlet",
                                                  "velocity_bottom_regularized"]


def test_batched_workload_distribution():
    w = B.batched_workload(1024, steps=5000, seed=7)
    assert len(w) == 1024
    nx = np.array([c.Nx for c in w])
    ny = np.array([c.Ny for c in w])
    assert (nx % 32 == 0).all() and (ny % 32 == 0).all()
    assert nx.min() >= 64 and nx.max() <= 1024 and ny.min() >= 64 and ny.max() <= 512
    kinds = {k: sum(c.kind == k for c in w) for k in ("poiseuille", "couette", "cavity", "tgv")}
    assert all(150 < n < 370 for n in kinds.values()), kinds
    assert abs(nx.mean() - 544) < 30 and abs(ny.mean() - 288) < 20
    for c in w:
        if c.kind == "poiseuille":
            assert 0.6 <= c.tau <= 1.2 and c.Fx == 1e-6
        if c.kind == "cavity":
            Re = 0.1 * (c.Nx - 2) / ((c.tau - 0.5) / 3.0)
            assert 99.9 < Re < 1000.1
    w2 = B.batched_workload(1024, steps=5000, seed=7)
    assert [(c.Nx, c.Ny, c.kind, c.tau) for c in w] == [(c.Nx, c.Ny, c.kind, c.tau) for c in w2]


def test_case_check_rules():
    imp = cases.LinerSpec().impedance()
    ok = [B.poiseuille(64, 64, 0.8, 10), B.couette(64, 64, 0.8, 10), B.cavity(64, 64, 300, 10),
          B.taylor_green(64, 64, 10), B.liner_duct(1024, 64, 10, imp)]
    for c in ok:
        B.case_check(c)
    bad = []
    c = B.liner_duct(1024, 64, 10, imp)
    c.gap = (0, 500)                        # opening reaches the inlet column
    bad.append(c)
    c = B.liner_duct(1024, 64, 10, imp)
    c.walls = B.WALL_TOP                    # impedance wall without a bottom wall
    bad.append(c)
    c = B.liner_duct(1024, 64, 10, imp)
    c.walls = B.WALL_BOTTOM                 # fluid top corners next to convective ends
    bad.append(c)
    c = B.couette(64, 64, 0.8, 10)
    c.per_y = 1                             # velocity_top on a periodic axis
    bad.append(c)
    c = B.poiseuille(64, 64, 0.8, 10)
    c.probes = [(5, 0)]                     # probe on a wall
    bad.append(c)
    c = B.poiseuille(64, 64, 0.5, 10)       # tau must exceed 1/2
    bad.append(c)
    c = B.poiseuille(64, 64, 0.8, 0)        # no steps
    bad.append(c)
    c = B.couette(64, 64, 0.8, 10)
    c.convective_ends = True                # with a moving lid
    bad.append(c)
    # over-damped impedance whose fast real pole (-2.18) the explicit step amplifies
    from lbc_amd import serving as SV
    fast = SV.LinerRequest(640, 64, 2.2405399916753885, 0.02347702808796827,
                           1.0223512526325425, 1e-3, 3649)
    assert not cases.ade_stable(fast.impedance())
    bad.append(B.liner_duct(1024, 64, 10, fast.impedance()))
    for c in bad:
        with pytest.raises(native.LbmxError):
            B.case_check(c)


def test_case_bytes():
    c = B.taylor_green(544, 288, 5000)
    n = 544 * 288
    plane = ((n + 63) // 64) * 64 + 64
    pops32 = 2 * 9 * plane * 4
    # the frame buffer: four rows and columns along each edge, one slot per population
    frame32 = 9 * (8 * (544 + 288) + 128) * 4
    got = B.case_bytes(c, native.F32_SHIFTED)
    assert pops32 + frame32 // 2 <= got <= pops32 + frame32 + 3 * n * 8 + 64 * 1024
    c.keep_moments = False
    assert B.case_bytes(c, native.F32_SHIFTED) < pops32 + frame32 + 64 * 1024
    assert B.case_bytes(c, native.F64) > 2 * pops32 - 1024
