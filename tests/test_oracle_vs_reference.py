"""Pins the oracle: the C restatement must reproduce the real reference library
(compiled from its own sources into oracle/_ref by oracle/Makefile) bit for bit,
on every scenario of the battery, over several steps, in both population
buffers, the lagged moment arrays and the momentum-exchange force.

Also pins Mei's initialization (mei_initialization.c) and sponge_layer.c.
"""
import numpy as np
import pytest

import oracle as O
from cases import all_scenarios
from parity import oracle_lattice, same_bits

SCENARIOS = all_scenarios(1)


@pytest.mark.parametrize("sc", SCENARIOS, ids=[s.name for s in SCENARIOS])
def test_step_bit_identical(sc, ref_lib):
    state = sc.initial_state()
    a = oracle_lattice(sc, state)
    b = oracle_lattice(sc, state)
    for step in range(8):
        bcs = sc.vals(step)
        La, Da = O.oracle_step(a, bcs, sc.per_x, sc.per_y, sc.Fx, sc.Fy, sc.use_ibb)
        Lb, Db = ref_lib.step(b, bcs, sc.per_x, sc.per_y, sc.Fx, sc.Fy, sc.use_ibb)
        for k in ("f", "f_new", "rho", "u", "v"):
            assert same_bits(getattr(a, k), getattr(b, k)), (sc.name, step, k)
        # the reference sums the force with OpenMP atomics in thread order; wall
        # contributions cancel, so the bound is absolute (per-link terms are O(1))
        assert Da == pytest.approx(Db, rel=1e-12, abs=1e-12)
        assert La == pytest.approx(Lb, rel=1e-12, abs=1e-12)


@pytest.mark.parametrize("sc", SCENARIOS[:10], ids=[s.name for s in SCENARIOS[:10]])
@pytest.mark.parametrize("iters,tol", [(5, 1e-30), (100, 1e-9)])
def test_mei_bit_identical(sc, iters, tol, ref_lib):
    f, fn, rho, u, v = sc.initial_state()
    a = oracle_lattice(sc, (f, fn, rho, u, v))
    b = oracle_lattice(sc, (f, fn, rho, u, v))
    O.oracle_mei(a, rho, u, v, sc.per_x, sc.per_y, iters, tol)
    ref_lib.mei(b, rho, u, v, sc.per_x, sc.per_y, iters, tol)
    for k in ("f", "f_new", "rho", "u", "v"):
        assert same_bits(getattr(a, k), getattr(b, k)), (sc.name, k)


def test_sponge_layer_bit_identical(ref_lib):
    import ctypes as C
    Nx, Ny = 40, 30
    a = np.zeros((Ny, Nx))
    b = np.zeros((Ny, Nx))
    args = (Nx, Ny, 0.5, 9.7, 7.2, 3.0, 12.0, 0.6, 3.0, 0.4)
    O.ORACLE.orc_sponge_layer(*args, a.ctypes.data_as(C.POINTER(C.c_double)))
    ref_lib.lib.sponge_layer(*args, b.ctypes.data_as(C.POINTER(C.c_double)))
    assert same_bits(a, b)
