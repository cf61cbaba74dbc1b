"""GPU parity of the drop-in C API (the reference's own entry points, as
liblbm_hip.so exports them) against the oracle.

Drivers call main_lbm once per step and swap f/f_new (cases/nit/nit.c:306-311);
these tests do exactly that through ctypes with Boundary tables built from the
library's own BC function pointers, and compare every step's rho/u/v and forces
and, at the end, both population buffers with the oracle's.  They also cover the
stand-alone functions (collision, streaming, bounce-back, IBB, force, moments,
every BC called directly), Mei's initialization, a broken swap protocol and a
user-written BC plug-in (a Python callback), which runs on the staged path.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from cases import Scenario, all_scenarios
from parity import FP64_RTOL, oracle_lattice, rel_err, same_bits

pytestmark = pytest.mark.gpu

SCENARIOS = all_scenarios(1)
IDS = [s.name for s in SCENARIOS]


@pytest.fixture(scope="module")
def dropin():
    from lbc_amd.capi import DropIn
    d = DropIn()
    yield d
    d.release()


def _host(sc, state):
    f, fn, rho, u, v = (np.array(a, copy=True) for a in state)
    return dict(f=f, f_new=fn, rho=rho, u=u, v=v, omega=sc.omega.copy(),
                mask=sc.mask.copy(), phi=sc.phi.copy())


@pytest.mark.parametrize("sc", SCENARIOS, ids=IDS)
def test_main_lbm_sequence(sc, dropin):
    state = sc.initial_state()
    ref = oracle_lattice(sc, state)
    h = _host(sc, state)
    fluid = sc.mask == 0
    for step in range(10):
        bcs = sc.vals(step)
        Lr, Dr = O.oracle_step(ref, bcs, sc.per_x, sc.per_y, sc.Fx, sc.Fy, sc.use_ibb)
        Lg, Dg = dropin.main_lbm(h["f"], h["f_new"], h["rho"], h["u"], h["v"], h["omega"],
                                 h["mask"], bcs, sc.per_x, sc.per_y, sc.Fx, sc.Fy, h["phi"],
                                 sc.use_ibb)
        h["f"], h["f_new"] = h["f_new"], h["f"]          # the driver's swap
        for k in ("rho", "u", "v"):
            assert rel_err(h[k], getattr(ref, k), fluid) <= FP64_RTOL, (step, k)
            assert same_bits(h[k], getattr(ref, k)), (step, k)
        assert Dg == pytest.approx(Dr, rel=1e-11, abs=1e-12)
        assert Lg == pytest.approx(Lr, rel=1e-11, abs=1e-12)
    dropin.sync()    # device state -> the caller's buffers
    assert same_bits(h["f"], ref.f)
    assert same_bits(h["f_new"], ref.f_new)


def test_broken_swap_protocol(dropin):
    """A caller that does not swap gets the reference's result anyway."""
    sc = [s for s in SCENARIOS if s.name == "channel_regularized_ibb"][0]
    state = sc.initial_state()
    ref = oracle_lattice(sc, state)
    h = _host(sc, state)
    for step in range(6):
        bcs = sc.vals(step)
        O.oracle_step(ref, bcs, sc.per_x, sc.per_y, sc.Fx, sc.Fy, sc.use_ibb)
        ref.f, ref.f_new = ref.f_new, ref.f             # undo the oracle's swap
        dropin.main_lbm(h["f"], h["f_new"], h["rho"], h["u"], h["v"], h["omega"], h["mask"],
                        bcs, sc.per_x, sc.per_y, sc.Fx, sc.Fy, h["phi"], sc.use_ibb)
    dropin.sync()
    assert same_bits(h["f"], ref.f)
    assert same_bits(h["f_new"], ref.f_new)
    assert same_bits(h["rho"], ref.rho)


BCFUNC = C.CFUNCTYPE(None, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double),
                     C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                     C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                     C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                     C.POINTER(C.c_double))


def test_user_plugin_boundary(dropin):
    """A BCFunction that is not the library's own (a user plug-in written in
    Python) is called on current host arrays; results equal the oracle running
    the built-in convective outlet it imitates."""
    sc = Scenario("plugin", 30, 14, [("periodic_y", 0), ("velocity_inlet", 0),
                                     ("convective_outlet", 29)], seed=21, use_ibb=1)
    sc.disk(9.2, 6.8, 2.7)
    calls = []

    def my_outlet(Nx, Ny, Q, index, val1, val2, f, f_new, rho, u, v, mask, cx, cy, w):
        fn = np.ctypeslib.as_array(f_new, shape=(Ny, Nx, Q))
        m = np.ctypeslib.as_array(mask, shape=(Ny, Nx))
        rows = m[:, index] == 0
        fn[rows, index, :] = fn[rows, index - 1, :]
        calls.append(index)

    cb = BCFUNC(my_outlet)
    state = sc.initial_state()
    ref = oracle_lattice(sc, state)
    h = _host(sc, state)
    for step in range(5):
        bcs = sc.vals(step)
        O.oracle_step(ref, bcs, sc.per_x, sc.per_y, sc.Fx, sc.Fy, sc.use_ibb)
        tbl = dropin.table(bcs)
        tbl[2].apply = C.cast(cb, C.c_void_p).value
        dropin.main_lbm(h["f"], h["f_new"], h["rho"], h["u"], h["v"], h["omega"], h["mask"],
                        bcs, sc.per_x, sc.per_y, sc.Fx, sc.Fy, h["phi"], sc.use_ibb, table=tbl)
        h["f"], h["f_new"] = h["f_new"], h["f"]
    assert calls == [29] * 5
    assert same_bits(h["f"], ref.f)
    assert same_bits(h["f_new"], ref.f_new)
    assert same_bits(h["rho"], ref.rho)


@pytest.mark.parametrize("sc", SCENARIOS[:12], ids=IDS[:12])
def test_standalone_stages(sc, dropin):
    """collision, bounce_back / interpolated_bounce_back, compute_force_mem,
    streaming, each BC and compute_macroscopic_fields called one by one, as
    main_lbm.c:33-61 calls them, equal the oracle's stage functions."""
    P = C.POINTER(C.c_double)
    I = C.POINTER(C.c_int)
    state = sc.initial_state()
    ref = oracle_lattice(sc, state)
    h = _host(sc, state)
    L = dropin.lib
    d = lambda a: a.ctypes.data_as(P)   # noqa: E731
    i = lambda a: a.ctypes.data_as(I)   # noqa: E731
    Nx, Ny = sc.Nx, sc.Ny
    cx, cy, w, opp = dropin.cx, dropin.cy, dropin.w, dropin.opp
    bcs = sc.vals(0)
    # oracle stages
    OR = O.ORACLE
    OR.orc_collision(Nx, Ny, d(ref.f), d(ref.rho), d(ref.u), d(ref.v), d(ref.omega), i(ref.mask),
                     sc.Fx, sc.Fy)
    if sc.use_ibb:
        OR.orc_interpolated_bounce_back(Nx, Ny, d(ref.f), d(ref.f_new), i(ref.mask), d(ref.phi),
                                        sc.per_x, sc.per_y)
    else:
        OR.orc_bounce_back(Nx, Ny, d(ref.f), d(ref.f_new), i(ref.mask), sc.per_x, sc.per_y)
    fx_r, fy_r = C.c_double(), C.c_double()
    OR.orc_force_mem(Nx, Ny, d(ref.f), d(ref.f_new), i(ref.mask), sc.per_x, sc.per_y,
                     C.byref(fx_r), C.byref(fy_r))
    OR.orc_streaming(Nx, Ny, d(ref.f), d(ref.f_new), i(ref.mask), sc.per_x, sc.per_y)
    tbl = O.make_bcs(bcs)
    for b in range(len(bcs)):
        OR.orc_apply_bc(Nx, Ny, C.byref(tbl[b]), d(ref.f), d(ref.f_new), d(ref.rho), d(ref.u),
                        d(ref.v), i(ref.mask))
    OR.orc_moments(Nx, Ny, d(ref.f_new), i(ref.mask), d(ref.rho), d(ref.u), d(ref.v), sc.Fx,
                   sc.Fy)
    # the same sequence through the drop-in's stand-alone functions
    f, fn = h["f"], h["f_new"]
    L.collision(Nx, Ny, 9, d(f), d(h["rho"]), d(h["u"]), d(h["v"]), d(h["omega"]),
                i(h["mask"]), i(cx), i(cy), d(w), sc.Fx, sc.Fy)
    assert same_bits(f, ref.f)
    if sc.use_ibb:
        L.interpolated_bounce_back(Nx, Ny, 9, d(f), d(fn), i(h["mask"]), d(h["phi"]), i(opp),
                                   i(cx), i(cy), sc.per_x, sc.per_y)
    else:
        L.bounce_back(Nx, Ny, 9, d(f), d(fn), i(h["mask"]), i(cx), i(cy), i(opp), sc.per_x,
                      sc.per_y)
    fx_g, fy_g = C.c_double(), C.c_double()
    L.compute_force_mem(Nx, Ny, 9, d(f), d(fn), i(h["mask"]), sc.per_x, sc.per_y, i(cx),
                        i(cy), i(opp), C.byref(fx_g), C.byref(fy_g))
    assert fx_g.value == pytest.approx(fx_r.value, rel=1e-11, abs=1e-12)
    assert fy_g.value == pytest.approx(fy_r.value, rel=1e-11, abs=1e-12)
    L.streaming(Nx, Ny, 9, d(f), d(fn), i(h["mask"]), i(cx), i(cy), sc.per_x, sc.per_y)
    gtbl = dropin.table(bcs)
    for b in range(len(bcs)):
        fnc = C.cast(gtbl[b].apply, BCFUNC)
        fnc(Nx, Ny, 9, gtbl[b].index, gtbl[b].val1, gtbl[b].val2, d(f), d(fn), d(h["rho"]),
            d(h["u"]), d(h["v"]), i(h["mask"]), i(cx), i(cy), d(w))
    assert same_bits(fn, ref.f_new)
    L.compute_macroscopic_fields(Nx, Ny, 9, d(fn), i(h["mask"]), i(cx), i(cy), d(h["rho"]),
                                 d(h["u"]), d(h["v"]), sc.Fx, sc.Fy)
    for k in ("rho", "u", "v"):
        assert same_bits(h[k], getattr(ref, k)), k


@pytest.mark.parametrize("sc", SCENARIOS[:9], ids=IDS[:9])
@pytest.mark.parametrize("iters,tol", [(6, 1e-30), (100, 1e-10)])
def test_mei_initialization(sc, iters, tol, dropin):
    f, fn, rho, u, v = sc.initial_state()
    ref = oracle_lattice(sc, (f, fn, rho, u, v))
    n_ref, _ = O.oracle_mei(ref, rho, u, v, sc.per_x, sc.per_y, iters, tol)
    h = _host(sc, (f, fn, rho, u, v))
    dropin.mei_initialization(h["f"], h["f_new"], rho.copy(), u.copy(), v.copy(), h["rho"],
                              h["u"], h["v"], h["omega"], sc.per_x, sc.per_y, iters, tol,
                              h["mask"])
    for k in ("f", "f_new", "rho", "u", "v"):
        assert same_bits(h[k], getattr(ref, k)), k
