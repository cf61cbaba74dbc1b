"""Several steps per launch on bounded lattices (the bounded gather form,
lbc_amd/csrc/hip/chan_kernels.h): walls and bodies with half-way or interpolated
bounce-back, periodic or closed axes, regularized inlet/outlet columns.

FP64 runs are checked bit for bit against the oracle (the reference's step,
cases/cylinder/cylinder.c's loop with constant boundary values), over calls whose
lengths start, continue and end launches part-way, with reads of the state in
between (which leave and re-enter the form).  Forces are summed in the one-step
path's fixed order, so the force record equals the one-step path's bit for bit
(and the oracle's within its own summation-order rounding).  Both FP32 storage
modes round exactly as the one-step path does, so they are compared with it bit
for bit.
"""
import numpy as np
import pytest

import oracle as O
from cases import Scenario
from parity import oracle_lattice, same_bits

pytestmark = pytest.mark.gpu


def two_omega(sc):
    sc.omega[:, sc.Nx // 2:] = 1.0 / 0.65
    return sc


CASES = [
    # the cylinder channel of cases/cylinder/cylinder.c with regularized ends
    Scenario("cylinder", 150, 48, [("velocity_inlet_regularized", 0),
                                   ("pressure_outlet_regularized", 149)], 41, use_ibb=1)
    .solid_rows(0, 47).disk(40.3, 23.6, 7.2),
    # (Nx - 1) is a multiple of the 58-column strip: the outlet starts a strip
    Scenario("cylinder_strip_edge", 117, 40, [("pressure_inlet_regularized", 0),
                                              ("velocity_outlet_regularized", 116)], 42,
             use_ibb=1).solid_rows(0, 39).disk(30.5, 19.2, 6.1),
    # plain bounce-back, two fluid classes (omega look-up table)
    two_omega(Scenario("cylinder_bb", 90, 35, [("velocity_inlet_regularized", 0),
                                               ("pressure_outlet_regularized", 89)], 45)
              .solid_rows(0, 34).disk(30.2, 17.1, 5.3)),
    # a different omega at every node (per-node omega array)
    Scenario("cylinder_random_omega", 70, 30, [("velocity_inlet_regularized", 0),
                                               ("pressure_outlet_regularized", 69)], 46,
             use_ibb=1).solid_rows(0, 29).disk(25.0, 14.5, 4.4).random_omega(),
    # Poiseuille: periodic x, walls, Guo body force
    Scenario("poiseuille", 130, 37, [("periodic_x", 0)], 43, Fx=2e-5).solid_rows(0, 36),
    # periodic both ways, one wall row, a body, interpolated bounce-back and force
    Scenario("periodic_walls_ibb", 96, 41, [("periodic_x", 0), ("periodic_y", 0)], 44,
             use_ibb=1, Fx=1e-5, Fy=-3e-6).solid_rows(0).disk(47.3, 20.4, 6.6),
    # narrower than one strip: both columns in one wavefront at every level
    Scenario("narrow", 7, 26, [("velocity_inlet_regularized", 0),
                               ("pressure_outlet_regularized", 6)], 47).solid_rows(0, 25),
    # no solid node: no links, a record of zero forces
    Scenario("open_channel", 80, 33, [("periodic_y", 0), ("velocity_inlet_regularized", 0),
                                      ("pressure_outlet_regularized", 79)], 48),
    # 37,196 wall links: 146 link blocks, so blocks b and b + 128 share a
    # partial-sum slot of the force reduction
    Scenario("long_walls", 6200, 20, [("velocity_inlet_regularized", 0),
                                      ("pressure_outlet_regularized", 6199)], 49)
    .solid_rows(0, 19),
]
IDS = [f"{s.name}_{s.Nx}x{s.Ny}" for s in CASES]
CHUNKS = [(13,), (6, 1, 9), (2, 5, 4)]


def where_differs(a, b):
    """First differing index and the values (for the failure message)."""
    d = ~((a == b) | (np.isnan(a) & np.isnan(b)))
    if not d.any():
        return "none"
    idx = tuple(int(i) for i in np.argwhere(d)[0])
    return f"{int(d.sum())} entries, first at {idx}: {a[idx]!r} vs {b[idx]!r}"


def consts(sc):
    return sc.vals(0)


def run_oracle(sc, state, nsteps):
    lat = oracle_lattice(sc, state)
    vals = consts(sc)
    forces = [O.oracle_step(lat, vals, sc.per_x, sc.per_y, sc.Fx, sc.Fy, sc.use_ibb)
              for _ in range(nsteps)]
    return lat, np.array(forces)


def run_gpu(sc, state, chunks, mode, precision=0, body=None):
    from lbc_amd import native
    lat = native.Lattice(sc.Nx, sc.Ny, precision=precision, per_x=sc.per_x, per_y=sc.per_y,
                         use_ibb=sc.use_ibb, Fx=sc.Fx, Fy=sc.Fy)
    lat.set_geometry(sc.mask, sc.omega, sc.phi)
    lat.set_boundaries(sc.bc_names)
    for b, (_, _, v1, v2) in enumerate(consts(sc)):
        lat.set_bc_values(b, v1, v2)
    lat.set_launch(mode)
    n = sum(chunks)
    lat.set_force_record(n, dt=0.5, cf=3.0)
    if body is not None:
        lat.set_force_body(body)
    lat.upload(*state)
    depth = lat.launch_depth()
    for c in chunks:
        lat.step(c)
        rho, u, v = lat.moments()
    out = {"depth": depth, "rec": lat.force_record(), "forces": lat.forces()}
    if body is not None:
        out["brec"] = lat.body_force_record()
    out["f"], out["f_new"] = lat.download(both=True)
    out["rho"], out["u"], out["v"] = rho, u, v
    lat.close()
    return out


@pytest.mark.parametrize("sc", CASES, ids=IDS)
@pytest.mark.parametrize("chunks", CHUNKS, ids=["-".join(map(str, c)) for c in CHUNKS])
def test_fp64_bounded_stream_matches_oracle(sc, chunks):
    from lbc_amd import native
    n = sum(chunks)
    state = sc.initial_state()
    ref, ref_forces = run_oracle(sc, state, n)
    got = run_gpu(sc, state, chunks, native.LAUNCH_STREAM)
    assert got["depth"] > 1, "the multi-step bounded launch was not chosen"
    for k in ("f", "f_new", "rho", "u", "v"):
        assert same_bits(got[k], getattr(ref, k)), \
            f"{sc.name} {chunks}: {k} differs: {where_differs(got[k], getattr(ref, k))}"
    # record rows {t, L, D}; the oracle sums the links in node order, the device
    # in link blocks: the totals agree to summation-order rounding, bounded here
    # by 1e-11 of the column's largest value (the walls' opposite forces cancel
    # in the lift of the 37,196-link channel)
    rec = got["rec"]
    assert rec.shape == (n, 3)
    np.testing.assert_array_equal(rec[:, 0], 0.5 * np.arange(n))
    for col, j in ((1, 0), (2, 1)):
        want = 3.0 * ref_forces[:, j]
        np.testing.assert_allclose(rec[:, col], want, rtol=1e-11,
                                   atol=1e-13 + 1e-11 * np.abs(want).max())


@pytest.mark.parametrize("precision", [0, 1, 2], ids=["f64", "f32", "f32wide"])
@pytest.mark.parametrize("sc", CASES, ids=IDS)
def test_bounded_stream_equals_one_step_path(sc, precision):
    """Populations, moments and both force records equal the one-step push path's
    bit for bit (same roundings in every storage mode)."""
    from lbc_amd import native
    body = (sc.mask != 0).astype(np.int32)
    body[0, :] = 0
    body[-1, :] = 0
    chunks = (6, 1, 9)
    state = sc.initial_state()
    a = run_gpu(sc, state, chunks, native.LAUNCH_STREAM, precision, body)
    b = run_gpu(sc, state, chunks, native.LAUNCH_SINGLE, precision, body)
    assert a["depth"] > 1 and b["depth"] == 1
    for k in ("f", "f_new", "rho", "u", "v", "rec", "brec"):
        assert same_bits(a[k], b[k]), \
            f"{sc.name} precision {precision}: {k} differs: {where_differs(a[k], b[k])}"
    assert a["forces"] == b["forces"]


def test_cylinder_channel_auto_rule_long_run():
    """The 2048 x 512 channel past a cylinder of the benchmark takes the bounded
    form under the default launch rule; 41 steps equal the one-step path's
    populations and force record bit for bit."""
    import math
    from lbc_amd import cases, native
    lat, _ = cases.cylinder_channel(Nx=2048, Ny=512)
    assert lat.launch_depth() > 1
    lat.set_force_record(41)
    lat.step(41)
    f, fn = lat.download(both=True)
    rec = lat.force_record()
    lat.close()
    one, _ = cases.cylinder_channel(Nx=2048, Ny=512)
    one.set_launch(native.LAUNCH_SINGLE)
    one.set_force_record(41)
    one.step(41)
    f1, fn1 = one.download(both=True)
    rec1 = one.force_record()
    one.close()
    assert same_bits(f, f1) and same_bits(fn, fn1)
    assert same_bits(rec, rec1)
    assert not math.isnan(rec[-1, 2])
