"""GPU parity of the native C ABI (include/lbmx.h) against the oracle.

Every scenario of the battery (all boundary types, bounce-back and interpolated
bounce-back, Guo forcing, periodic and non-periodic edges, stale entries,
corners, boundaries at interior indices) is run for several steps from the same
seeded state on the oracle and on the MI355X.  FP64 must agree to 1e-12
relative in rho and u; in practice the HIP path is bit-identical and the tests
also assert that.  FP32-shifted storage is checked
against the FP64 oracle with a stated tolerance.
"""
import numpy as np
import pytest

import oracle as O
from cases import Scenario, all_scenarios
from parity import FP64_RTOL, oracle_lattice, rel_err, same_bits

pytestmark = pytest.mark.gpu

SCENARIOS = all_scenarios(1) + [s for s in all_scenarios(3) if s.name in (
    "cylinder_ibb", "testbc_regularized", "nit_like", "late_periodic_overwrite")]
IDS = [f"{s.name}_{s.Nx}x{s.Ny}" for s in SCENARIOS]
NSTEPS = 12


def run_gpu(sc, state, nsteps, precision=0, exact=False):
    from lbc_amd import native
    lat = native.Lattice(sc.Nx, sc.Ny, precision=precision, per_x=sc.per_x, per_y=sc.per_y,
                         use_ibb=sc.use_ibb, Fx=sc.Fx, Fy=sc.Fy, exact=exact)
    lat.set_geometry(sc.mask, sc.omega, sc.phi)
    lat.set_boundaries(sc.bc_names)
    f, fn, rho, u, v = state
    lat.upload(f, fn, rho, u, v)
    forces = []
    for step in range(nsteps):
        for b, (_, _, v1, v2) in enumerate(sc.vals(step)):
            lat.set_bc_values(b, v1, v2)
        lat.step(1)
        forces.append(lat.forces())
    out = {"forces": np.array(forces)}
    if exact:
        out["f"], out["f_new"] = lat.download(both=True)
    else:
        out["f"] = lat.download()
    out["rho"], out["u"], out["v"] = lat.moments()
    lat.close()
    return out


def run_oracle(sc, state, nsteps):
    lat = oracle_lattice(sc, state)
    forces = []
    for step in range(nsteps):
        forces.append(O.oracle_step(lat, sc.vals(step), sc.per_x, sc.per_y, sc.Fx, sc.Fy,
                                    sc.use_ibb))
    return lat, np.array(forces)


@pytest.mark.parametrize("sc", SCENARIOS, ids=IDS)
@pytest.mark.parametrize("exact", [False, True], ids=["fast", "exact"])
def test_fp64_matches_oracle(sc, exact):
    state = sc.initial_state()
    ref, ref_forces = run_oracle(sc, state, NSTEPS)
    got = run_gpu(sc, state, NSTEPS, exact=exact)
    fluid = sc.mask == 0
    bitwise = {k: same_bits(got[k], getattr(ref, k)) for k in ("f", "rho", "u", "v")}
    print(f"{sc.name} exact={exact} bit-identical: {bitwise}")
    assert rel_err(got["rho"], ref.rho, fluid) <= FP64_RTOL
    assert rel_err(got["u"], ref.u, fluid) <= FP64_RTOL
    assert rel_err(got["v"], ref.v, fluid) <= FP64_RTOL
    assert np.all(np.isnan(got["rho"][~fluid]))
    # populations, solid nodes included (frozen), must agree everywhere
    scale = np.max(np.abs(ref.f))
    assert np.max(np.abs(got["f"] - ref.f)) <= FP64_RTOL * scale
    if exact:
        assert np.max(np.abs(got["f_new"] - ref.f_new)) <= FP64_RTOL * scale
    # forces: same links, different (fixed) summation order
    np.testing.assert_allclose(got["forces"], ref_forces, rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("sc", SCENARIOS[:16], ids=IDS[:16])
def test_fp64_bit_identical(sc):
    """Stronger than the tolerance: the FP64 step rounds exactly like the reference."""
    state = sc.initial_state()
    ref, _ = run_oracle(sc, state, 6)
    got = run_gpu(sc, state, 6)
    assert same_bits(got["f"], ref.f)
    assert same_bits(got["rho"], ref.rho)
    assert same_bits(got["u"], ref.u)
    assert same_bits(got["v"], ref.v)


GATHER_CASES = [
    Scenario("tgv_gather", 40, 24, [("periodic_x", 0), ("periodic_y", 0)], 31),
    Scenario("tgv_gather_force", 36, 28, [("periodic_x", 0), ("periodic_y", 0)], 32,
             Fx=1e-5, Fy=-2e-5),
    Scenario("tgv_gather_odd", 75, 19, [("periodic_y", 0), ("periodic_x", 0)], 33),
]


def _gather_run(sc, state, chunks, precision=0):
    """Runs the chunks of steps with single lbmx_step calls (>= 2 steps use the
    gather form), reading moments between chunks (which returns to the
    reference's buffers); returns (f, f_new, rho, u, v)."""
    from lbc_amd import native
    lat = native.Lattice(sc.Nx, sc.Ny, precision=precision, per_x=sc.per_x, per_y=sc.per_y,
                         Fx=sc.Fx, Fy=sc.Fy)
    lat.set_geometry(sc.mask, sc.omega, sc.phi)
    lat.set_boundaries(sc.bc_names)
    lat.upload(*state)
    for n in chunks:
        lat.step(n)
        rho, u, v = lat.moments()
    f, fn = lat.download(both=True)
    lat.close()
    return f, fn, rho, u, v


@pytest.mark.parametrize("sc", GATHER_CASES, ids=[s.name for s in GATHER_CASES])
@pytest.mark.parametrize("chunks", [(9,), (3, 1, 5), (1, 2, 2, 4)], ids=["9", "3-1-5", "1-2-2-4"])
def test_gather_form_bit_identical(sc, chunks):
    """Periodic all-fluid lattices run multi-step calls in gather form (buffers
    hold post-collision populations); the result, both reference buffers
    included, must equal the oracle bit for bit."""
    state = sc.initial_state()
    ref, _ = run_oracle(sc, state, sum(chunks))
    f, fn, rho, u, v = _gather_run(sc, state, chunks)
    assert same_bits(f, ref.f)
    assert same_bits(fn, ref.f_new)
    assert same_bits(rho, ref.rho) and same_bits(u, ref.u) and same_bits(v, ref.v)


LAUNCH_CASES = GATHER_CASES + [
    # several column strips and row segments, the last of each partial
    Scenario("tgv_strips", 300, 200, [("periodic_x", 0), ("periodic_y", 0)], 34,
             Fx=2e-6, Fy=1e-6),
    Scenario("tgv_thin", 3, 150, [("periodic_x", 0), ("periodic_y", 0)], 35),
    Scenario("tgv_flat", 130, 3, [("periodic_x", 0), ("periodic_y", 0)], 36),
    # strips that tile the width exactly (56 columns: FP64 five levels, FP32
    # storage with FP64 collision four; 48: FP32 eight levels), and a strip wider
    # than the lattice
    Scenario("tgv_exact_strips", 112, 41, [("periodic_x", 0), ("periodic_y", 0)], 37),
    Scenario("tgv_exact_strips32", 96, 9, [("periodic_x", 0), ("periodic_y", 0)], 38,
             Fx=-1e-6),
    Scenario("tgv_narrow", 17, 64, [("periodic_x", 0), ("periodic_y", 0)], 39),
    # the three-step tiles at their other heights: 4 rows (1104 tiles; the cases
    # above get 2), and 8 rows (1161 tiles), the last tile row partial in both
    Scenario("tgv_tiles4", 200, 1101, [("periodic_x", 0), ("periodic_y", 0)], 40, Fy=1e-6),
    Scenario("tgv_tiles8", 520, 1030, [("periodic_x", 0), ("periodic_y", 0)], 41),
]


def _launch_run(sc, state, chunks, mode, precision=0):
    from lbc_amd import native
    lat = native.Lattice(sc.Nx, sc.Ny, precision=precision, per_x=1, per_y=1,
                         Fx=sc.Fx, Fy=sc.Fy)
    lat.set_launch(mode)
    lat.set_geometry(sc.mask, sc.omega, sc.phi)
    lat.set_boundaries(sc.bc_names)
    lat.upload(*state)
    for n in chunks:
        lat.step(n)
    rho, u, v = lat.moments()
    f, fn = lat.download(both=True)
    lat.close()
    return f, fn, rho, u, v


@pytest.mark.parametrize("sc", LAUNCH_CASES, ids=[s.name for s in LAUNCH_CASES])
@pytest.mark.parametrize("mode", [1, 2, 3], ids=["single", "tiles", "stream"])
def test_launch_modes_bit_identical(sc, mode):
    """One step per launch, tiles (3 steps per launch) and row streaming (5 steps
    per launch, forced on these small lattices), over runs of 19 and 6 + 11 steps:
    every buffer and moment equals the oracle bit for bit."""
    state = sc.initial_state()
    ref, _ = run_oracle(sc, state, 17)
    got = _launch_run(sc, state, (6, 11), mode)
    for a, b in zip(got, (ref.f, ref.f_new, ref.rho, ref.u, ref.v)):
        assert same_bits(a, b)
    ref, _ = run_oracle(sc, state, 19)
    got = _launch_run(sc, state, (19,), mode)
    for a, b in zip(got, (ref.f, ref.f_new, ref.rho, ref.u, ref.v)):
        assert same_bits(a, b)


@pytest.mark.parametrize("sc", LAUNCH_CASES[-8:], ids=[s.name for s in LAUNCH_CASES[-8:]])
def test_fp64_stream_four_step_remainder(sc):
    """An FP64 run at five steps per launch whose last four steps go in one
    four-level launch (10 = entry + 5 + 4 steps; 4 + 9 = entry + 3 tiles, then 5 +
    4 in gather form) equals the oracle bit for bit."""
    from lbc_amd import native
    state = sc.initial_state()
    for chunks in ((10,), (4, 9)):
        ref, _ = run_oracle(sc, state, sum(chunks))
        got = _launch_run(sc, state, chunks, native.LAUNCH_STREAM)
        for a, b in zip(got, (ref.f, ref.f_new, ref.rho, ref.u, ref.v)):
            assert same_bits(a, b)


@pytest.mark.parametrize("sc", LAUNCH_CASES[-8:], ids=[s.name for s in LAUNCH_CASES[-8:]])
def test_fp32_stream_four_step_remainder(sc):
    """An FP32 run at eight steps per launch whose last four to seven steps start
    with one four-level launch (13 = entry + 8 + 4; 4 + 13 = entry + 3 tiles, then
    8 + 4 + 1) rounds exactly as one step per launch."""
    from lbc_amd import native
    state = sc.initial_state()
    for chunks in ((13,), (4, 13)):
        ref = _launch_run(sc, state, chunks, native.LAUNCH_SINGLE, native.F32_SHIFTED)
        got = _launch_run(sc, state, chunks, native.LAUNCH_STREAM, native.F32_SHIFTED)
        for a, b in zip(got, ref):
            assert same_bits(a, b)


@pytest.mark.parametrize("precision", [1, 2], ids=["f32", "f32wide"])
@pytest.mark.parametrize("sc", LAUNCH_CASES, ids=[s.name for s in LAUNCH_CASES])
def test_launch_modes_fp32_identical(sc, precision):
    """FP32 storage: tiles (3 steps per launch) and row streaming (8 steps with
    FP32 collision, 4 with FP64 collision) round exactly as one step per launch."""
    state = sc.initial_state()
    ref = _launch_run(sc, state, (25,), 1, precision)
    for mode in (2, 3):
        got = _launch_run(sc, state, (25,), mode, precision)
        for a, b in zip(got, ref):
            assert same_bits(a, b)


@pytest.mark.parametrize("Nx,Ny,plan", [
    # 17 strips x 57 segments of 16 rows, the last of each partial
    (900, 900, (4, 56, 16)),
    # 18 strips x 55 segments of 19 rows (16 rows would need a second round of
    # workgroups), the last of each partial
    (1000, 1030, (4, 56, 19)),
    # 32-row segments
    (1280, 1280, (4, 56, 32)),
    # a narrow lattice: one full 56-column strip and a partial 8-column one
    (64, 90000, (4, 56, 32))], ids=["16row", "19row", "32row", "narrow"])
def test_fp32_four_step_stream_identical(Nx, Ny, plan):
    """FP32 storage and collision at four steps per launch (the automatic choice
    below 6 M cells) rounds exactly as one step per launch, with a force, for the
    segment heights and strip shapes the automatic rule reaches."""
    from lbc_amd import native
    sc = Scenario(f"tgv_f32_stream4_{Nx}x{Ny}", Nx, Ny, [("periodic_x", 0), ("periodic_y", 0)],
                  42, Fx=1e-6, Fy=-2e-6)
    assert native.launch_plan(sc.Nx, sc.Ny, native.F32_SHIFTED) == plan
    state = sc.initial_state()
    ref = _launch_run(sc, state, (14,), native.LAUNCH_SINGLE, native.F32_SHIFTED)
    got = _launch_run(sc, state, (14,), native.LAUNCH_AUTO, native.F32_SHIFTED)
    for a, b in zip(got, ref):
        assert same_bits(a, b)


@pytest.mark.parametrize("Nx,Ny,plan", [
    # four steps: 16-, 19- and 32-row segments, the last strip and segment partial
    (900, 900, (4, 56, 16)), (1000, 1030, (4, 56, 19)), (1280, 1280, (4, 56, 32)),
    # five steps: 128-row segments, and a narrow lattice of one full 56-column strip
    # and a partial 8-column one
    (2100, 1900, (5, 56, 128)), (64, 90000, (5, 56, 128)),
    # a flat lattice: 215 strips (the last partial) x segments of 128, 128 and 44 rows
    (12000, 300, (5, 56, 128))],
    ids=["d4_16row", "d4_19row", "d4_32row", "d5_128row", "d5_narrow", "d5_flat"])
@pytest.mark.parametrize("force", [False, True], ids=["noforce", "force"])
def test_fp64_auto_stream_identical(Nx, Ny, plan, force):
    """FP64 row streaming at the depths and segment shapes the automatic rule
    reaches (four steps per launch on mid-size lattices, five on large ones), in the
    force-free and the forced kernel, rounds exactly as one step per launch."""
    from lbc_amd import native
    F = (1e-6, -2e-6) if force else (0.0, 0.0)
    sc = Scenario(f"tgv_f64_stream_{Nx}x{Ny}", Nx, Ny, [("periodic_x", 0), ("periodic_y", 0)],
                  43, Fx=F[0], Fy=F[1])
    assert native.launch_plan(sc.Nx, sc.Ny, native.F64) == plan
    state = sc.initial_state()
    ref = _launch_run(sc, state, (16,), native.LAUNCH_SINGLE)
    got = _launch_run(sc, state, (16,), native.LAUNCH_AUTO)
    for a, b in zip(got, ref):
        assert same_bits(a, b)


def test_launch_depth_rule():
    """LBMX_LAUNCH_AUTO streams rows once the lattice gives k_stream 200
    workgroups of 32 rows (400 with FP64 collision on FP32 storage; FP32 eight
    steps from 6 M cells), and runs the tile kernels on smaller lattices; small
    lattices with walls take one step per launch unless streaming is forced, when
    the bounded multi-step launch takes them."""
    from lbc_amd import native
    cases = [(2048, 2048, 0, 5), (256, 256, 0, 3), (1024, 1024, 0, 4), (2048, 2048, 1, 4),
             (1024, 1024, 1, 4), (256, 256, 1, 3), (2048, 2048, 2, 4), (768, 768, 0, 4),
             (576, 576, 0, 3), (704, 704, 0, 4), (768, 768, 2, 3), (896, 896, 2, 4),
             (768, 768, 1, 4), (576, 576, 1, 3),
             (3072, 3072, 1, 8)]
    for Nx, Ny, prec, depth in cases:
        lat = native.Lattice(Nx, Ny, precision=prec, per_x=1, per_y=1)
        lat.set_uniform(1.0 / 0.8)
        lat.set_boundaries([("periodic_x", 0), ("periodic_y", 0)])
        assert lat.launch_depth() == depth, (Nx, Ny, prec)
        lat.set_launch(native.LAUNCH_SINGLE)
        assert lat.launch_depth() == 1
        lat.set_launch(native.LAUNCH_STREAM)
        assert lat.launch_depth() == (8 if prec == 1 else 5 if prec == 0 else 4)
        lat.close()
    lat = native.Lattice(64, 32, precision=0, per_x=1, per_y=0)
    mask = np.zeros((32, 64), np.int32)
    mask[0, :] = mask[-1, :] = 1
    lat.set_geometry(mask, np.full((32, 64), 1.25), None)
    lat.set_boundaries([("periodic_x", 0)])
    assert lat.launch_depth() == 1
    lat.set_launch(native.LAUNCH_STREAM)
    assert lat.launch_depth() > 1
    lat.close()


@pytest.mark.parametrize("precision", [1, 2], ids=["f32", "f32wide"])
@pytest.mark.parametrize("sc", GATHER_CASES, ids=[s.name for s in GATHER_CASES])
def test_gather_form_fp32_equals_push(sc, precision):
    """FP32 storage (FP32 or FP64 arithmetic): gather form and push form round
    identically."""
    state = sc.initial_state()
    a = _gather_run(sc, state, (8,), precision=precision)
    b = _gather_run(sc, state, (1,) * 8, precision=precision)
    for x, y in zip(a, b):
        assert same_bits(x, y)


# FP32 storage against the FP64 oracle: relative L2 error of the velocity after
# NSTEPS.  Bars: 1e-5 with FP32 arithmetic (LBMX_F32_SHIFTED), 1e-6 with FP64
# arithmetic on the FP32-stored deviations (LBMX_F32_WIDE).
FP32_VEL_L2 = {1: 1e-5, 2: 1e-6}


@pytest.mark.parametrize("precision", [1, 2], ids=["f32", "f32wide"])
@pytest.mark.parametrize("sc", SCENARIOS[:16], ids=IDS[:16])
def test_fp32_shifted_close_to_fp64(sc, precision):
    state = sc.initial_state()
    ref, ref_forces = run_oracle(sc, state, NSTEPS)
    got = run_gpu(sc, state, NSTEPS, precision=precision)
    fluid = sc.mask == 0
    du = np.sqrt(np.sum((got["u"] - ref.u)[fluid] ** 2 + (got["v"] - ref.v)[fluid] ** 2))
    nu = np.sqrt(np.sum(ref.u[fluid] ** 2 + ref.v[fluid] ** 2))
    assert du / nu <= FP32_VEL_L2[precision], du / nu
    assert rel_err(got["rho"], ref.rho, fluid) <= 1e-6


FORCE_CASES = [s for s in SCENARIOS if s.use_ibb or s.mask.any()][:4]


@pytest.mark.parametrize("sc", FORCE_CASES, ids=[f"{s.name}_{s.Nx}x{s.Ny}" for s in FORCE_CASES])
def test_force_record_matches_oracle(sc):
    """The device force record (cylinder.c:162-185 layout {t, L*cf, D*cf}) equals
    the oracle's per-step (L, D), up to the fixed summation order."""
    from lbc_amd import native
    state = sc.initial_state()
    _, ref_forces = run_oracle(sc, state, NSTEPS)
    lat = native.Lattice(sc.Nx, sc.Ny, precision=0, per_x=sc.per_x, per_y=sc.per_y,
                         use_ibb=sc.use_ibb, Fx=sc.Fx, Fy=sc.Fy)
    lat.set_geometry(sc.mask, sc.omega, sc.phi)
    lat.set_boundaries(sc.bc_names)
    lat.upload(*state)
    dt, cf = 1.0 / 300.0, 2.5
    lat.set_force_record(NSTEPS, dt=dt, cf=cf)
    for step in range(NSTEPS):
        for b, (_, _, v1, v2) in enumerate(sc.vals(step)):
            lat.set_bc_values(b, v1, v2)
        lat.step(1)
    rec = lat.force_record()
    lat.close()
    assert rec.shape == (NSTEPS, 3)
    assert np.array_equal(rec[:, 0], np.arange(NSTEPS) * dt)
    np.testing.assert_allclose(rec[:, 1], ref_forces[:, 0] * cf, rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(rec[:, 2], ref_forces[:, 1] * cf, rtol=1e-11, atol=1e-12)


PUSH_LONG = [s for s in all_scenarios(2) if s.name in (
    "cylinder_ibb", "channel_regularized_ibb", "testbc_regularized", "inner_index_bcs",
    "zouhe_box", "ldc_walls")]


@pytest.mark.parametrize("sc", PUSH_LONG, ids=[s.name for s in PUSH_LONG])
def test_push_multistep_bit_identical(sc):
    """200 steps in one lbmx_step call on the push path (bulk and links in one
    launch, the boundary table and stale edges in planned launches), fixed
    boundary values: bit-identical to the oracle."""
    from lbc_amd import native
    nsteps = 200
    state = sc.initial_state()
    vals = sc.vals(0)
    lat0 = oracle_lattice(sc, state)
    O.oracle_run(lat0, nsteps, vals, sc.per_x, sc.per_y, sc.Fx, sc.Fy, sc.use_ibb)
    lat = native.Lattice(sc.Nx, sc.Ny, precision=0, per_x=sc.per_x, per_y=sc.per_y,
                         use_ibb=sc.use_ibb, Fx=sc.Fx, Fy=sc.Fy)
    lat.set_geometry(sc.mask, sc.omega, sc.phi)
    lat.set_boundaries(sc.bc_names)
    for b, (_, _, v1, v2) in enumerate(vals):
        lat.set_bc_values(b, v1, v2)
    lat.upload(*state)
    lat.step(nsteps)
    rho, u, v = lat.moments()
    f = lat.download()
    lat.close()
    assert same_bits(rho, lat0.rho) and same_bits(u, lat0.u) and same_bits(v, lat0.v)
    assert same_bits(f, lat0.f)


def tgv256_state(N=256, u0=0.01):
    """Taylor-Green vortex, k = 2 pi / N, x = i, y = j; rho' = p'/cs^2."""
    from cases import feq
    k = 2 * np.pi / N
    j, i = np.mgrid[0:N, 0:N].astype(np.float64)
    u = u0 * np.sin(k * i) * np.cos(k * j)
    v = -u0 * np.cos(k * i) * np.sin(k * j)
    rho = 1 - 3 * u0 ** 2 / 4 * (np.cos(2 * k * i) + np.cos(2 * k * j))
    f = feq(rho, u, v)
    return f, f.copy(), rho, u, v


def tgv256_gpu(nsteps):
    from lbc_amd import native
    N = 256
    lat = native.Lattice(N, N, precision=0, per_x=1, per_y=1)
    lat.set_uniform(1.0 / 0.8)
    lat.set_boundaries([("periodic_x", 0), ("periodic_y", 0)])
    lat.upload(*tgv256_state())
    lat.step(nsteps)
    out = lat.moments()
    lat.close()
    return out


def test_tgv256_bit_identical():
    """The reference's own Taylor-Green case (256 x 256, tau = 0.8, u0 = 0.01) for
    500 steps in one call (gather form): equal to the oracle bit for bit."""
    N = 256
    lat0 = O.Lattice(N, N)
    lat0.f, lat0.f_new, lat0.rho, lat0.u, lat0.v = tgv256_state()
    lat0.omega[:] = 1.0 / 0.8
    O.oracle_run(lat0, 500, [("periodic_x", 0, None, None), ("periodic_y", 0, None, None)], 1, 1)
    rho, u, v = tgv256_gpu(500)
    assert same_bits(rho, lat0.rho) and same_bits(u, lat0.u) and same_bits(v, lat0.v)


def test_tgv256_analytic_decay():
    """10,000 steps of the same case against u0 exp(-2 nu k^2 t): relative L2 velocity
    deviation below 3e-3 (the oracle's own deviation is 2.23e-3: O(k^2) lattice
    error plus the start from equilibrium populations)."""
    N, u0, nu, T = 256, 0.01, 0.1, 10000
    k = 2 * np.pi / N
    _, u, v = tgv256_gpu(T)
    j, i = np.mgrid[0:N, 0:N].astype(np.float64)
    d = np.exp(-2 * nu * k * k * T)
    ua, va = u0 * np.sin(k * i) * np.cos(k * j) * d, -u0 * np.cos(k * i) * np.sin(k * j) * d
    err = np.sqrt((((u - ua) ** 2 + (v - va) ** 2).sum()) / ((ua ** 2 + va ** 2).sum()))
    print(f"TGV 256^2, {T} steps: L2 deviation from the analytic vortex {err:.3e}")
    assert err < 3e-3


def oracle_body_force(lat0, sc, body):
    """compute_force_mem over the links into `body` only, in the reference's step
    order up to the force (collision in place, then (interpolated) bounce-back
    into f_new; main_lbm.c:31-43) on copies of the oracle lattice.  With
    body = the full solid mask this is the step's own (L, D)."""
    import ctypes as C
    Nx, Ny = lat0.Nx, lat0.Ny
    f, fn = lat0.f.copy(), lat0.f_new.copy()
    p = O._p
    O.ORACLE.orc_collision(Nx, Ny, p(f), p(lat0.rho), p(lat0.u), p(lat0.v), p(lat0.omega),
                           p(lat0.mask), sc.Fx, sc.Fy)
    if sc.use_ibb:
        O.ORACLE.orc_interpolated_bounce_back(Nx, Ny, p(f), p(fn), p(lat0.mask), p(lat0.phi),
                                              sc.per_x, sc.per_y)
    else:
        O.ORACLE.orc_bounce_back(Nx, Ny, p(f), p(fn), p(lat0.mask), sc.per_x, sc.per_y)
    fx, fy = C.c_double(), C.c_double()
    O.ORACLE.orc_force_mem(Nx, Ny, p(f), p(fn), p(np.ascontiguousarray(body, dtype=np.int32)),
                           sc.per_x, sc.per_y, C.byref(fx), C.byref(fy))
    return fy.value, fx.value


BODY_CASES = [s for s in all_scenarios(2) if s.name in ("channel_regularized_ibb", "cylinder_ibb")]


@pytest.mark.parametrize("sc", BODY_CASES, ids=[s.name for s in BODY_CASES])
def test_body_force_record(sc):
    """Drag and lift of the disk alone (walls excluded) per step, recorded on the
    device, against the oracle's momentum exchange over the disk's links; the
    total (walls included) is unchanged."""
    from lbc_amd import native
    body = (sc.mask != 0) & (sc.phi > -1e29)          # the disk, not the wall rows
    assert body.any()
    state = sc.initial_state()
    lat0 = oracle_lattice(sc, state)
    lat = native.Lattice(sc.Nx, sc.Ny, precision=0, per_x=sc.per_x, per_y=sc.per_y,
                         use_ibb=sc.use_ibb, Fx=sc.Fx, Fy=sc.Fy)
    lat.set_geometry(sc.mask, sc.omega, sc.phi)
    lat.set_boundaries(sc.bc_names)
    lat.upload(*state)
    lat.set_force_body(body)
    lat.set_force_record(NSTEPS, dt=0.5, cf=2.0)
    ref_body, ref_total = [], []
    for step in range(NSTEPS):
        vals = sc.vals(step)
        ref_body.append(oracle_body_force(lat0, sc, body))
        check_total = oracle_body_force(lat0, sc, sc.mask)
        ref_total.append(O.oracle_step(lat0, vals, sc.per_x, sc.per_y, sc.Fx, sc.Fy, sc.use_ibb))
        assert check_total == ref_total[-1]        # the procedure is the step's own
        for b, (_, _, v1, v2) in enumerate(vals):
            lat.set_bc_values(b, v1, v2)
        lat.step(1)
    rec, brec = lat.force_record(), lat.body_force_record()
    Lb, Db = lat.body_forces()
    lat.close()
    ref_body, ref_total = np.array(ref_body), np.array(ref_total)
    assert np.array_equal(brec[:, 0], rec[:, 0])
    np.testing.assert_allclose(rec[:, 1:], ref_total * 2.0, rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(brec[:, 1:], ref_body * 2.0, rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose([Lb, Db], ref_body[-1], rtol=1e-11, atol=1e-12)
    if (sc.phi < -1e29).any():                         # the walls do carry force
        assert np.abs(ref_body - ref_total).max() > 1e-6


@pytest.mark.parametrize("precision", [1, 2], ids=["f32", "f32wide"])
@pytest.mark.parametrize("mode", [1, 3], ids=["single", "stream"])
def test_fp32_storage_with_other_shift_density(precision, mode):
    """FP32 storage shifted by rho0 = 1.3 on a state of density ~1.3 (the FP32
    velocity then divides by rho0 as well): one step per launch and row streaming
    within the FP32 bars of the FP64 oracle."""
    from lbc_amd import native
    sc = LAUNCH_CASES[3]
    f, fn, rho, u, v = sc.initial_state()
    s = 1.3
    state = (f * s, fn * s, rho * s, u, v)
    ref, _ = run_oracle(sc, state, 17)
    lat = native.Lattice(sc.Nx, sc.Ny, precision=precision, per_x=1, per_y=1, Fx=sc.Fx,
                         Fy=sc.Fy, rho0=s)
    lat.set_launch(mode)
    lat.set_geometry(sc.mask, sc.omega, sc.phi)
    lat.set_boundaries(sc.bc_names)
    lat.upload(*state)
    lat.step(17)
    r, uu, vv = lat.moments()
    lat.close()
    fluid = sc.mask == 0
    du = np.sqrt(np.sum((uu - ref.u)[fluid] ** 2 + (vv - ref.v)[fluid] ** 2))
    nu = np.sqrt(np.sum(ref.u[fluid] ** 2 + ref.v[fluid] ** 2))
    assert du / nu <= FP32_VEL_L2[precision], du / nu
    assert rel_err(r, ref.rho, fluid) <= 1e-6


def test_force_total_is_deterministic():
    """The force total is a fixed-order reduction over the link partials: two runs
    of the same scenario give the same forces bit for bit, every step."""
    sc = FORCE_CASES[0]
    state = sc.initial_state()
    a = run_gpu(sc, state, NSTEPS)["forces"]
    b = run_gpu(sc, state, NSTEPS)["forces"]
    assert np.array_equal(a, b)



def test_set_launch_rejects_unknown_mode():
    from lbc_amd import native
    lat = native.Lattice(64, 32, per_x=1, per_y=1)
    try:
        with pytest.raises(native.LbmxError):
            lat.set_launch(7)
        with pytest.raises(native.LbmxError):
            lat.set_launch(-1)
        lat.set_launch(native.LAUNCH_STREAM)   # a valid mode still works after
    finally:
        lat.close()
