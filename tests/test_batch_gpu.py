"""GPU parity of the batched engine (include/lbmx_batch.h).

* FP64 batch == the oracle, bit for bit, for every case family of the batched
  workload (Poiseuille, Couette, lid-driven cavity, Taylor-Green) and the liner
  duct, with lattices of different sizes and lengths in one batch and a lattice
  joining mid-run.
* FP32-shifted batch == the FP32-shifted single-lattice path (lbmx.h), bit for bit.
* Results do not depend on the launch configuration, the HIP-graph chunk length
  or the batch's other members (tile-shape invariance), and graph replays equal
  eager launches.
"""
import numpy as np
import pytest

import oracle as O
from lbc_amd import batch as B
from lbc_amd import cases, native

pytestmark = pytest.mark.gpu


def small_impedance():
    return cases.LinerSpec(R=1.0, f_res=1 / 40, M=8.0).impedance()


def mixed_cases():
    imp = small_impedance()
    return [
        B.poiseuille(96, 34, 0.8, 400, Fx=1e-5),
        B.couette(64, 40, 0.7, 300),
        B.cavity(66, 50, 200.0, 350),
        B.taylor_green(64, 96, 250),
        B.liner_duct(384, 40, 400, imp, amp=1e-3, x0=64.0, halfwidth=8.0, liner=(128, 256),
                     sponge=24, probes=[(96, 20), (300, 20)]),
    ]


def oracle_moments(case, nsteps):
    lat = O.Lattice(case.Nx, case.Ny)
    rho, u, v = case.initial_fields()
    lat.f = cases.equilibrium(rho, u, v)
    lat.f_new = lat.f.copy()
    lat.rho, lat.u, lat.v = rho.copy(), u.copy(), v.copy()
    lat.omega = np.ascontiguousarray(np.broadcast_to(case.omega_column(), (case.Ny, case.Nx)))
    lat.mask = case.solid_mask()
    O.oracle_run(lat, nsteps, case.boundary_table(), case.per_x, case.per_y, case.Fx, case.Fy)
    return lat.rho, lat.u, lat.v


def oracle_liner(case, nsteps):
    import ctypes as C
    cfg = O.OrcLinerCfg()
    cfg.Nx, cfg.Ny = case.Nx, case.Ny
    cfg.liner_x0, cfg.liner_x1 = case.gap
    cfg.sponge = case.sponge
    cfg.probe_x, cfg.probe_y = case.probes[1]
    cfg.tau, cfg.tau_sponge, cfg.rho0 = case.tau, case.tau_sponge, case.rho0
    cfg.amp, cfg.x0, cfg.halfwidth, cfg.gain = case.amp, case.x0, case.halfwidth, case.gain
    cfg.imp.Zinf = case.liner["Zinf"]
    for n in ("A", "lambdas", "B", "C", "alpha", "beta"):
        arr = getattr(cfg.imp, n)
        arr[0], arr[1] = case.liner[n]
    probe = np.zeros(nsteps)
    rho = np.zeros((case.Ny, case.Nx))
    dp = C.POINTER(C.c_double)
    O.ORACLE.orc_liner_run(C.byref(cfg), nsteps, probe.ctypes.data_as(dp), None, None,
                           rho.ctypes.data_as(dp))
    return probe, rho


def same(a, b):
    return np.array_equal(a, b, equal_nan=True)


def test_fp64_batch_equals_oracle():
    cs = mixed_cases()
    b = B.Batch(native.F64, budget=1 << 30)
    ids = [b.add(c) for c in cs[:3]]
    b.run(100)
    ids += [b.add(c) for c in cs[3:]]          # join after 100 steps of the others
    b.run(400)
    b.synchronize()
    for i, c in zip(ids, cs):
        assert b.steps_done(i) == c.steps
        rho, u, v = b.moments(i)
        if c.kind == "liner":
            probe, rho_ref = oracle_liner(c, c.steps)
            assert same(rho, rho_ref), c.kind
            assert np.array_equal(b.probes(i)[:c.steps, 1], probe), c.kind
        else:
            r0, u0, v0 = oracle_moments(c, c.steps)
            assert same(rho, r0) and same(u, u0) and same(v, v0), c.kind
    b.close()


@pytest.mark.parametrize("precision", [native.F32_SHIFTED, native.F32_WIDE],
                         ids=["f32", "f32wide"])
def test_fp32_batch_equals_single_lattice(precision):
    cs = mixed_cases()
    b = B.Batch(precision, budget=1 << 30)
    ids = [b.add(c) for c in cs]
    b.run(400)
    b.synchronize()
    for i, c in zip(ids, cs):
        lat = c.single_lattice(precision, record_steps=c.steps + 1)
        lat.step(c.steps)
        rho, u, v = lat.moments()
        rb, ub, vb = b.moments(i)
        assert same(rb, rho) and same(ub, u) and same(vb, v), c.kind
        if c.probes:
            ph = lat.probe_history()
            assert np.array_equal(b.probes(i)[:c.steps], ph[:c.steps]), c.kind
        lat.close()
    b.close()


def run_batch(cs, precision, tuning=None, chunk=None, steps=None, extra=()):
    b = B.Batch(precision, budget=1 << 30, chunk=chunk)
    if tuning is not None:
        b.set_tuning(*tuning)
    for e in extra:
        b.add(e)
    ids = [b.add(c) for c in cs]
    b.run(steps or max(c.steps for c in cs))
    b.synchronize()
    out = [(b.moments(i), b.probes(i)) for i in ids]
    b.close()
    return out


def equal_runs(r1, r2):
    for (m1, p1), (m2, p2) in zip(r1, r2):
        for a, c in zip(m1, m2):
            if not same(a, c):
                return False
        if not np.array_equal(p1, p2):
            return False
    return True


@pytest.mark.parametrize("tuning,chunk", [((64, 256, 2, 0), 1), ((128, 1024, 16, 1), 7),
                                          ((256, 2048, 4, 2), 64), ((256, 512, 8, 3), 5),
                                          ((128, 256, 16, 4), 32), ((64, 256, 16, 5), 3),
                                          ((256, 1024, 16, 5), 9)])
def test_results_independent_of_launch_shape(tuning, chunk):
    cs = mixed_cases()
    ref = run_batch(cs, native.F32_SHIFTED)
    other = run_batch(cs, native.F32_SHIFTED, tuning=tuning, chunk=chunk,
                      extra=[B.taylor_green(160, 64, 123)])
    assert equal_runs(ref, other)


def test_autotune_keeps_results():
    cs = mixed_cases()
    ref = run_batch(cs, native.F32_SHIFTED)
    b = B.Batch(native.F32_SHIFTED, budget=1 << 30)
    ids = [b.add(c) for c in cs]
    t = b.tune(trial_steps=3)
    assert t == b.tuning() and t[1] % t[0] == 0
    b.run(400)
    b.synchronize()
    got = [(b.moments(i), b.probes(i)) for i in ids]
    b.close()
    assert equal_runs(ref, got)


# relative L2 velocity error against FP64 on the batched families: 1e-5 with FP32
# arithmetic, 1e-7 with FP64 arithmetic on the FP32-stored deviations
FP32_BATCH_L2 = {native.F32_SHIFTED: 1e-5, native.F32_WIDE: 1e-7}


@pytest.mark.parametrize("precision", [native.F32_SHIFTED, native.F32_WIDE],
                         ids=["f32", "f32wide"])
def test_fp32_batch_close_to_fp64(precision):
    cs = mixed_cases()[:4]
    r64 = run_batch(cs, native.F64)
    r32 = run_batch(cs, precision)
    for c, (m64, _), (m32, _) in zip(cs, r64, r32):
        fluid = ~np.isnan(m64[0])
        du = np.hypot(m32[1] - m64[1], m32[2] - m64[2])[fluid]
        ref = np.hypot(m64[1], m64[2])[fluid]
        err = np.sqrt((du ** 2).sum() / (ref ** 2).sum())
        print(f"{c.kind}: L2 velocity error (precision {precision}) vs FP64 = {err:.2e}")
        assert err < FP32_BATCH_L2[precision], c.kind


def test_remove_and_reuse_memory():
    cs = mixed_cases()
    ref = run_batch(cs[:2], native.F32_SHIFTED)
    b = B.Batch(native.F32_SHIFTED, budget=64 << 20)
    big = B.taylor_green(1024, 512, 10)
    assert B.case_bytes(big) > (32 << 20)
    t = b.add(big)
    b.run(10)
    b.synchronize()
    b.remove(t)
    with pytest.raises(native.LbmxError):
        b.add(B.taylor_green(2048, 1024, 10))       # over the budget
    ids = [b.add(c) for c in cs[:2]]                 # reuse the freed block
    b.run(400)
    b.synchronize()
    got = [(b.moments(i), b.probes(i)) for i in ids]
    b.close()
    assert equal_runs(ref, got)


def test_many_grid_changes_keep_results():
    """Lattices join and leave between chunks often enough that the one-tile-per-
    workgroup grid (and its captured graphs) is replaced more than 32 times without
    an explicit synchronisation; every lattice still ends bit-identical to its own
    single-lattice run."""
    rng = np.random.default_rng(7)
    b = B.Batch(native.F32_SHIFTED, budget=1 << 30, chunk=4)
    done = []
    live = {}
    for round_ in range(40):
        nx = int(rng.integers(2, 12)) * 32
        ny = int(rng.integers(2, 8)) * 32
        c = B.taylor_green(nx, ny, int(rng.integers(6, 30))) if round_ % 2 else \
            B.poiseuille(nx, ny, 0.8, int(rng.integers(6, 30)), Fx=1e-5)
        live[b.add(c)] = c
        b.run(5)                     # no lbmx_batch_synchronize: replaced graphs pile up
        for i, cc in list(live.items()):
            if b.steps_done(i) >= cc.steps:
                done.append((cc, b.moments(i)))
                b.remove(i)
                del live[i]
    b.run(40)
    b.synchronize()
    for i, cc in live.items():
        done.append((cc, b.moments(i)))
    b.close()
    assert len(done) == 40
    for c, (rb, ub, vb) in done:
        lat = c.single_lattice(native.F32_SHIFTED)
        lat.step(c.steps)
        rho, u, v = lat.moments()
        lat.close()
        assert same(rb, rho) and same(ub, u) and same(vb, v), c.kind
