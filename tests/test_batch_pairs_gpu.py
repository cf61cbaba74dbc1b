"""GPU parity of the batched engine's pair mode (lbmx_batch_set_pairs: two steps
per pass, the core's intermediate step in LDS, the edge band through the frame
buffer).

Pair mode must give, bit for bit, the moments and probe histories of one step per
launch (which the single-step tests pin against the oracle in FP64):
* every case family and the liner duct, odd and even lengths, lattices too small
  to have a core (all band) or with a core of one column, periodic and walled;
  wide cores (side columns inside the core tiles: periodic, walled and
  convective sides) narrower than a tile, 65 and 66 columns wide, the smallest
  (9 x 9), and narrow cores (a side with stale entries: open ends, one wall);
* lattices joining the batch at later epochs, several graph chunk lengths;
* the FP64 liner duct against the oracle's liner driver directly;
* the divergence guard: a NaN lattice trips alone.
"""
import numpy as np
import pytest

from lbc_amd import batch as B
from lbc_amd import cases, native

import test_batch_gpu as TB

pytestmark = pytest.mark.gpu

PRECS = [native.F64, native.F32_WIDE, native.F32_SHIFTED]
IDS = ["f64", "f32wide", "f32"]


def zoo():
    imp = TB.small_impedance()
    return [
        B.poiseuille(96, 34, 0.8, 401, Fx=1e-5),
        B.couette(64, 40, 0.7, 300),
        B.cavity(66, 50, 200.0, 351),
        B.taylor_green(64, 96, 250),
        B.taylor_green(8, 6, 77),              # no core: every node in the band
        B.poiseuille(7, 9, 0.9, 53, Fx=2e-5),  # a core one column wide
        B.cavity(12, 11, 20.0, 64),            # frame buffer is the whole lattice
        B.couette(130, 19, 0.65, 129),
        B.liner_duct(384, 40, 401, imp, amp=1e-3, x0=64.0, halfwidth=8.0, liner=(128, 256),
                     sponge=24, probes=[(96, 20), (300, 20)]),
        B.liner_duct(257, 27, 300, imp, amp=5e-4, x0=40.0, halfwidth=6.0, liner=(90, 170),
                     sponge=12, probes=[(3, 3), (200, 13)]),
        # open, non-periodic ends (stale entries on both sides): a narrow core
        B.Case(96, 40, 201, tau=0.8, walls=B.WALL_BOTTOM | B.WALL_TOP, Fx=1e-5, kind="open"),
        # a wall on the left only, stale entries on the right: a narrow core
        B.Case(70, 33, 150, tau=0.7, walls=B.WALL_BOTTOM | B.WALL_LEFT,
               top_velocity=(0.02, 0.0), kind="leftwall"),
        # wide cores narrower than a tile, one column past a tile, a tile plus the ring
        B.couette(40, 30, 0.7, 120),
        B.poiseuille(65, 20, 0.9, 99, Fx=1e-5),
        B.taylor_green(66, 12, 90),
        B.cavity(9, 9, 10.0, 40),             # the smallest wide core
        # convective ends, last tile two columns wide (the 257-column duct above,
        # one column wide, keeps a narrow core)
        B.liner_duct(258, 27, 300, imp, amp=5e-4, x0=40.0, halfwidth=6.0, liner=(90, 170),
                     sponge=12, probes=[(3, 3), (257, 13)]),
    ]


def run(cases_, prec, pairs, chunk=32, schedule=None):
    """Runs every case to its end; schedule: [(epoch, [indices added then])]."""
    b = B.Batch(prec, budget=512 << 20, chunk=chunk)
    if pairs:
        b.set_pairs(True)
    schedule = schedule or [(0, list(range(len(cases_))))]
    ids = {}
    epoch = 0
    for at, idx in schedule:
        if at > epoch:
            b.run(at - epoch)
            epoch = at
        for i in idx:
            ids[i] = b.add(cases_[i])
    longest = max(cases_[i].steps + at for at, idx in schedule for i in idx)
    rest = longest - epoch
    b.run(rest + (rest & 1))
    b.synchronize()
    out = {}
    for i, j in ids.items():
        out[i] = (b.moments(j), b.probes(j), b.status(j))
    b.close()
    return out


def assert_same(a, b):
    for i in a:
        (ma, pa, ga), (mb, pb, gb) = a[i], b[i]
        for x, y in zip(ma, mb):
            assert TB.same(x, y), i
        assert TB.same(pa, pb), i
        assert ga == gb == 0, (i, ga, gb)


@pytest.mark.parametrize("prec", PRECS, ids=IDS)
def test_pairs_equal_single_steps(prec):
    cs = zoo()
    assert_same(run(cs, prec, True), run(cs, prec, False))


@pytest.mark.parametrize("prec", [native.F64, native.F32_WIDE], ids=["f64", "f32wide"])
def test_pairs_staggered_admission_and_chunks(prec):
    cs = zoo()
    sched = [(0, [0, 2, 4, 8, 10, 14]), (6, [1, 5, 11, 13]), (40, [3, 6, 7, 9, 12, 15])]
    ref = run(cs, prec, False, chunk=32, schedule=sched)
    for chunk in (2, 6, 32):
        assert_same(run(cs, prec, True, chunk=chunk, schedule=sched), ref)


def test_pairs_fp64_equal_oracle():
    """Open ends, a single side wall, the small wide cores and the wide duct
    against the oracle."""
    cs = zoo()[10:]
    got = run(cs, native.F64, True)
    for i, c in enumerate(cs):
        if c.kind == "liner":
            probe, rho_ref = TB.oracle_liner(c, c.steps)
            assert TB.same(got[i][0][0], rho_ref), c.kind
            assert np.array_equal(got[i][1][:c.steps, 1], probe), c.kind
            continue
        ref = TB.oracle_moments(c, c.steps)
        for x, y in zip(got[i][0], ref):
            assert TB.same(x, y), c.kind


def test_pairs_liner_fp64_equals_oracle():
    c = zoo()[8]
    b = B.Batch(native.F64, budget=64 << 20)
    b.set_pairs(True)
    i = b.add(c)
    b.run(c.steps + 1)
    b.synchronize()
    rho, _, _ = b.moments(i)
    hist = b.probes(i)
    b.close()
    probe, rho_ref = TB.oracle_liner(c, c.steps)
    assert TB.same(rho, rho_ref)
    assert np.array_equal(hist[:c.steps, 1], probe)


def test_pairs_mode_rules():
    b = B.Batch(native.F32_WIDE, budget=64 << 20)
    b.set_pairs(True)
    with pytest.raises(native.LbmxError):
        b.run(3)                        # pair mode advances in steps of two
    with pytest.raises(native.LbmxError):
        b.set_chunk(7)
    i = b.add(B.cavity(40, 40, 100.0, 10))
    b.run(2)
    with pytest.raises(native.LbmxError):
        b.set_pairs(False)              # a lattice has started
    b.run(8)
    b.synchronize()
    assert b.steps_done(i) == 10
    b.close()


def test_pairs_nan_lattice_fails_alone():
    steps = 40
    healthy = [B.poiseuille(96, 40, 0.8, steps, Fx=1e-5), B.taylor_green(64, 64, steps)]
    base = B.taylor_green(48, 32, steps)
    rho = np.ones((32, 48))
    rho[10, 20] = np.nan
    zero = np.zeros((32, 48))
    c = B.Case(**{k: getattr(base, k) for k in base.__dataclass_fields__})
    c.init = B.INIT_FIELDS
    b = B.Batch(native.F32_WIDE, budget=128 << 20)
    b.set_pairs(True)
    ids = [b.add(h) for h in healthy]
    bad = b.add(c, fields=(rho, zero, zero))
    b.run(steps)
    b.synchronize()
    assert b.status(bad) & B.GUARD_NAN
    mix = [b.moments(i) for i in ids]
    b.close()
    alone = run(healthy, native.F32_WIDE, False)
    for k, m in enumerate(mix):
        for x, y in zip(m, alone[k][0]):
            assert TB.same(x, y)
