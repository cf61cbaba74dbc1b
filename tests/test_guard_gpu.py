"""GPU checks of the divergence guard and the FP64 totals.

* A batch holding lattices whose initial fields carry a NaN at one node, or a
  negative density or a speed above 0.2 on a 5 x 5 block: each of those trips its
  own guard bit (LBMX_GUARD_NAN / _RHO / _SPEED; a block's edges may add the speed
  bit), the word does not depend on the launch shape, and the healthy lattices
  beside them end bit for bit as they do alone.
* A served request with a NaN pulse fails alone with LBMX_ERR_DIVERGED; every
  other request's spectrum and totals equal those of the same case run alone in a
  batch, bit for bit.
* lbmx_totals / lbmx_batch_totals equal the same sums formed on the host from the
  downloaded moments (relative 1e-13 for the mass), repeat bit for bit, conserve
  mass on a periodic lattice and follow the vortex's energy decay.
"""
import math
import time

import numpy as np
import pytest

from lbc_amd import batch as B
from lbc_amd import native
from lbc_amd import serving as SV

pytestmark = pytest.mark.gpu


def fields_case(base, rho, u, v):
    c = B.Case(**{k: getattr(base, k) for k in base.__dataclass_fields__})
    c.init = B.INIT_FIELDS
    c.keep_moments = True
    return c, (rho, u, v)


def run_alone(c, prec, steps, fields=None):
    b = B.Batch(prec, budget=B.case_bytes(c, prec) + (16 << 20))
    i = b.add(c, fields=fields)
    b.run(steps)
    b.synchronize()
    out = b.moments(i)
    g = b.status(i)
    b.close()
    return out, g


@pytest.mark.parametrize("prec", [native.F64, native.F32_WIDE], ids=["f64", "f32wide"])
def test_bad_lattices_fail_alone(prec):
    steps = 60
    healthy = [B.poiseuille(96, 40, 0.8, steps, Fx=1e-5), B.cavity(64, 48, 300.0, steps),
               B.taylor_green(64, 64, steps)]
    for c in healthy:
        c.keep_moments = True
    Ny, Nx = 32, 48
    rho = np.ones((Ny, Nx))
    u = np.zeros((Ny, Nx))
    v = np.zeros((Ny, Nx))
    base = B.taylor_green(Nx, Ny, steps)
    bad = {}
    r1 = rho.copy()
    r1[10, 20] = np.nan
    bad[B.GUARD_NAN] = fields_case(base, r1, u, v)
    # blocks of 5 x 5 nodes, so that the streamed moments at their centre keep the
    # planted value
    r2 = rho.copy()
    r2[5:10, 7:12] = -0.5
    bad[B.GUARD_RHO] = fields_case(base, r2, u, v)
    u3 = u.copy()
    u3[12:17, 28:33] = 0.35
    bad[B.GUARD_SPEED] = fields_case(base, rho, u3, v)

    b = B.Batch(prec, budget=256 << 20)
    ids = [b.add(c) for c in healthy[:2]]
    bad_ids = {g: b.add(c, fields=f) for g, (c, f) in bad.items()}
    ids.append(b.add(healthy[2]))
    b.run(steps)
    b.synchronize()
    words = {g: b.status(i) for g, i in bad_ids.items()}
    for g, w in words.items():
        assert w & g, (g, w)
    assert words[B.GUARD_NAN] == B.GUARD_NAN
    # the same batch under another launch shape records the same words
    b2 = B.Batch(prec, budget=256 << 20)
    b2.set_tuning(64, 256, 16, 0)
    ids2 = {g: b2.add(c, fields=f) for g, (c, f) in bad.items()}
    b2.run(steps)
    b2.synchronize()
    assert {g: b2.status(i) for g, i in ids2.items()} == words
    b2.close()
    for c, i in zip(healthy, ids):
        assert b.status(i) == 0
        mix = b.moments(i)
        alone, g = run_alone(c, prec, steps)
        assert g == 0
        for a, s in zip(mix, alone):
            assert np.array_equal(a, s, equal_nan=True), c.kind
    b.close()


def wait_all(server, n, timeout=120.0):
    got = {}
    t0 = time.time()
    while len(got) < n:
        for r in server.poll():
            got[r.tag] = r
        assert time.time() - t0 < timeout, f"only {len(got)} of {n} results"
        time.sleep(0.002)
    return got


def arrays(res):
    n = res.nfreq
    return (np.array(res.tl_db[:n]), np.array(res.up_abs[:n]), np.array(res.down_abs[:n]),
            np.array([res.mass, res.energy]))


def solo(r, prec):
    c = r.case(keep_moments=True)
    b = B.Batch(prec, budget=B.case_bytes(c, prec) + (64 << 20))
    j = b.add(c)
    b.run(r.steps)
    b.synchronize()
    res = b.spectrum(j, r.tl_spec())
    res.mass, res.energy = b.totals(j)
    b.close()
    return res


def test_served_nan_request_fails_alone():
    prec = native.F32_WIDE
    rs = SV.draw_requests(4, seed=12, Nx=(512, 1024), Ny=(64, 96), steps=(1400, 2500))
    x = rs[1]
    nan_req = SV.LinerRequest(x.Nx, x.Ny, x.R, x.f_res, x.M, float("nan"), x.steps)
    s = B.Server(prec, budget=2 << 30, chunk=32)
    for i, r in enumerate(rs[:2]):
        s.submit(r.case(), r.tl_spec(), i)
    s.submit(nan_req.case(), nan_req.tl_spec(), 99)
    for i, r in enumerate(rs[2:], start=2):
        time.sleep(0.003)
        s.submit(r.case(), r.tl_spec(), i)
    got = wait_all(s, len(rs) + 1)
    st = s.stats()
    s.close()
    assert got[99].status == B.ERR_DIVERGED and got[99].guard & B.GUARD_NAN
    assert st["running"] == 0 and st["waiting"] == 0
    for i, r in enumerate(rs):
        assert got[i].status == 0 and got[i].guard == 0
        for a, c in zip(arrays(got[i]), arrays(solo(r, prec))):
            assert np.array_equal(a, c), i
        assert got[i].mass > 0.9 * r.Nx * (r.Ny - 2)


def host_totals(rho, u, v):
    fl = ~np.isnan(rho)
    return float(rho[fl].sum()), float((0.5 * rho * (u * u + v * v))[fl].sum())


@pytest.mark.parametrize("prec", [native.F64, native.F32_SHIFTED], ids=["f64", "f32"])
def test_lattice_totals(prec):
    from lbc_amd import cases
    lat, _ = cases.cylinder_channel(512, 128, 10.0, 128.0, 64.0, 0.05, 100.0, prec)
    lat.step(200)
    m, e = lat.totals()
    assert (m, e) == lat.totals()              # fixed order: repeats bit for bit
    rho, u, v = lat.moments()
    hm, he = host_totals(rho, u, v)
    lat.close()
    assert abs(m - hm) <= 1e-13 * hm and abs(e - he) <= 1e-12 * he, (m, hm, e, he)
    assert e > 0.0
    # periodic lattice: mass is conserved across the gather-form launches
    N = 384
    lat = native.Lattice(N, N, precision=prec, per_x=1, per_y=1)
    lat.set_uniform(1.0 / 0.8)
    lat.set_boundaries([("periodic_x", 0), ("periodic_y", 0)])
    lat.init_taylor_green(0.01, 2 * math.pi / N, 2 * math.pi / N)
    m0, e0 = lat.totals()
    lat.step(77)
    m1, e1 = lat.totals()
    lat.close()
    bar = 1e-12 if prec == native.F64 else 1e-7
    assert abs(m1 - m0) <= bar * m0, (m0, m1)
    nu = (0.8 - 0.5) / 3.0
    k = 2 * math.pi / N
    # the vortex's energy decays as exp(-4 nu k^2 t)
    assert abs(e1 / e0 - math.exp(-4 * nu * k * k * 77)) < 1e-3


def test_batch_totals_match_host_sums():
    c = B.cavity(96, 64, 400.0, 300)
    c.keep_moments = True
    b = B.Batch(native.F64, budget=64 << 20)
    i = b.add(c)
    b.run(c.steps)
    b.synchronize()
    m, e = b.totals(i)
    assert (m, e) == b.totals(i)
    rho, u, v = b.moments(i)
    b.close()
    hm, he = host_totals(rho, u, v)
    assert abs(m - hm) <= 1e-13 * hm and abs(e - he) <= 1e-12 * he
