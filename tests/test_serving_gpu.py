"""GPU checks of the request server (continuous batching, lbmx_server).

* A request's result is the same, bit for bit, alone and in a mix of requests
  that overlap it on the device (admitted before, during and after it).
* The server's transmission-loss spectrum equals the spectrum of the probe
  histories of the same case run directly in a batch (host DFT, the stated
  definition); a rigid duct's TL is Kirchhoff's viscous wall attenuation, and a
  liner near resonance attenuates well beyond it.
* Admission respects the memory budget (requests wait until memory frees up),
  every request completes, and a request that can never fit fails cleanly.
TL itself is this project's definition ("parity unpinned": the reference keeps
its TL post-processing in an unvendored toolkit).
"""
import time

import numpy as np
import pytest

from lbc_amd import batch as B
from lbc_amd import native
from lbc_amd import serving as SV

pytestmark = pytest.mark.gpu

PREC = native.F32_WIDE


def requests(n, seed=3):
    return SV.draw_requests(n, seed=seed, Nx=(512, 1024), Ny=(64, 96), steps=(800, 1500))


def wait_all(server, n, timeout=120.0):
    got = {}
    t0 = time.time()
    while len(got) < n:
        for r in server.poll():
            got[r.tag] = r
        assert time.time() - t0 < timeout, f"only {len(got)} of {n} results"
        time.sleep(0.002)
    return got


def as_arrays(res):
    n = res.nfreq
    return (np.array(res.tl_db[:n]), np.array(res.up_abs[:n]), np.array(res.down_abs[:n]))


def test_same_result_alone_and_in_mix():
    rs = SV.draw_requests(6, seed=3, Nx=(512, 1024), Ny=(64, 96), steps=(3000, 4000))
    x = rs[2]
    s = B.Server(PREC, budget=4 << 30, chunk=32)
    s.submit(x.case(), x.tl_spec(), 100)
    alone = wait_all(s, 1)[100]
    # x arrives while two others run; three more arrive while x runs
    for i in (0, 1):
        s.submit(rs[i].case(), rs[i].tl_spec(), i)
    time.sleep(0.005)
    s.submit(x.case(), x.tl_spec(), 101)
    for i in (3, 4, 5):
        time.sleep(0.003)
        s.submit(rs[i].case(), rs[i].tl_spec(), i)
    mix = wait_all(s, 6)
    s.close()
    assert alone.status == 0 and all(r.status == 0 for r in mix.values()), \
        [(t, r.status, r.guard) for t, r in mix.items()]
    for a, b in zip(as_arrays(alone), as_arrays(mix[101])):
        assert np.array_equal(a, b)
    # x really shared the device: its residency overlapped other requests'
    me = mix[101]
    overlap = [t for t, r in mix.items() if t != 101 and r.t_admit < me.t_done and
               me.t_admit < r.t_done]
    assert len(overlap) >= 2, overlap


def test_channel_requests_served_beside_liners():
    """Channel solves (Poiseuille channels of different sizes, no spectrum) are
    admitted and run beside liner requests: each returns status 0, no spectrum,
    and the FP64 mass and energy of its final state equal, bit for bit, those of
    the same case run alone in a batch in the server's step mode."""
    chans = [B.poiseuille(256, 64, 0.8, 600), B.poiseuille(384, 96, 1.0, 900, Fx=2e-6),
             B.poiseuille(128, 160, 0.7, 701)]
    liners = SV.draw_requests(3, seed=5, Nx=(512, 768), Ny=(64, 96), steps=(800, 1200))
    s = B.Server(PREC, budget=4 << 30, chunk=32)
    for i, r in enumerate(liners):
        s.submit(r.case(), r.tl_spec(), 10 + i)
        time.sleep(0.002)
        s.submit(chans[i], None, i)
    got = wait_all(s, 6)
    s.close()
    for i, c in enumerate(chans):
        r = got[i]
        assert r.status == 0 and r.nfreq == 0 and r.steps == c.steps
        c.keep_moments = True
        b = B.Batch(PREC, budget=B.case_bytes(c, PREC) + (64 << 20), chunk=32)
        b.set_pairs(True)
        j = b.add(c)
        b.run(c.steps + (c.steps & 1))   # pair mode advances in twos; the lattice stops at its steps
        b.synchronize()
        mass, energy = b.totals(j)
        b.close()
        assert r.mass == mass and r.energy == energy, (i, r.mass, mass, r.energy, energy)
    for i in range(3):
        assert got[10 + i].status == 0 and got[10 + i].nfreq == SV.NFREQ


def test_tl_equals_spectrum_of_batch_probes():
    x = requests(3, seed=5)[1]
    s = B.Server(PREC, budget=2 << 30, chunk=16)
    s.submit(x.case(), x.tl_spec(), 7)
    res = wait_all(s, 1)[7]
    s.close()
    b = B.Batch(PREC, budget=2 << 30)
    i = b.add(x.case())
    b.run(x.steps)
    b.synchronize()
    hist = b.probes(i)
    b.close()
    tl, up, dn = SV.tl_from_history(hist, x)
    tl_s, up_s, dn_s = as_arrays(res)
    assert np.abs(up_s - up).max() <= 1e-10 * up.max()
    assert np.abs(dn_s - dn).max() <= 1e-10 * dn.max()
    band = up > 1e-2 * up.max()
    assert np.abs(tl_s - tl)[band].max() < 1e-6
    assert np.isfinite(tl_s).all()


def kirchhoff_tl(f, L, H, tau):
    """Viscous (Stokes-layer) attenuation of a plane wave over length L in a 2D
    channel of height H between no-slip walls (Kirchhoff): amplitude
    exp(-alpha L), alpha = sqrt(omega nu / 2) / (H c), in dB."""
    nu = (tau - 0.5) / 3.0
    alpha = np.sqrt(2.0 * np.pi * f * nu / 2.0) / (H * SV.CS)
    return 20.0 * np.log10(np.e) * alpha * L


def test_rigid_duct_and_liner_spectra():
    """Rigid duct: the TL between the probes is the walls' viscous attenuation
    (Kirchhoff's closed form, within 0.5 dB at the low end of the band); liner of
    matched resistance: TL well above the rigid duct's near its resonance."""
    rigid = SV.LinerRequest(1024, 64, 1.0, 1 / 40, 4.0, 5e-4, 1500, rigid=True)
    liner = SV.LinerRequest(1024, 64, 1.0, 1 / 40, 4.0, 5e-4, 1500)
    s = B.Server(PREC, budget=2 << 30)
    s.submit(rigid.case(), rigid.tl_spec(), 0)
    s.submit(liner.case(), liner.tl_spec(), 1)
    got = wait_all(s, 2)
    s.close()
    f = np.linspace(SV.F_LO, SV.F_HI, SV.NFREQ)
    tl0, up0, _ = as_arrays(got[0])
    tl1, _, _ = as_arrays(got[1])
    band = up0 > 1e-2 * up0.max()
    kirch = kirchhoff_tl(f, L=512, H=62, tau=0.51)
    print("f:", np.round(f[band], 4))
    print("rigid TL (dB):", np.round(tl0[band], 3))
    print("Kirchhoff (dB):", np.round(kirch[band], 3))
    print("liner TL (dB):", np.round(tl1[band], 3))
    low = band & (f <= 1.0 / 60.0)
    assert low.sum() >= 6
    assert np.abs(tl0[low] - kirch[low]).max() < 0.5
    near = band & (np.abs(f - liner.f_res) < 0.01)
    assert (tl1[near] - tl0[near]).max() > 3.0


def test_budget_admission_and_completion():
    r = requests(1, seed=9)[0]
    one = B.case_bytes(r.case(keep_moments=True), PREC)   # served cases keep their moments
    budget = int(2.2 * one)             # room for two requests at a time
    s = B.Server(PREC, budget=budget, chunk=32)
    for i in range(5):
        s.submit(r.case(), r.tl_spec(), i)
    got = wait_all(s, 5)
    st = s.stats()
    s.close()
    assert all(x.status == 0 for x in got.values())
    assert st["peak_bytes"] <= budget
    order = sorted(got.values(), key=lambda x: x.t_admit)
    assert order[2].t_admit >= min(order[0].t_done, order[1].t_done) - 1e-3
    assert st["finished"] == 5 and st["running"] == 0 and st["waiting"] == 0
    tl = [as_arrays(x)[0] for x in got.values()]
    assert all(np.array_equal(tl[0], t) for t in tl[1:])


def test_request_that_cannot_fit_fails_cleanly():
    big = SV.LinerRequest(8192, 256, 1.0, 1 / 40, 4.0, 5e-4, SV.min_steps(8192))
    small = requests(1)[0]
    s = B.Server(PREC, budget=int(1.5 * B.case_bytes(small.case(keep_moments=True), PREC)))
    s.submit(big.case(), big.tl_spec(), 0)
    s.submit(small.case(), small.tl_spec(), 1)
    got = wait_all(s, 2)
    s.close()
    assert got[0].status == -4            # LBMX_ERR_NOMEM
    assert got[1].status == 0
