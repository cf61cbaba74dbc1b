"""GPU checks of the device-side TDIBC (lumped NIT tube and per-node liner duct).

* NIT tube, as cases/nit/nit.c builds it (Mei's initialization on the device,
  periodic_y + pressure_inlet + velocity_outlet_regularized, wall velocity from
  the lumped TDIBC controller): the history equals, as printed, the first 64 rows
  the reference's own nit driver writes, and the oracle's rows bit for bit.
* The bounded "uniform" variant over 2400 steps (the pulse reaches the wall near
  step 1900): bit-identical to the oracle's stored rows.
* The full uniform run (25,047 steps) tracks the reference's published
  tdibc_2000Hz_89Pa.txt series (physics-level agreement; that file comes from a
  nit version not in the tree: the tree's nit.c, built from the reference's
  own sources, diverges from it within ~60 steps).
* Liner duct (per-node TDIBC on a regularized wall, sponge ends): FP64 equals
  the oracle bit for bit; FP32-shifted against FP64 within a stated tolerance.
"""
import ctypes as C
import os

import numpy as np
import pytest

import nit_case as N
import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def impedance():
    z = N.impedance_fields()
    return {k: (tuple(v) if k != "Zinf" else v) for k, v in z.items()}


def test_nit_matches_reference_driver_rows():
    from lbc_amd import cases
    rows = N.reference_rows()
    lat = cases.nit_tube(cases.NitSpec(variant=0), impedance(), record_steps=len(rows))
    lat.step(len(rows))
    rec = lat.tdibc_history()
    lat.close()
    assert N.format_rows(rec) == rows
    ref, _ = N.oracle_run(len(rows))
    assert np.array_equal(rec, ref, equal_nan=True)


def test_nit_uniform_bit_identical_2400():
    from lbc_amd import cases
    fix = np.load(os.path.join(HERE, "golden", "nit_uniform_oracle_2400.npy"))
    lat = cases.nit_tube(cases.NitSpec(variant=1), impedance(), record_steps=fix.shape[0])
    lat.step(fix.shape[0])
    rec = lat.tdibc_history()
    lat.close()
    assert np.array_equal(rec, fix)
    assert np.abs(fix[:, 1]).max() > 100.0      # the pulse has arrived (p' in Pa)


def test_nit_uniform_full_run_tracks_published_series():
    from lbc_amd import cases
    s = cases.NitSpec(variant=1)
    nt = s.derived()["Nt"]
    lat = cases.nit_tube(s, impedance(), record_steps=nt)
    lat.step(nt)
    rec = lat.tdibc_history()
    lat.close()
    _, micro, _ = N.load_golden()
    assert rec.shape[0] == micro.shape[0] == nt
    pub = micro[:, 1:] / 1e6
    ep = np.abs(rec[:, 1] - pub[:, 0]).max() / np.abs(pub[:, 0]).max()
    ev = np.abs(rec[:, 2] - pub[:, 1]).max() / np.abs(pub[:, 1]).max()
    print(f"uniform NIT vs published series: max |dp'|/max|p'| = {ep:.4f}, "
          f"max |dv|/max|v| = {ev:.4f}")
    assert ep < 0.05 and ev < 0.05


def small_liner():
    from lbc_amd.cases import LinerSpec
    return LinerSpec(Nx=384, Ny=40, tau=0.51, amp=1e-3, x0=64.0, halfwidth=8.0, liner_x0=128,
                     liner_x1=256, R=1.0, f_res=1 / 40, M=8.0, sponge=24, tau_sponge=0.9,
                     probe_x=300, probe_y=20)


def oracle_liner(s, nt):
    cfg = O.OrcLinerCfg()
    for k in ("Nx", "Ny", "liner_x0", "liner_x1", "sponge", "probe_x", "probe_y"):
        setattr(cfg, k, getattr(s, k))
    for k in ("tau", "tau_sponge", "rho0", "amp", "x0", "halfwidth"):
        setattr(cfg, k, float(getattr(s, k)))
    cfg.gain = -1.0
    z = s.impedance()
    cfg.imp.Zinf = z["Zinf"]
    for name in ("A", "lambdas", "B", "C", "alpha", "beta"):
        arr = getattr(cfg.imp, name)
        arr[0], arr[1] = z[name]
    probe = np.zeros(nt)
    wall = np.zeros(s.liner_x1 - s.liner_x0 + 1)
    rho = np.zeros((s.Ny, s.Nx))
    dp = C.POINTER(C.c_double)
    O.ORACLE.orc_liner_run(C.byref(cfg), nt, probe.ctypes.data_as(dp), wall.ctypes.data_as(dp),
                           None, rho.ctypes.data_as(dp))
    return probe, wall, rho


def gpu_liner(s, nt, precision):
    from lbc_amd import cases
    lat = cases.liner_duct(s, precision=precision, record_steps=nt)
    lat.step(nt)
    probe = lat.probe_history()[:, 0]
    rho, _, _ = lat.moments()
    lat.close()
    wall = cases.CS * cases.CS * (rho[0, s.liner_x0:s.liner_x1 + 1] - s.rho0) * 1.0
    return probe, wall, rho


def test_liner_fp64_bit_identical():
    s = small_liner()
    nt = 400
    p_ref, w_ref, r_ref = oracle_liner(s, nt)
    p_gpu, w_gpu, r_gpu = gpu_liner(s, nt, precision=0)
    assert np.array_equal(p_gpu, p_ref)
    assert np.array_equal(r_gpu, r_ref, equal_nan=True)
    assert np.array_equal(w_gpu, w_ref)
    assert np.abs(w_ref).max() > 0.0


# FP32-shifted vs FP64 on the liner: wall pressure and transmitted probe, relative
# to the pulse scale (amp * cs^2) -- stated tolerance 1e-5 (observed ~4e-7 with the
# bias-free FP32 collision; a rounding bias that accumulates over the run would
# exceed it).
FP32_LINER_TOL = 1e-5


def test_liner_fp32_close_to_fp64():
    s = small_liner()
    nt = 400
    p64, w64, _ = gpu_liner(s, nt, precision=0)
    p32, w32, _ = gpu_liner(s, nt, precision=1)
    scale = s.amp / 3.0
    ew = np.abs(w32 - w64).max() / scale
    ep = np.abs(p32 - p64).max() / s.amp
    print(f"liner FP32 vs FP64: wall p' err {ew:.2e}, probe err {ep:.2e} (relative to pulse)")
    assert ew < FP32_LINER_TOL and ep < FP32_LINER_TOL
