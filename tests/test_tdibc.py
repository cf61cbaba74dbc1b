"""CPU checks of the time-domain impedance boundary (cases/nit/nit.c:13-47).

* The oracle's restatement of the whole NIT loop reproduces, character for
  character, the first 64 rows the reference's own nit driver prints
  (tests/golden/nit_ref_2000Hz_89Pa_first64.txt), including the rows where the
  current nit.c diverges (up to 17 printed significant digits).
* The explicit-Euler ODE update equals its closed form for a constant pressure,
  with the in-place (aliased) complex-pair update the reference performs.
* The mass-spring-damper mapping used for liners (lbc_amd.cases.msd_ade)
  reproduces the analytic admittance 1/(R + i(wM - K/w)) when the ODE is driven
  by a sinusoid.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O
import nit_case as N


def test_oracle_nit_matches_reference_driver():
    rows = N.reference_rows()
    rec, cfg = N.oracle_run(len(rows))
    assert (cfg.Nx, cfg.Ny, cfg.nt) == (14461, 65, 25047)   # nit.c:128-130 at 2000 Hz
    assert N.format_rows(rec) == rows


def _tdibc(z):
    st = O.OrcTdibc()
    st.Zinf = z["Zinf"]
    for name in ("A", "lambdas", "B", "C", "alpha", "beta"):
        arr = getattr(st, name)
        arr[0], arr[1] = z[name]
    return st


def test_ade_constant_pressure_closed_form():
    z = N.impedance_fields()
    st = _tdibc(z)
    dt, p, n = 2.59e-7, 3.7, 400
    got = [O.ORACLE.orc_tdibc_velocity(C.byref(st), p, dt) for _ in range(n)]
    # closed form: phi_n = p((1 + l dt)^n - 1)/l; the complex pair follows the
    # aliased update psi' = psi + dt(p + a psi - b chi), chi' = chi + dt(b psi' + a chi)
    steps = np.arange(1, n + 1)
    v = np.zeros(n)
    for i in range(2):
        lam = z["lambdas"][i]
        phi = p * ((1 + lam * dt) ** steps - 1) / lam
        a, b = z["alpha"][i], z["beta"][i]
        M = np.array([[1 + dt * a, -dt * b], [dt * b * (1 + dt * a), 1 + dt * a - dt * dt * b * b]])
        c = np.array([dt * p, dt * b * dt * p])
        s = np.zeros(2)
        psi = np.zeros(n)
        chi = np.zeros(n)
        for k in range(n):
            s = M @ s + c
            psi[k], chi[k] = s
        v += -z["A"][i] * phi - 2.0 * (z["B"][i] * psi + z["C"][i] * chi)
    v += z["Zinf"] * p
    np.testing.assert_allclose(got, v, rtol=1e-9, atol=1e-14)
    # a non-aliased update (chi from the old psi) would differ at this precision
    assert abs(got[-1] - v[-1]) < 1e-6 * abs(v[-1])


@pytest.mark.parametrize("M,R,f_res", [(8.0, 1.0, 1 / 40), (4.0, 0.5, 1 / 20), (12.0, 3.0, 1 / 80)])
def test_msd_mapping_matches_impedance(M, R, f_res):
    from lbc_amd.cases import CS, msd_ade
    Mr, Rr = M * CS, R * CS
    w0 = 2 * math.pi * f_res
    K = Mr * w0 * w0
    z = msd_ade(Mr, Rr, K)
    st = _tdibc(z)
    w = 0.8 * w0
    dt = 0.01                      # small step: the scheme approaches the ODE
    n_per = int(round(2 * math.pi / w / dt))
    n = 60 * n_per
    t = np.arange(n) * dt
    v = np.array([O.ORACLE.orc_tdibc_velocity(C.byref(st), math.sin(w * tk), dt) for tk in t])
    tail = slice(n - 10 * n_per, n)
    # fit v = Re(Y e^{i w t}) over the last periods
    A = np.stack([np.sin(w * t[tail]), np.cos(w * t[tail])], axis=1)
    cs, cc = np.linalg.lstsq(A, v[tail], rcond=None)[0]
    Y_num = cs + 1j * cc           # response to sin: Im(Y e^{iwt}) = Re(Y) sin + Im(Y) cos
    Y_ref = 1.0 / (Rr + 1j * (w * Mr - K / w))
    assert abs(Y_num - Y_ref) <= 0.02 * abs(Y_ref)
