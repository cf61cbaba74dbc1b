"""Normal-incidence reflection of the liner's impedance wall, on the oracle.

The liner couples the impedance ODE (nit.c:13-47, per wall node) to the fluid
through velocity_bottom_regularized (velocity_bottom_regularized.c:17-71) with the
wall's lagged density as p' and gain -1 (the oracle's orc_liner_run order,
oracle/lbm_oracle.c:702-712).  Its parts are pinned elsewhere; this pins the
composition against physics: a plane wave meeting a wall of impedance Z reflects
with R = (Z - rho0 c0) / (Z + rho0 c0).

The scheme lags the wall pressure by a step and takes the wall density from the
node above, so the measured |R| departs from R at first order in the frequency:
on tubes scaled by s (resonance 1/(40 s) per step, pulse and tube s times larger)
the largest departure over 0.5 .. 1.6 times the resonance must halve as s doubles,
and the notch at resonance (matched resistance: R = 0) must deepen and stay at the
resonance.  Mistakes planted in the oracle's run (the gain's sign, a 10 % error in
the reactance slope, a 20 % error in the resistance) fail these checks.
"""
import math

import numpy as np
import pytest

from lbc_amd import batch as B
from lbc_amd import cases

import tube as T


def oracle_runs(scale, imp_sim, gain=-1.0):
    c, fw = T.tube(scale, imp_sim, gain=gain)
    cr, fr = T.tube(scale, None)
    hw, _ = T.oracle_tube(c, fw, c.steps)
    hr, _ = T.oracle_tube(cr, fr, cr.steps)
    return hw, hr


def test_admittance_closed_forms():
    """ade_admittance of the mass-spring-damper fit is s / (M s^2 + R s + K), for a
    pole pair and for the over-damped fit with two real poles."""
    w = np.linspace(0.01, 0.5, 50)
    s = 1j * w
    for M, R, K in [(8 * T.CS, T.CS, 8 * T.CS * (2 * math.pi / 40) ** 2),
                    (0.3, 2.0, 0.05)]:
        Y = cases.ade_admittance(cases.msd_ade(M, R, K), w)
        assert np.allclose(Y, s / (M * s * s + R * s + K), rtol=1e-10, atol=0)
    # matched resistance at resonance: no reflection; a rigid wall (Y = 0): full
    imp = T.msd(1)
    assert abs(cases.normal_reflection(imp, 2 * math.pi / 40)) < 1e-12
    z = dict(Zinf=0.0, A=(0.0, 0.0), lambdas=(-1.0, -1.0), B=(0.0, 0.0), C=(0.0, 0.0),
             alpha=(-1.0, -1.0), beta=(0.0, 0.0))
    assert abs(cases.normal_reflection(z, 0.1) - 1.0) < 1e-15


def test_tube_case_is_valid():
    c, _ = T.tube(1, T.msd(1))
    B.case_check(c)
    cr, _ = T.tube(1, None)
    B.case_check(cr)
    c.per_x = 0                     # an end-to-end opening needs a periodic row
    with pytest.raises(Exception):
        B.case_check(c)


def test_oracle_reflection_converges_to_impedance_theory():
    res = {}
    for s in (1, 2, 4):
        res[s] = T.reflection_checks(s, *oracle_runs(s, T.msd(s)), T.msd(s))
    print({s: tuple(round(x, 4) for x in r) for s, r in res.items()})
    assert T.passes(res), res


@pytest.mark.parametrize("mistake", ["gain_sign", "reactance_slope", "resistance"])
def test_planted_mistakes_fail(mistake):
    res = {}
    for s in ((1, 2) if mistake == "gain_sign" else (1, 2, 4)):
        if mistake == "gain_sign":
            runs = oracle_runs(s, T.msd(s), gain=+1.0)
        elif mistake == "reactance_slope":
            runs = oracle_runs(s, T.msd(s, M=8.8))
        else:
            runs = oracle_runs(s, T.msd(s, R=1.2))
        with np.errstate(all="ignore"):
            res[s] = T.reflection_checks(s, *runs, T.msd(s))
    assert not T.passes(res), res
