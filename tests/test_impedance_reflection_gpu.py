"""Normal-incidence reflection of the liner's impedance wall on the batched engine.

* FP64 tubes (impedance wall and rigid) are bit-identical to the oracle's run of
  the same tube, one step per launch and in pair mode;
* the reflection read from the engine's probe histories converges to
  R = (Z - rho0 c0) / (Z + rho0 c0) at first order in the frequency over tube
  scales 2 / 4 / 8 (bounds and their origin in tests/tube.py);
* FP32 storage with the FP64 collision (the served precision) gives the same |R|
  within 1e-4.
"""
import numpy as np
import pytest

from lbc_amd import batch as B
from lbc_amd import native

import tube as T

pytestmark = pytest.mark.gpu


def run_tubes(tubes, prec, pairs=False):
    """tubes: [(case, fields)] -> [(probe history [steps], final rho)]."""
    b = B.Batch(prec, budget=1 << 30)
    if pairs:
        b.set_pairs(True)
    ids = [b.add(c, fields=f) for c, f in tubes]
    longest = max(c.steps for c, _ in tubes)
    b.run(longest + (longest & 1))
    b.synchronize()
    out = []
    for i, (c, _) in zip(ids, tubes):
        assert b.steps_done(i) == c.steps
        assert b.status(i) == 0
        out.append((b.probes(i)[:c.steps, 0].copy(), b.moments(i)[0]))
    b.close()
    return out


@pytest.mark.parametrize("pairs", [False, True], ids=["single", "pairs"])
def test_tube_fp64_equals_oracle(pairs):
    tubes = [T.tube(1, T.msd(1), Nx=16), T.tube(1, None, Nx=16)]
    got = run_tubes(tubes, native.F64, pairs)
    for (c, f), (hist, rho) in zip(tubes, got):
        h_ref, rho_ref = T.oracle_tube(c, f, c.steps)
        assert np.array_equal(hist, h_ref), c.liner is not None
        assert np.array_equal(rho, rho_ref, equal_nan=True), c.liner is not None


def test_tube_reflection_converges_to_impedance_theory():
    scales = (2, 4, 8)
    tubes = []
    for s in scales:
        tubes += [T.tube(s, T.msd(s), Nx=16), T.tube(s, None, Nx=16)]
    got = run_tubes(tubes, native.F64, pairs=True)
    res = {}
    for k, s in enumerate(scales):
        res[s] = T.reflection_checks(s, got[2 * k][0], got[2 * k + 1][0], T.msd(s))
    print({s: tuple(round(x, 4) for x in r) for s, r in res.items()})
    assert T.passes(res), res


def test_tube_reflection_fp32_storage():
    s = 4
    tubes = [T.tube(s, T.msd(s), Nx=16), T.tube(s, None, Nx=16)]
    ref = run_tubes(tubes, native.F64)
    got = run_tubes(tubes, native.F32_WIDE)
    f_res = 1.0 / (40.0 * s)
    fs = f_res * np.linspace(0.5, 1.6, 23)
    R64 = T.reflection(ref[0][0], ref[1][0], s, fs)
    R32 = T.reflection(got[0][0], got[1][0], s, fs)
    print("max |R32 - R64|", float(np.max(np.abs(R32 - R64))))
    assert np.max(np.abs(R32 - R64)) < 1e-4
