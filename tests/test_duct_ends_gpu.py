"""Reflection of the served ducts' ends on the batched engine (FP64, pair mode):
convective ends inside a tau sponge 3 Nx / 128 wide, as serving.LinerRequest
builds them, at Nx = 2048 / 4096 / 8192, against a convective end alone (method
in tests/ductend.py; the oracle's measurement and the reference's alternatives
are in tests/test_duct_ends.py).  Measured: the sponge end reflects |R| <= 0.25
at 1/160 cycles per step (less for wider sponges) and <= 0.09 from 1/40 up; a
convective end alone ~0.27 at every frequency."""
import numpy as np
import pytest

from lbc_amd import batch as B
from lbc_amd import native

import ductend as DE

pytestmark = pytest.mark.gpu

NY = 64
FREQS = np.linspace(1.0 / 160.0, 1.0 / 20.0, 10)


def test_served_duct_end_reflection():
    b = B.Batch(native.F64, budget=4 << 30)
    b.set_pairs(True)
    runs = []
    for Nx in (2048, 4096, 8192):
        sp = max(8, 3 * Nx // 128)
        ids = {}
        for end in ("rigid", "sponge", "convective"):
            c, mid = DE.duct(Nx, NY, sp, end)
            ids[end] = (b.add(c), c.steps)
        runs.append((Nx, mid, ids))
    longest = max(s for _, _, ids in runs for _, s in ids.values())
    b.run(longest + (longest & 1))
    b.synchronize()
    low = {}
    for Nx, mid, ids in runs:
        h = {e: b.probes(i)[:s, 0].copy() for e, (i, s) in ids.items()}
        for e, (i, _) in ids.items():
            assert b.status(i) == 0
        Rs = DE.reflection(h["sponge"], h["rigid"], mid, FREQS)
        Rc = DE.reflection(h["convective"], h["rigid"], mid, FREQS)
        print(Nx, np.round(Rs, 3).tolist(), np.round(Rc, 3).tolist())
        assert np.all(Rs < 0.3)
        assert np.all(Rs[FREQS >= 1.0 / 40.0] < 0.1)
        assert 0.22 < np.median(Rc) < 0.32
        low[Nx] = Rs[0]
    b.close()
    assert low[8192] < low[2048]        # a wider sponge absorbs longer waves
