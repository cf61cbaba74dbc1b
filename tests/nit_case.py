"""The reference's impedance-liner (NIT) case, cases/nit/nit.c, as test data.

Golden output: tests/golden/nit_2000Hz_89Pa.npz (the reference's own
cases/nit/sol/tdibc_2000Hz_89Pa.txt and impedance_89Pa.txt, see
tests/golden/make_nit_fixture.py).  Rows are compared as the strings nit.c:273
prints ("%lf %lf %lf"), so agreement is exact to the printed precision.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                      "nit_2000Hz_89Pa.npz")
P_A, F_EXC = 89.0, 2000.0


def load_golden():
    d = np.load(GOLDEN)
    return d["impedance"], d["micro"], d["negsign"]


def golden_rows(n=None):
    _, micro, neg = load_golden()
    rows = []
    for r in range(micro.shape[0] if n is None else n):
        cols = []
        for c in range(3):
            v = int(micro[r, c])
            s = f"{abs(v) // 1_000_000}.{abs(v) % 1_000_000:06d}"
            cols.append(("-" if neg[r, c] else "") + s)
        rows.append(" ".join(cols))
    return rows


def _cfmt(x):
    """printf("%lf") of one double, including glibc's signed "nan"/"-nan"."""
    if x != x:
        return "-nan" if np.signbit(x) else "nan"
    return "%f" % x


def format_rows(rec):
    """nit.c:273 formatting ("%lf %lf %lf") of (t, p', v_ac) rows."""
    return [" ".join(_cfmt(c) for c in row) for row in rec]


def reference_rows():
    """The first rows printed by the reference's own nit driver (see
    tests/golden/make_nit_ref_fixture.py)."""
    p = os.path.join(os.path.dirname(GOLDEN), "nit_ref_2000Hz_89Pa_first64.txt")
    return open(p).read().splitlines()


def impedance_fields():
    imp = load_golden()[0]
    return {"Zinf": imp[0], "A": imp[1:3], "lambdas": imp[3:5], "B": imp[5:7],
            "C": imp[7:9], "alpha": imp[9:11], "beta": imp[11:13]}


def oracle_run(nt):
    """nt steps of the oracle's restatement of nit.c; returns (rec[nt,3], cfg)."""
    import oracle as O
    cfg = O.OrcNitCfg()
    O.ORACLE.orc_nit_config(P_A, F_EXC, C.byref(cfg))
    z = impedance_fields()
    st = O.OrcTdibc()
    st.Zinf = z["Zinf"]
    for name in ("A", "lambdas", "B", "C", "alpha", "beta"):
        arr = getattr(st, name)
        arr[0], arr[1] = z[name]
    rec = np.zeros((nt, 3))
    O.ORACLE.orc_nit_run(C.byref(cfg), P_A, F_EXC, C.byref(st), nt,
                         rec.ctypes.data_as(C.POINTER(C.c_double)), None, None)
    return rec, cfg
