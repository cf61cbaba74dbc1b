"""Writes tests/golden/nit_uniform_oracle_2400.npy: the oracle's (t, p', v_ac)
history of the NIT tube, 89 Pa / 2000 Hz, "uniform" variant (every row of the
impedance wall gets v_ac/cu; orc_nit_cfg.variant = 1), for 2400 steps at full
precision.  The pulse reaches the wall around step 1900, so the rows cover the
TDIBC under real load.  The oracle takes a few minutes for this; the GPU test
compares against the stored rows bit for bit.

    python tests/golden/make_nit_uniform_fixture.py
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

STEPS = 2400


def main():
    import oracle as O
    import nit_case as N
    cfg = O.OrcNitCfg()
    O.ORACLE.orc_nit_config(N.P_A, N.F_EXC, C.byref(cfg))
    cfg.variant = 1
    z = N.impedance_fields()
    st = O.OrcTdibc()
    st.Zinf = z["Zinf"]
    for name in ("A", "lambdas", "B", "C", "alpha", "beta"):
        arr = getattr(st, name)
        arr[0], arr[1] = z[name]
    rec = np.zeros((STEPS, 3))
    O.ORACLE.orc_nit_run(C.byref(cfg), N.P_A, N.F_EXC, C.byref(st), STEPS,
                         rec.ctypes.data_as(C.POINTER(C.c_double)), None, None)
    out = os.path.join(HERE, f"nit_uniform_oracle_{STEPS}.npy")
    np.save(out, rec)
    print(f"wrote {out}")


if __name__ == "__main__":
    main()
