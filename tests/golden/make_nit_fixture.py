"""Writes tests/golden/nit_2000Hz_89Pa.npz from data files the reference ships.

Inputs (read only, from the reference tree):
  cases/nit/impedance_89Pa.txt          impedance fit read by nit.c:49-74
                                        (Zinf, then A, lambdas, B, C, alpha, beta)
  cases/nit/sol/tdibc_2000Hz_89Pa.txt   nit.c:273 output of `./nit.o 89 2000`:
                                        rows "t p' v_ac" printed with %lf

The %lf columns have exactly six decimals, so they are stored losslessly as
integers in units of 1e-6 (int64), together with the raw impedance numbers.

    python tests/golden/make_nit_fixture.py [/root/reference]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def parse_impedance(path):
    vals = [float(x) for x in open(path).read().split()]
    assert len(vals) == 13, len(vals)
    return np.array(vals, dtype=np.float64)


def parse_series(path):
    rows = []
    for line in open(path):
        parts = line.split()
        if len(parts) != 3:
            continue
        micro = []
        for s in parts:
            neg = s.startswith("-")
            a, b = s.lstrip("-").split(".")
            assert len(b) == 6, s
            v = int(a) * 1_000_000 + int(b)
            micro.append(-v if neg else v)
        # keep the sign of "-0.000000" (a printed negative zero)
        rows.append(micro + [int(p.startswith("-")) for p in parts])
    return np.array(rows, dtype=np.int64)


def main(ref="/root/reference"):
    imp = parse_impedance(os.path.join(ref, "cases/nit/impedance_89Pa.txt"))
    ser = parse_series(os.path.join(ref, "cases/nit/sol/tdibc_2000Hz_89Pa.txt"))
    out = os.path.join(HERE, "nit_2000Hz_89Pa.npz")
    np.savez_compressed(out, impedance=imp, micro=ser[:, :3], negsign=ser[:, 3:].astype(np.int8))
    print(f"wrote {out}: {ser.shape[0]} rows, {os.path.getsize(out)} bytes")


if __name__ == "__main__":
    main(*sys.argv[1:])
