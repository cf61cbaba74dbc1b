"""CPU baseline timing (TEST INFRASTRUCTURE, bench.py's cpu_baseline leg only).

Times the reference library's ``main_lbm`` (``oracle/_ref/liblbm_ref.so``, built
from lbc's own ``src/`` by ``oracle/Makefile`` with the reference's flags,
``gcc -O3 -fopenmp``, no ``-march``) -- or, where that library was not built,
this project's C restatement (``liblbm_oracle.so``) -- on the periodic
Taylor-Green vortex N x N, tau = 0.8, u0 = 0.01, FP64, with the driver's f/f_new
swap after every call (cases/taylorgreen/taylorgreen.c's step loop without its
output).  Only the step loop is timed.  Prints one JSON object.

Run as its own process so that the OpenMP settings of the environment
(``OMP_NUM_THREADS``, ``OMP_PROC_BIND``, ``OMP_PLACES``) take effect when libgomp
starts:

    OMP_NUM_THREADS=16 OMP_PROC_BIND=close OMP_PLACES=cores \\
        python -m oracle.cpu_bench --size 2048 --seconds 10
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))


def tgv_state(lat, N, u0=0.01):
    """Taylor-Green vortex at equilibrium (cases/taylorgreen/taylorgreen.c set-up)."""
    k = 2.0 * math.pi / N
    j, i = np.mgrid[0:N, 0:N].astype(np.float64)
    lat.u[:] = u0 * np.sin(k * i) * np.cos(k * j)
    lat.v[:] = -u0 * np.cos(k * i) * np.sin(k * j)
    lat.rho[:] = 1.0 - 0.75 * u0 * u0 * (np.cos(2 * k * i) + np.cos(2 * k * j))
    cx = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1], dtype=np.float64)
    cy = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1], dtype=np.float64)
    w = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
    cu = lat.u[..., None] * cx + lat.v[..., None] * cy
    uu = (lat.u ** 2 + lat.v ** 2)[..., None]
    lat.f[:] = w * lat.rho[..., None] * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * uu)
    lat.f_new[:] = lat.f
    lat.omega[:] = 1.0 / 0.8


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--seconds", type=float, default=10.0)
    ap.add_argument("--max-steps", type=int, default=400)
    args = ap.parse_args()

    import oracle as O
    ref = O.reference()
    N = args.size
    lat = O.Lattice(N, N)
    tgv_state(lat, N)
    bcs = [("periodic_x", 0, None, None), ("periodic_y", 0, None, None)]
    if ref is None:
        kind, binary = "port", "oracle/liblbm_oracle.so"
        step = lambda: O.oracle_step(lat, bcs, 1, 1)   # noqa: E731
    else:
        kind, binary = "reference", "oracle/_ref/liblbm_ref.so"
        step = lambda: ref.step(lat, bcs, 1, 1)        # noqa: E731
    step()   # first call: thread team start-up, page faults
    n, t0 = 0, time.perf_counter()
    while n < args.max_steps:
        step()
        n += 1
        if time.perf_counter() - t0 >= args.seconds and n >= 3:
            break
    dt = time.perf_counter() - t0
    print(json.dumps({"MLUPS": N * N * n / dt / 1e6, "steps": n, "seconds": dt, "kind": kind,
                      "binary": binary, "N": N,
                      "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
                      "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
                      "omp_places": os.environ.get("OMP_PLACES")}), flush=True)


if __name__ == "__main__":
    main()
