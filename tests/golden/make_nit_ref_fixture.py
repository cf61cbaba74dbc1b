"""Writes tests/golden/nit_ref_2000Hz_89Pa_first64.txt by running the reference's own
NIT driver.

The reference ships cases/nit/sol/tdibc_2000Hz_89Pa.txt, but its current
cases/nit/nit.c does not reproduce that file: built from the reference sources
(oracle/Makefile target ref-nit) and run as `nit 89 2000`, it diverges within
~60 steps (the orifice-scaled wall velocity of nit.c:286 destabilises the
tau = 0.5005 lattice).  Those first rows are still the reference's exact
behaviour for Mei's initialization, pressure_inlet, velocity_outlet_regularized,
periodic_y and the TDIBC ODE, and near the blow-up they print up to 17
significant digits, so they pin the whole nit path.

    python tests/golden/make_nit_ref_fixture.py
"""
import os
import shutil
import subprocess
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
ROWS = 64


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref-nit"], check=True)
    run = os.path.join(ROOT, "oracle", "_ref", "nit_run")
    out = os.path.join(run, "sol", "tdibc_2000Hz_89Pa.txt")
    shutil.rmtree(os.path.join(run, "sol"), ignore_errors=True)
    p = subprocess.Popen([os.path.join(ROOT, "oracle", "_ref", "nit"), "89", "2000"], cwd=run,
                         stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    try:
        t0 = time.time()
        while time.time() - t0 < 300:
            time.sleep(1)
            if os.path.exists(out) and sum(1 for _ in open(out)) > ROWS + 40:
                break
    finally:
        p.kill()
        p.wait()
    rows = open(out).read().splitlines()[:ROWS]
    assert len(rows) == ROWS, len(rows)
    dst = os.path.join(HERE, f"nit_ref_2000Hz_89Pa_first{ROWS}.txt")
    with open(dst, "w") as fh:
        fh.write("\n".join(rows) + "\n")
    shutil.rmtree(os.path.join(run, "sol"), ignore_errors=True)
    print(f"wrote {dst}")


if __name__ == "__main__":
    main()
