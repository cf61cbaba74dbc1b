"""Writes tests/golden/vtk_ref_13x7.vtk with the reference's own write_vtk_binary_2D
(src/utils/write_vtk_binary_2D.c, compiled into oracle/_ref/liblbm_ref.so by
oracle/Makefile) and the inputs it was given to tests/golden/vtk_ref_13x7.npz.

The inputs are seeded moments of a 13 x 7 lattice with NaN on three nodes (the
reference's compute_macroscopic_fields marks solid nodes with NaN), so the fixture
pins the header text, the big-endian float32 blocks and the NaN conversion.

    python tests/golden/make_vtk_fixture.py
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
NX, NY, DX, CU, CRHO = 13, 7, 0.0125, 2.75, 1.184


def inputs():
    rng = np.random.default_rng(20261203)
    rho = 1.0 + 0.01 * rng.standard_normal((NY, NX))
    u = 0.05 * rng.standard_normal((NY, NX))
    v = 0.05 * rng.standard_normal((NY, NX))
    for j, i in ((0, 0), (3, 5), (6, 12)):
        rho[j, i] = u[j, i] = v[j, i] = np.nan
    return u, v, rho


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    lib = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "liblbm_ref.so"))
    dp = C.POINTER(C.c_double)
    lib.write_vtk_binary_2D.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_double, dp, dp, dp,
                                        C.c_double, C.c_double]
    lib.write_vtk_binary_2D.restype = None
    u, v, rho = inputs()
    out = os.path.join(HERE, "vtk_ref_13x7.vtk")
    lib.write_vtk_binary_2D(out.encode(), NX, NY, DX, *(a.ctypes.data_as(dp) for a in (u, v, rho)),
                            CU, CRHO)
    np.savez(os.path.join(HERE, "vtk_ref_13x7.npz"), u=u, v=v, rho=rho,
             params=np.array([DX, CU, CRHO]))
    print(f"wrote {out} ({os.path.getsize(out)} bytes)")


if __name__ == "__main__":
    main()
