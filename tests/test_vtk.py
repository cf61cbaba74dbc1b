"""VTK output (write_vtk_binary_2D, lbm_utils.h).

CPU: the drop-in host writer reproduces, byte for byte, the file the reference's
own writer produced from the same inputs (tests/golden/vtk_ref_13x7.vtk, NaN nodes
included), and the file decodes to (float)(u * cu), (float)(v * cu), 0 and
(float)(rho * crho) in big-endian order.

GPU: lbmx_write_vtk packs the same file on the device; it must equal the host
writer's file for the same moments, before the first step (uploaded arrays, with
NaNs of either sign and with payloads) and after steps of a lattice with solids.
"""
import os

import numpy as np
import pytest

from lbc_amd import native

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden_inputs():
    z = np.load(os.path.join(GOLDEN, "vtk_ref_13x7.npz"))
    dx, cu, crho = z["params"]
    return z["u"], z["v"], z["rho"], float(dx), float(cu), float(crho)


def test_host_writer_matches_reference(tmp_path):
    u, v, rho, dx, cu, crho = golden_inputs()
    out = tmp_path / "out.vtk"
    native.write_vtk_host(str(out), dx, u, v, rho, cu, crho)
    with open(os.path.join(GOLDEN, "vtk_ref_13x7.vtk"), "rb") as f:
        ref = f.read()
    assert out.read_bytes() == ref


def decode(data, Nx, Ny):
    head = b"VECTORS velocity float\n"
    a = data.index(head) + len(head)
    vel = np.frombuffer(data[a:a + 12 * Nx * Ny], dtype=">f4").reshape(Ny, Nx, 3)
    mid = b"\nSCALARS density float 1\nLOOKUP_TABLE default\n"
    assert data[a + 12 * Nx * Ny:].startswith(mid)
    b = a + 12 * Nx * Ny + len(mid)
    den = np.frombuffer(data[b:b + 4 * Nx * Ny], dtype=">f4").reshape(Ny, Nx)
    assert len(data) == b + 4 * Nx * Ny
    return data[:a], vel, den


def test_reference_file_decodes():
    u, v, rho, dx, cu, crho = golden_inputs()
    Ny, Nx = rho.shape
    with open(os.path.join(GOLDEN, "vtk_ref_13x7.vtk"), "rb") as f:
        data = f.read()
    header, vel, den = decode(data, Nx, Ny)
    assert header.startswith(b"# vtk DataFile Version 3.0\n")
    assert f"DIMENSIONS {Nx} {Ny} 1".encode() in header
    assert f"POINT_DATA {Nx * Ny}".encode() in header
    fluid = ~np.isnan(rho)
    assert np.array_equal(vel[..., 0][fluid], (u * cu).astype(np.float32)[fluid])
    assert np.array_equal(vel[..., 1][fluid], (v * cu).astype(np.float32)[fluid])
    assert np.all(vel[..., 2] == 0)
    assert np.array_equal(den[fluid], (rho * crho).astype(np.float32)[fluid])
    assert np.all(np.isnan(den[~fluid])) and np.all(np.isnan(vel[..., 0][~fluid]))


@pytest.mark.gpu
def test_device_pack_equals_host_writer(tmp_path):
    from cases import Scenario
    sc = Scenario("vtk", 70, 33, [("velocity_inlet_regularized", 0),
                                  ("pressure_outlet_regularized", 69)], seed=5, use_ibb=1)
    sc.solid_rows(0, 32).disk(20.3, 15.6, 5.2)
    f, fn, rho, u, v = sc.initial_state()
    # NaNs of both signs and with payloads on a few nodes of the uploaded arrays
    bits = np.array([0x7FF8000000000000, 0xFFF8000000000000, 0x7FF80000DEADBEEF,
                     0xFFFC123456789ABC], dtype=np.uint64).view(np.float64)
    for n, (j, i) in enumerate(((1, 1), (5, 9), (17, 40), (31, 68))):
        rho[j, i] = bits[n]
        u[j, i] = bits[(n + 1) % 4]
        v[j, i] = bits[(n + 2) % 4]
    lat = native.Lattice(sc.Nx, sc.Ny, use_ibb=1)
    lat.set_geometry(sc.mask, sc.omega, sc.phi)
    lat.set_boundaries(sc.bc_names)
    for b, (_, _, v1, v2) in enumerate(sc.vals(0)):
        lat.set_bc_values(b, v1, v2)
    lat.upload(f, fn, rho, u, v)
    dx, cu, crho = 0.01, 3.5, 1.2
    a, b = tmp_path / "dev0.vtk", tmp_path / "host0.vtk"
    lat.write_vtk(str(a), dx, cu, crho)
    native.write_vtk_host(str(b), dx, u, v, rho, cu, crho)
    assert a.read_bytes() == b.read_bytes()
    lat.step(7)
    r2, u2, v2 = lat.moments()
    a, b = tmp_path / "dev7.vtk", tmp_path / "host7.vtk"
    lat.write_vtk(str(a), dx, cu, crho)
    native.write_vtk_host(str(b), dx, u2, v2, r2, cu, crho)
    lat.close()
    assert a.read_bytes() == b.read_bytes()
