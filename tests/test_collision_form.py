"""The device's FP64 collision (lbc_amd/csrc/hip/d2q9.h: moments_ref, collide_ref)
forms c.u and c.F with lattice components in {-1, 0, 1} as sums, differences and
signed copies, leaving out the products with zero components that the reference
evaluates (collision.c:30-37, compute_macroscopic_fields.c:27-41).  This restates
that form in NumPy, operation for operation in the same order (float64, no fused
operations), and checks it against the oracle's reference-form routines on
values chosen to reach the corner cases: exact zeros of either sign, rest states
(exact cancellations), zero momentum, subnormal sums, with and without a body
force.  The values must agree exactly (a zero may differ only in its sign, which
== ignores)."""
import ctypes as C

import numpy as np

import oracle as O

W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
CX = [0, 1, 0, -1, 0, 1, -1, -1, 1]
CY = [0, 0, 1, 0, -1, 1, 1, -1, -1]


def cdot(cx, cy, ax, ay):
    if cx == 0:
        return ay if cy > 0 else -ay
    if cy == 0:
        return ax if cx > 0 else -ax
    if cx > 0:
        return ax + ay if cy > 0 else ax - ay
    return ay - ax if cy > 0 else -ax - ay


def moments_form(f, Fx, Fy):
    s = f[..., 0].copy()
    for k in range(1, 9):
        s = s + f[..., k]
    sx = ((((f[..., 1] - f[..., 3]) + f[..., 5]) - f[..., 6]) - f[..., 7]) + f[..., 8]
    sy = ((((f[..., 2] - f[..., 4]) + f[..., 5]) + f[..., 6]) - f[..., 7]) - f[..., 8]
    r = np.where(s < 1e-16, 1e-16, s)
    return r, (sx + 0.5 * Fx) / r, (sy + 0.5 * Fy) / r


def collide_form(f, rho, ux, uy, om, force, Fx, Fy):
    usq = ux * ux + uy * uy
    uF = ux * Fx + uy * Fy
    out = np.empty_like(f)
    for k in range(9):
        w = W[k]
        if k == 0:
            cu = 0.0
            eq = w * rho * (1.0 - 1.5 * usq)
        else:
            cu = cdot(CX[k], CY[k], ux, uy)
            eq = w * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq)
        r = -om * (f[..., k] - eq)
        if force:
            if k == 0:
                Fk = w * -(3.0 * uF)
            else:
                cF = cdot(CX[k], CY[k], Fx, Fy)
                Fk = w * (3.0 * cF + 9.0 * cu * cF - 3.0 * uF)
            r = r + (1 - 0.5 * om) * Fk
        out[..., k] = f[..., k] + r
    return out


def sample(Ny, Nx, seed):
    rng = np.random.default_rng(seed)
    f = W * (1.0 + 0.01 * (rng.random((Ny, Nx, 9)) - 0.5))
    kind = rng.integers(0, 6, (Ny, Nx))
    f[kind == 1] = np.where(rng.random((np.sum(kind == 1), 9)) < 0.3, 0.0, f[kind == 1])
    f[kind == 2] = np.where(rng.random((np.sum(kind == 2), 9)) < 0.3, -0.0, f[kind == 2])
    f[kind == 3] = W                                          # rest state
    z = f[kind == 4]                                          # zero momentum
    z[:, 1], z[:, 5], z[:, 6], z[:, 2] = z[:, 3], z[:, 7], z[:, 8], z[:, 4]
    f[kind == 4] = z
    f[kind == 5] = (rng.random((np.sum(kind == 5), 9)) - 0.5) * 1e-300   # subnormal sums
    return f


def p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def test_zero_free_form_equals_reference_form():
    Ny, Nx = 64, 96
    mask = np.zeros((Ny, Nx), np.int32)
    for seed, (Fx, Fy) in enumerate([(0.0, 0.0), (1e-5, -3e-6), (0.0, 2e-6), (-0.0, 0.0)]):
        f = sample(Ny, Nx, seed)
        om = 0.5 + 1.49 * np.random.default_rng(seed + 7).random((Ny, Nx))
        rho, u, v = (np.zeros((Ny, Nx)) for _ in range(3))
        O.ORACLE.orc_moments(Nx, Ny, p(f), mask.ctypes.data_as(C.POINTER(C.c_int)), p(rho),
                             p(u), p(v), Fx, Fy)
        r2, u2, v2 = moments_form(f, Fx, Fy)
        assert np.array_equal(rho, r2) and np.array_equal(u, u2) and np.array_equal(v, v2)
        ref = f.copy()
        O.ORACLE.orc_collision(Nx, Ny, p(ref), p(rho), p(u), p(v), p(om),
                               mask.ctypes.data_as(C.POINTER(C.c_int)), Fx, Fy)
        force = Fx != 0.0 or Fy != 0.0
        got = collide_form(f, rho, u, v, om, force, Fx, Fy)
        assert np.array_equal(ref, got)


def test_form_check_detects_a_dropped_term():
    """The comparison is sharp: dropping the 4.5 (c.u)^2 term of one direction
    in the restated form must be caught."""
    Ny, Nx = 16, 16
    mask = np.zeros((Ny, Nx), np.int32)
    f = sample(Ny, Nx, 99)
    om = np.full((Ny, Nx), 1.25)
    rho, u, v = moments_form(f, 0.0, 0.0)
    ref = f.copy()
    O.ORACLE.orc_collision(Nx, Ny, p(ref), p(rho), p(u), p(v), p(om),
                           mask.ctypes.data_as(C.POINTER(C.c_int)), 0.0, 0.0)
    got = collide_form(f, rho, u, v, om, False, 0.0, 0.0)
    cu = u + v
    got[..., 5] += om * W[5] * rho * 4.5 * cu * cu        # the term, removed again
    assert not np.array_equal(ref, got)
