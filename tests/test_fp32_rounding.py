"""Rounding bias of the FP32-storage collision (lbc_amd/csrc/hip/step_kernels.h:
moments_f32, collide_f32), restated in NumPy float32 on the CPU.

A rounding that repeats identically at a slowly changing node step after step
accumulates in the velocity instead of averaging out.  Two such roundings were
in the FP32 step: the velocity divided by rho0 + dr rounded to FP32's grid
around rho0, and the FP32 weights 4/9, 1/9, 1/36, all 2^-27 too large, scaling
the equilibrium's momentum each step.  The device step avoids both: u =
(j / rho0)(1 + q) with q = -dr / (rho0 + dr), and the equilibrium formed as
fma(w, X, w_lo X) with the weight's remainder w_lo.  On a 64 x 64 Taylor-Green
vortex the old form drifts from FP64 past 1e-5 within 1,500 steps, the device's
form stays below 2e-6.  Fused multiply-adds are emulated as one float64
evaluation rounded to float32 (at most one extra rounding)."""
import math

import numpy as np

W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
CX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
CY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
f32 = np.float32


def fma(a, b, c):
    return (np.asarray(a, np.float64) * b + c).astype(f32)


def stream(g):
    return np.stack([np.roll(np.roll(g[k], CY[k], axis=0), CX[k], axis=1) for k in range(9)])


def tgv(N, u0=0.01):
    k = 2 * np.pi / N
    j, i = np.mgrid[0:N, 0:N].astype(np.float64)
    u = u0 * np.sin(k * i) * np.cos(k * j)
    v = -u0 * np.cos(k * i) * np.sin(k * j)
    rho = 1 - 3 * u0 ** 2 / 4 * (np.cos(2 * k * i) + np.cos(2 * k * j))
    usq = u * u + v * v
    return np.stack([W[q] * rho * (1 + 3 * (CX[q] * u + CY[q] * v)
                                   + 4.5 * (CX[q] * u + CY[q] * v) ** 2 - 1.5 * usq)
                     for q in range(9)])


def step64(f, om):
    f = stream(f)
    rho = f.sum(0)
    u = (CX[:, None, None] * f).sum(0) / rho
    v = (CY[:, None, None] * f).sum(0) / rho
    usq = u * u + v * v
    out = np.empty_like(f)
    for q in range(9):
        cu = CX[q] * u + CY[q] * v
        out[q] = f[q] - om * (f[q] - W[q] * rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * usq))
    return out, u, v


def step32(g, om, device_form):
    """FP32-shifted step (g = f - w rho0, rho0 = 1, no body force)."""
    g = stream(g)
    s56, s78 = g[5] + g[6], g[7] + g[8]
    dr = (((g[0] + g[1]) + (g[2] + g[3])) + g[4]) + (s56 + s78)
    jx = (g[1] - g[3]) + (g[5] - g[6]) + (g[8] - g[7])
    jy = ((g[2] - g[4]) + s56) - s78
    rho = f32(1) + dr
    if device_form:
        q = (-dr * (f32(1) / rho)).astype(f32)
        ux, uy = fma(jx, q, jx), fma(jy, q, jy)
    else:
        inv = (f32(1) / rho).astype(f32)
        ux, uy = (jx * inv).astype(f32), (jy * inv).astype(f32)
    usq15 = f32(1.5) * fma(ux, ux, uy * uy)
    omf = f32(om)
    out = np.empty_like(g)
    for q in range(9):
        w = f32(W[q])
        if q == 0:
            t = -usq15
        else:
            cu = (CX[q] * ux + CY[q] * uy).astype(f32)   # exact: components in {-1, 0, 1}
            t = fma(cu, fma(f32(4.5), cu, f32(3)), -usq15)
        X = fma(rho, t, dr)
        geq = fma(w, X, f32(W[q] - np.float64(w)) * X) if device_form else (w * X).astype(f32)
        out[q] = fma(-omf, g[q] - geq, g[q])
    return out, ux, uy


def drift(device_form, N=64, steps=1500, tau=0.8):
    om = 1 / tau
    f = tgv(N)
    g = (f - W[:, None, None]).astype(f32)
    for _ in range(steps):
        f, u, v = step64(f, om)
        g, u32, v32 = step32(g, om, device_form)
    u32, v32 = u32.astype(np.float64), v32.astype(np.float64)
    return math.sqrt(float(((u32 - u) ** 2 + (v32 - v) ** 2).sum()) / float((u * u + v * v).sum()))


def test_device_form_does_not_drift():
    assert drift(True) < 2e-6


def test_old_form_drifts():
    """The check has teeth: the form with both biased roundings drifts past 1e-5."""
    assert drift(False) > 1e-5
