"""Benchmark of the lbc D2Q9 hot path on one MI355X.

Headline (``value``): FP64 MLUPS of the fused step on the periodic Taylor-Green
lattice 16384 x 16384 (tau = 0.8, u0 = 0.01), the top of the roofline sweep of
the channel/cylinder configuration (1024^2 .. 16384^2).  The lattice is
initialised on the device; W untimed steps, then K timed steps between two
stream synchronisations.

Multi-step runs of this lattice take four time steps per launch (k_stream: each
workgroup walks down a column strip with one wavefront per time level, the
levels' rows kept in LDS), bit-identical to one launch per step;
LBMX_NO_TEMPORAL=1 switches that off.

Also reported in the same JSON line:
  roofline      one-pass bytes of one launch of the dominant kernel (144 B per
                cell in FP64: 9 populations read + 9 written once, the least HBM
                traffic of a launch that advances every cell D steps) / its
                duration, from HIP events on the lattice's stream (the difference
                of two runs of D*m steps), vs the 8 TB/s HBM peak; ``traffic`` is
                the measured HBM bytes per launch from the rocprofv3 FETCH_SIZE /
                WRITE_SIZE summary committed under profiles/, ``limiter`` what
                the rocprofv3 SQ counters say holds the launch back;
                ``secondary.headline_kernel`` has the per-step algorithmic rate
                (D times the one-pass bytes per launch) and the counters
  cpu_baseline  the reference library (oracle/_ref, compiled from lbc's own
                sources with its flags) timed on this host through main_lbm on a
                bounded sample of the same workload
  secondary     FP32-storage MLUPS at 16384^2 (FP32 collision and FP64 collision)
                and both modes' velocity error against FP64 after 2,000 steps
                at 4096^2,
                a plain device copy's GB/s (the two-buffer ceiling), the roofline
                sweep 1024^2 .. 16384^2 in FP64 and FP32, the 256^2 Taylor-Green
                case against the analytic decay,
                the FP64 cylinder channel 2048 x 512 (interpolated bounce-back,
                regularized inlet/outlet), the batched workload (1,024 lattices x
                5,000 steps) and the serving sweep (Poisson arrivals of liner-duct
                requests), each against sequential single-lattice FP64 solves

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--size N]
                    [--precision f64|f32|f32wide]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def tgv_lattice(N, precision):
    from lbc_amd import native
    lat = native.Lattice(N, N, precision=precision, per_x=1, per_y=1)
    lat.set_uniform(1.0 / 0.8)
    lat.set_boundaries([("periodic_x", 0), ("periodic_y", 0)])
    k = 2.0 * math.pi / N
    lat.init_taylor_green(0.01, k, k)
    return lat


def time_lattice(lat, steps, warmup):
    """(wall seconds, device ms from HIP events) for `steps` steps."""
    lat.step(warmup)
    lat.synchronize()
    t0 = time.perf_counter()
    ms = lat.time_steps(steps)
    lat.synchronize()
    return time.perf_counter() - t0, ms


INFINITY_CACHE_BYTES = 256 << 20

# the FP64 gather-form kernel a long run of the periodic lattice launches, by steps
# per launch (lbmx_launch_depth)
FP64_KERNELS = {
    1: "k_pull<double, false>",
    2: "k_pull2<double, false, 8> (two steps per launch, 64 x 8 tiles)",
    3: "k_pull3<double, false, 8> (three steps per launch, 64 x 8 tiles)",
    4: "k_stream<double, false, 4, 64, 3, 3> (four steps per launch, row streaming)",
    5: "k_stream<double, false, 5, 64, 3, 3, true> (five steps per launch, row streaming, "
       "force-free)",
}


def roofline_sweep(sizes=(1024, 2048, 4096, 8192, 16384), seconds=2.0):
    """Periodic Taylor-Green step at each N in FP64 and FP32 storage (FP32
    collision), each point timed for at least `seconds`; moments are not read
    inside the timed region.  Points whose two population buffers fit the 256 MB
    Infinity Cache are labelled: their rate is not an HBM rate.  Two bandwidths:
    the one-pass rate (144 / 72 B per cell per launch: a D-step launch reads and
    writes the lattice once, the least HBM traffic it can have) against the HBM
    peak, and the per-step algorithmic rate (those bytes per cell-update), which
    exceeds the peak once a launch takes several steps."""
    from lbc_amd import native
    rows = []
    for N in sizes:
        for name, prec, bpc in (("fp64", native.F64, 144.0), ("fp32", native.F32_SHIFTED, 72.0)):
            lat = tgv_lattice(N, prec)
            _, probe_ms = time_lattice(lat, 20, 10)
            depth = lat.launch_depth()
            steps = max(50, int(math.ceil(seconds * 1e3 / max(probe_ms / 20, 1e-6))))
            steps = (steps + depth - 1) // depth * depth
            wall, ms = time_lattice(lat, steps, 10)
            lat.close()
            cells = N * N
            gbs = bpc * cells / (ms / steps * 1e-3) / 1e9
            one_pass = gbs / depth
            rows.append({"N": N, "precision": name, "steps": steps, "steps_per_launch": depth,
                         "MLUPS": round(cells * steps / wall / 1e6, 1),
                         "one_pass_GB_per_s": round(one_pass, 1),
                         "one_pass_frac_hbm_peak": round(one_pass / HBM_PEAK_GBS, 4),
                         "per_step_alg_GB_per_s": round(gbs, 1),
                         # two buffers of bpc / 2 bytes per cell
                         "cache_resident": bool(bpc * cells <= INFINITY_CACHE_BYTES)})
            log(f"sweep {name} {N}^2: {rows[-1]['MLUPS']:.0f} MLUPS, one pass {one_pass:.0f} "
                f"GB/s at {depth} steps per launch ({steps} steps"
                f"{', cache resident' if rows[-1]['cache_resident'] else ''})")
    return {"workload": "taylor_green_periodic_NxN_roofline_sweep", "points": rows}


def copy_ceiling(gib=8.0):
    """GB/s (bytes read + written) of a plain device copy of `gib` GiB with 16-byte
    loads and non-temporal stores (lbmx_copy_rate): the practical ceiling a
    two-buffer stencil step is held against."""
    from lbc_amd import native
    return native.copy_rate(gib)


def tgv256_bench(steps=10000):
    """The reference's own Taylor-Green case, 256 x 256, FP64, tau = 0.8, u0 = 0.01,
    `steps` steps: rate and the relative L2 deviation of (u, v) from the decaying
    analytic vortex u0 exp(-2 nu k^2 t)."""
    from lbc_amd import native
    N, u0, tau = 256, 0.01, 0.8
    k, nu = 2.0 * math.pi / N, (tau - 0.5) / 3.0
    lat = tgv_lattice(N, native.F64)
    lat.synchronize()
    t0 = time.perf_counter()
    lat.step(steps)
    lat.synchronize()
    wall = time.perf_counter() - t0
    _, u, v = lat.moments()
    lat.close()
    j, i = np.mgrid[0:N, 0:N].astype(np.float64)
    damp = math.exp(-2.0 * nu * k * k * steps)
    ua = u0 * np.sin(k * i) * np.cos(k * j) * damp
    va = -u0 * np.cos(k * i) * np.sin(k * j) * damp
    err = math.sqrt((((u - ua) ** 2 + (v - va) ** 2).sum()) / ((ua ** 2 + va ** 2).sum()))
    return {"workload": f"taylor_green_periodic_256x256_fp64_{steps}_steps",
            "MLUPS": round(N * N * steps / wall / 1e6, 1), "seconds": round(wall, 3),
            "l2_velocity_error_vs_analytic": float(f"{err:.3e}")}


def fp32_accuracy(N=4096, steps=2000):
    """Relative L2 velocity error of both FP32 storage modes against FP64 on the
    periodic Taylor-Green lattice N^2 after `steps` steps (the same multi-step
    kernels as the speed lines)."""
    from lbc_amd import native

    def vel(prec):
        lat = tgv_lattice(N, prec)
        lat.step(steps)
        _, u, v = lat.moments()
        tot = lat.totals()
        lat.close()
        return u, v, tot

    u64, v64, t64 = vel(native.F64)
    den = float((u64 ** 2 + v64 ** 2).sum())
    out = {"workload": f"taylor_green_periodic_{N}x{N}_{steps}_steps"}
    for prec, key in ((native.F32_SHIFTED, "fp32_collision"), (native.F32_WIDE, "fp64_collision")):
        u, v, t = vel(prec)
        out[key + "_vel_l2_vs_fp64"] = float(
            f"{math.sqrt(float(((u - u64) ** 2 + (v - v64) ** 2).sum()) / den):.3e}")
        out[key + "_mass_rel_diff"] = float(f"{abs(t[0] - t64[0]) / t64[0]:.3e}")
        out[key + "_energy_rel_diff"] = float(f"{abs(t[1] - t64[1]) / t64[1]:.3e}")
    return out


CYL = dict(Nx=2048, Ny=512, R=32.0, xc=512.0, yc=256.0, umax=0.05, Re=100.0)


def cylinder_lattice(precision, start="inflow", kick=0.0, warm=0):
    """Channel past a cylinder, 2048 x 512, Re = 100 on u_max and D (the channel
    configuration).  start="inflow": every fluid node starts at the inlet's
    parabolic profile, as the reference's cylinder driver starts from its inflow
    velocity (cases/cylinder/cylinder.c:56); start="rest": u = 0, whose sudden
    inflow launches a density step of u_max / cs ~ 9 % that rings between the
    reflecting inlet and outlet.  kick: a transverse velocity kick * u_max *
    exp(-r^2 / R^2) centred 3 R behind the cylinder, which breaks the wake's
    symmetry so that shedding starts without waiting for round-off to grow.
    warm: steps run once and undone by uploading the initial state again, so that
    a timed run starts with the launch tables built and the kernels loaded."""
    from lbc_amd import cases as CS_
    from lbc_amd import native
    Nx, Ny, R, xc, yc, umax = (CYL[k] for k in ("Nx", "Ny", "R", "xc", "yc", "umax"))
    nu = umax * 2 * R / CYL["Re"]
    tau = 3.0 * nu + 0.5
    j, i = np.mgrid[0:Ny, 0:Nx]
    mask = ((j == 0) | (j == Ny - 1)).astype(np.int32)
    d = np.sqrt((i - xc) ** 2 + (j - yc) ** 2) - R
    mask[d <= 0] = 1
    phi = np.where((j == 0) | (j == Ny - 1), -1e30, d)
    lat = native.Lattice(Nx, Ny, precision=precision, per_x=0, per_y=0, use_ibb=1)
    lat.set_geometry(mask, np.full((Ny, Nx), 1.0 / tau), phi)
    lat.set_force_body(d <= 0)          # the cylinder alone, besides the walls-included total
    lat.set_boundaries([("velocity_inlet_regularized", 0),
                        ("pressure_outlet_regularized", Nx - 1)])
    y = np.arange(Ny) - 0.5
    H = Ny - 2.0
    uin = np.clip(4 * umax * y * (H - y) / H ** 2, 0, None)
    lat.set_bc_values(0, uin, np.zeros(Ny))
    lat.set_bc_values(1, np.ones(Ny), None)
    rho = np.ones((Ny, Nx))
    zero = np.zeros((Ny, Nx))
    u = np.where(mask == 0, uin[:, None], 0.0) if start == "inflow" else zero
    v = zero
    if kick:
        v = np.where(mask == 0, kick * umax * np.exp(-((i - xc - 3 * R) ** 2 + (j - yc) ** 2)
                                                       / (R * R)), 0.0)
    f = CS_.equilibrium(rho, u, v)
    lat.upload(f, f, rho, u, v)
    if warm:
        lat.step(warm)
        lat.upload(f, f, rho, u, v)
    return lat


def force_stats(body, window_conv=20.0):
    """Force coefficients of the cylinder alone over the last `window_conv`
    convective times D / u_max (cases/cylinder/process.py:105-150: C = F /
    (rho U^2 R), mean and standard deviation, Strouhal number of the lift
    spectrum's peak), with U = u_max; the same with the mean inflow 2/3 u_max is
    given for the drag.  The lift peak is located on an 8x zero-padded spectrum."""
    D, U = 2 * CYL["R"], CYL["umax"]
    T = D / U
    n = len(body)
    w = int(min(n, round(window_conv * T)))
    L, Dr = body[n - w:, 1], body[n - w:, 2]
    q = 0.5 * 1.0 * U * U * D
    CL, CD = L / q, Dr / q
    spec = np.abs(np.fft.rfft(CL - CL.mean(), n=8 * w))
    freq = np.fft.rfftfreq(8 * w, 1.0)
    k = int(np.argmax(spec[1:]) + 1)
    return {"window_steps": w, "window_convective_times": round(w / T, 1),
            "cd_mean": round(float(CD.mean()), 4), "cd_std": round(float(CD.std()), 4),
            "cd_mean_on_mean_inflow": round(float(CD.mean() * 9.0 / 4.0), 4),
            "cl_mean": round(float(CL.mean()), 4), "cl_rms": round(float(CL.std()), 4),
            "strouhal": round(float(freq[k] * D / U), 4)}


def batched_bench(n=1024, steps=5000, sample=48, seed=2024, tune=True, precision=None,
                  pairs=True):
    """The batched workload: n independent lattices (Poiseuille / Couette / cavity /
    Taylor-Green, sizes uniform in {64..1024} x {64..512}) advanced `steps` steps in
    FP32-shifted storage with the collision in FP64 (LBMX_F32_WIDE), one
    HIP-graph-captured launch pair per step, autotuned launch shape.
    Baseline: the same lattices one after another on the single-lattice FP64 path
    (its plain per-step launches), timed on a seeded random sample of `sample`
    lattices and scaled by n/sample (an unbiased estimate of the full sequential
    time); the sample also gives the FP32-vs-FP64 velocity error."""
    from lbc_amd import batch as B
    from lbc_amd import native
    prec = native.F32_WIDE if precision is None else precision
    cases = B.batched_workload(n, steps=steps, seed=seed)
    rng = np.random.default_rng(seed + 1)
    pick = sorted(rng.choice(n, size=min(sample, n), replace=False).tolist())
    for i, c in enumerate(cases):
        c.keep_moments = i in pick
    budget = sum(B.case_bytes(c, prec) for c in cases) + (256 << 20)
    b = B.Batch(prec, budget=budget, chunk=32)
    if pairs:
        b.set_pairs(True)     # two steps per pass (the tuner then keeps its shapes)
    ids = [b.add(c) for c in cases]
    t0 = time.perf_counter()
    tuning = b.tune(trial_steps=3) if tune else b.tuning()
    t_tune = time.perf_counter() - t0
    b.synchronize()
    t0 = time.perf_counter()
    dev_ms = b.time(steps)
    b.synchronize()
    wall = time.perf_counter() - t0
    cells = sum(c.Nx * c.Ny for c in cases)
    gbs = 72.0 * cells * steps / (dev_ms * 1e-3) / 1e9
    log(f"batched {n} lattices x {steps} steps (precision {prec}): {wall:.2f} s wall, "
        f"{dev_ms / 1e3:.2f} s "
        f"device, {n / wall:.1f} solves/s, {cells * steps / wall / 1e6:.0f} MLUPS, "
        f"{gbs:.0f} GB/s, tuning {tuning} ({t_tune:.1f} s)")

    # the divergence guard of every lattice (small lid-driven cavities near tau = 1/2
    # are unstable in FP64 too); diverged lattices are left out of the error figures
    diverged = {}
    for i, c in enumerate(cases):
        if b.status(ids[i]):
            diverged.setdefault(c.kind, []).append(i)
    seq = 0.0
    err_char, err_field = {}, {}
    for i in pick:
        c = cases[i]
        lat = c.single_lattice(native.F64)
        lat.synchronize()
        t1 = time.perf_counter()
        lat.step(c.steps)
        lat.synchronize()
        seq += time.perf_counter() - t1
        r64, u64, v64 = lat.moments()
        lat.close()
        r32, u32, v32 = b.moments(ids[i])
        if i in diverged.get(c.kind, []):
            continue
        fl = ~np.isnan(r64)
        d2 = (u32 - u64)[fl] ** 2 + (v32 - v64)[fl] ** 2
        den = float((u64[fl] ** 2 + v64[fl] ** 2).sum())
        e_char = math.sqrt(d2.mean()) / characteristic_velocity(c)
        err_char.setdefault(c.kind, []).append(e_char)
        err_field.setdefault(c.kind, [])
        if den > 0.0:   # a lattice still at rest has no field-relative error
            err_field[c.kind].append(math.sqrt(d2.sum() / den))
    b.close()
    seq_total = seq * n / len(pick)
    worst = max(max(v) for v in err_char.values())
    log(f"sequential FP64 (sample {len(pick)}): {seq:.2f} s -> {seq_total:.1f} s for {n}, "
        f"{n / seq_total:.2f} solves/s; velocity error vs FP64 (RMS / U_char) max {worst:.2e}; "
        + ", ".join(f"{k}: {max(v):.1e} (field-relative {max(err_field[k], default=0.0):.1e})"
                    for k, v in sorted(err_char.items())))
    return {
        "workload": f"batched_{n}_lattices_{steps}_steps_"
                    + {native.F32_WIDE: "fp32_storage_fp64_collision",
                       native.F32_SHIFTED: "fp32_storage_fp32_collision",
                       native.F64: "fp64"}[prec],
        "solves_per_s": round(n / wall, 2), "wall_s": round(wall, 3),
        "device_s": round(dev_ms / 1e3, 3), "MLUPS": round(cells * steps / wall / 1e6, 1),
        # per-step algorithmic bytes (72 B per cell-update); a pass of pair mode
        # advances two steps, so the one-pass rate is half of it
        "per_step_alg_GB_per_s": round(gbs, 1),
        "one_pass_GB_per_s": round(gbs / (2 if pairs else 1), 1),
        "one_pass_frac_hbm_peak": round(gbs / (2 if pairs else 1) / HBM_PEAK_GBS, 4),
        "cells": cells, "steps_per_pass": 2 if pairs else 1,
        "tuning": None if pairs else {"threads": tuning[0], "tile": tuning[1],
                                      "groups_per_cu": tuning[2], "order": tuning[3]},
        "sequential_fp64_solves_per_s": round(n / seq_total, 3),
        "sequential_sample": len(pick),
        "diverged_lattices_by_family": {k: len(v) for k, v in sorted(diverged.items())},
        "speedup_vs_sequential": round((n / wall) / (n / seq_total), 2),
        "velocity_err_vs_fp64_rms_over_uchar_max": float(f"{worst:.3e}"),
        "velocity_err_vs_fp64_by_family": {
            k: {"rms_over_uchar_max": float(f"{max(v):.3e}"),
                "field_relative_l2_max": float(f"{max(err_field[k], default=0.0):.3e}"), "n": len(v)}
            for k, v in sorted(err_char.items())},
    }


def serving_bench(rates=(10, 30, 100, 300, 1000), n=128, seed=2026, budget_gib=200,
                  precision=None):
    """The serving loop: n seeded liner-duct requests (Nx in [1024, 8192], Ny in
    [64, 256], 2,000-20,000 steps, seeded impedance) arriving as a Poisson stream
    at each rate, admitted by a continuous-batching server (FP32 storage, FP64
    collision) under an HBM budget; each returns a TL spectrum and the FP64 mass
    and energy of its final state.  Throughput = requests / (last completion -
    first arrival); latency = completion - submission, split into the queue wait
    (submission -> admission) and the device residency (admission -> completion).
    Baseline: every one of the same n requests solved one after another with plain
    single-lattice FP64 launches (stepping time only), which also gives each served
    result's TL, mass and energy error."""
    from lbc_amd import batch as B
    from lbc_amd import native
    from lbc_amd import serving as SV
    prec = native.F32_WIDE if precision is None else precision
    reqs = SV.draw_requests(n, seed=seed)
    # launch configuration tuned once on a representative batch
    tb = B.Batch(prec, budget=sum(B.case_bytes(r.case(), prec) for r in reqs[:16]) + (64 << 20))
    for r in reqs[:16]:
        tb.add(r.case())
    tuning = tb.tune(trial_steps=3)
    tb.close()
    try:
        server = B.Server(prec, budget=budget_gib << 30, chunk=64, tuning=tuning)
        budget = budget_gib
    except Exception as e:   # a box with less free HBM: report the budget actually used
        log(f"server with a {budget_gib} GiB budget failed ({e}); using 64 GiB")
        server = B.Server(prec, budget=64 << 30, chunk=64, tuning=tuning)
        budget = 64

    def pct(a):
        return {"p50_s": round(float(np.percentile(a, 50)), 4),
                "p95_s": round(float(np.percentile(a, 95)), 4),
                "p99_s": round(float(np.percentile(a, 99)), 4),
                "mean_s": round(float(np.mean(a)), 4)}

    sweep = []
    served = {}
    for rate in rates:
        arrivals = SV.poisson_arrivals(n, rate, seed=seed + int(rate))
        res, wall = SV.serve(server, reqs, arrivals)
        bad = [r.status for r in res.values() if r.status != 0]
        t0 = min(r.t_submit for r in res.values())
        t1 = max(r.t_done for r in res.values())
        lat = np.array([r.t_done - r.t_submit for r in res.values()])
        wait = np.array([r.t_admit - r.t_submit for r in res.values()])
        resid = np.array([r.t_done - r.t_admit for r in res.values()])
        row = {"rate_per_s": rate, "requests": n, "solves_per_s": round(n / (t1 - t0), 3),
               **pct(lat), "queue_wait": pct(wait), "device_residency": pct(resid),
               "failed": len(bad)}
        sweep.append(row)
        served = res
        log(f"serving rate {rate}/s: {n / (t1 - t0):.2f} solves/s, latency p50 "
            f"{row['p50_s']:.2f} p99 {row['p99_s']:.2f} s (queue p50 "
            f"{row['queue_wait']['p50_s']:.2f} s, residency p50 "
            f"{row['device_residency']['p50_s']:.2f} s), failed {len(bad)}")
    peak = server.stats()["peak_bytes"]
    server.close()

    seq_t, tl_err, mass_err, energy_err = [], [], [], []
    for i, r in enumerate(reqs):
        lat = r.case().single_lattice(native.F64, record_steps=r.steps + 1)
        lat.synchronize()
        t1 = time.perf_counter()
        lat.step(r.steps)
        lat.synchronize()
        seq_t.append(time.perf_counter() - t1)
        hist = lat.probe_history()
        m64, e64 = lat.totals()
        lat.close()
        tl64, up64, _ = SV.tl_from_history(hist, r)
        s32 = served[i]
        tl32 = np.array(s32.tl_db[:SV.NFREQ])
        band = up64 > 1e-2 * up64.max()
        tl_err.append(float(np.abs(tl32 - tl64)[band].max()))
        mass_err.append(abs(s32.mass - m64) / m64)
        energy_err.append(abs(s32.energy - e64) / e64)
        if (i + 1) % 16 == 0:
            log(f"sequential FP64: {i + 1} of {n} requests, {sum(seq_t):.1f} s")
    seq_total = float(np.sum(seq_t))
    seq_rate = n / seq_total
    sat = max(row["solves_per_s"] for row in sweep)
    log(f"sequential FP64, all {n} requests: {seq_total:.1f} s ({seq_rate:.2f} solves/s); served "
        f"max {sat:.2f} solves/s ({sat / seq_rate:.2f}x); TL difference vs FP64 max "
        f"{max(tl_err):.2e} dB; mass {max(mass_err):.1e}, energy {max(energy_err):.1e} relative")
    return {
        "workload": f"serving_liner_requests_{n}_per_rate_"
                    + {native.F32_WIDE: "fp32_storage_fp64_collision",
                       native.F32_SHIFTED: "fp32_storage_fp32_collision"}[prec],
        "budget_GiB": budget, "peak_bytes": int(peak), "chunk_steps": 64,
        "tuning": {"threads": tuning[0], "tile": tuning[1], "groups_per_cu": tuning[2],
                   "order": tuning[3]},
        "sweep": sweep,
        "sequential_fp64_solves_per_s": round(seq_rate, 3), "sequential_sample": n,
        "speedup_vs_sequential_at_saturation": round(sat / seq_rate, 2),
        "tl_max_abs_diff_db_vs_fp64": float(f"{max(tl_err):.3e}"),
        "mass_max_rel_diff_vs_fp64": float(f"{max(mass_err):.3e}"),
        "energy_max_rel_diff_vs_fp64": float(f"{max(energy_err):.3e}"),
    }


def dropin_bench(steps=100):
    """The drop-in boundary as the reference's drivers use it: main_lbm called once
    per step on host arrays with the driver's swap (cases/cylinder/cylinder.c:170-190),
    on the channel past a cylinder 2048 x 512.  The library keeps the lattice on the
    device between calls but must return rho, u, v (and the forces) to the caller
    every step, so the rate is set by the PCIe copies, not by the kernels."""
    from lbc_amd import capi, cases
    Nx, Ny, R, xc, yc, umax = 2048, 512, 32.0, 512.0, 256.0, 0.05
    nu = umax * 2 * R / 100.0
    tau = 3.0 * nu + 0.5
    j, i = np.mgrid[0:Ny, 0:Nx]
    walls = (j == 0) | (j == Ny - 1)
    d = np.sqrt((i - xc) ** 2 + (j - yc) ** 2) - R
    mask = (walls | (d <= 0)).astype(np.int32)
    phi = np.where(walls, -1e30, d)
    omega = np.full((Ny, Nx), 1.0 / tau)
    y = np.arange(Ny) - 0.5
    H = Ny - 2.0
    uin = np.clip(4 * umax * y * (H - y) / H ** 2, 0.0, None)
    bcs = [("velocity_inlet_regularized", 0, uin, np.zeros(Ny)),
           ("pressure_outlet_regularized", Nx - 1, np.ones(Ny), None)]
    rho = np.ones((Ny, Nx))
    u = np.zeros((Ny, Nx))
    v = np.zeros((Ny, Nx))
    f = cases.equilibrium(rho, u, v)
    f_new = f.copy()
    D = capi.DropIn()
    tbl = D.table(bcs)

    def run(n):
        nonlocal f, f_new
        for _ in range(n):
            D.main_lbm(f, f_new, rho, u, v, omega, mask, bcs, 0, 0, 0.0, 0.0, phi, 1, table=tbl)
            f, f_new = f_new, f
    run(5)
    t0 = time.perf_counter()
    run(steps)
    wall = time.perf_counter() - t0
    D.release()
    out = {"workload": f"dropin_main_lbm_cylinder_2048x512_fp64_{steps}_steps",
           "ms_per_step": round(wall / steps * 1e3, 3),
           "MLUPS": round(Nx * Ny * steps / wall / 1e6, 1),
           "host_bytes_per_step": 3 * 8 * Nx * Ny,
           "note": "rho, u, v returned to host arrays every call (the drivers read them)"}
    log(f"drop-in main_lbm: {out}")
    return out


def vtk_output_times(lat, reps=3):
    """One VTK output step of the lattice: the device-packed writer
    (lbmx_write_vtk) against downloading rho/u/v and the host writer; seconds
    (best of reps) and whether the two files are identical."""
    import tempfile
    from lbc_amd import native
    with tempfile.TemporaryDirectory() as d:
        a, b = os.path.join(d, "dev.vtk"), os.path.join(d, "host.vtk")
        td, th = [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            lat.write_vtk(a, 1.0, 1.0, 1.0)
            td.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            rho, u, v = lat.moments()
            native.write_vtk_host(b, 1.0, u, v, rho)
            th.append(time.perf_counter() - t0)
        with open(a, "rb") as fa, open(b, "rb") as fb:
            same = fa.read() == fb.read()
    return {"device_pack_s": round(min(td), 4), "download_and_host_writer_s": round(min(th), 4),
            "identical": same}


def cylinder_bench(steps=50000, shedding=True):
    """Channel past a cylinder, 2048 x 512, Re = 100: FP64 and FP32 storage with the
    FP64 collision, `steps` steps each from the inflow profile with the drag/lift
    history recorded on the device every step (the cylinder driver's {t, L, D}
    record, walls included, and the cylinder alone beside it), and FP64 once more
    from rest.  Reported: MLUPS, the force statistics of the cylinder alone over
    the last 20 convective times (mean C_D, C_L rms, St; references of
    cases/cylinder/process.py: the C_D correlation of unbounded uniform flow, 1.45
    at Re = 100 (:63), St 0.21 (:225), C_L std 0.4), the FP32 history's largest
    deviation from FP64, relative to the FP64 run's largest |D| and |L|, the FP64
    totals, and one VTK output step packed on the device vs on the host."""
    from lbc_amd import native
    out = {"workload": f"cylinder_channel_2048x512_re100_ibb_regularized_{steps}_steps",
           "reference_values": {"cd_mean_correlation_unbounded": 1.449, "strouhal": 0.21,
                                "cl_std": 0.4}}
    rec = {}
    runs = (("fp64", native.F64, "inflow"), ("fp32_fp64collision", native.F32_WIDE, "inflow"),
            ("fp64_from_rest", native.F64, "rest"))
    for name, prec, start in runs:
        lat = cylinder_lattice(prec, start, warm=8)
        lat.set_force_record(steps)
        lat.synchronize()
        t0 = time.perf_counter()
        lat.step(steps)
        lat.synchronize()
        wall = time.perf_counter() - t0
        rec[name] = lat.force_record()
        body = lat.body_force_record()
        stats = force_stats(body)
        mass, energy = lat.totals()
        if name == "fp64":
            out["vtk_output"] = vtk_output_times(lat)
        lat.close()
        out[name] = {"MLUPS": round(2048 * 512 * steps / wall / 1e6, 1),
                     "ms_per_step": round(wall / steps * 1e3, 4), "forces": stats,
                     "mass": mass, "kinetic_energy": energy}
        log(f"cylinder {name}: {out[name]}")
    r64, r32 = rec["fp64"], rec["fp32_fp64collision"]
    n = min(len(r64), len(r32))
    o32 = out["fp32_fp64collision"]
    o32["drag_err"] = float(f"{np.abs(r32[:n, 2] - r64[:n, 2]).max() / np.abs(r64[:n, 2]).max():.3e}")
    o32["lift_err"] = float(f"{np.abs(r32[:n, 1] - r64[:n, 1]).max() / np.abs(r64[:n, 1]).max():.3e}")
    o32["mass_rel_diff"] = float(f"{abs(o32['mass'] - out['fp64']['mass']) / out['fp64']['mass']:.3e}")
    o32["energy_rel_diff"] = float(
        f"{abs(o32['kinetic_energy'] - out['fp64']['kinetic_energy']) / out['fp64']['kinetic_energy']:.3e}")
    out["vs_fp64"] = round(o32["MLUPS"] / out["fp64"]["MLUPS"], 3)
    if not shedding:
        return out
    # shedding needs longer than the configuration's run: FP64 for 200,000 steps
    # (156 convective times) with a wake kick, statistics over the last 20
    long_steps = 200000
    lat = cylinder_lattice(native.F64, "inflow", kick=0.3)
    lat.set_force_record(long_steps)
    lat.step(long_steps)
    out["fp64_shedding_200k_steps"] = {"kick": 0.3, "forces": force_stats(lat.body_force_record())}
    out["fp64_shedding_200k_steps"]["forces"]["strouhal_on_mean_inflow"] = round(
        out["fp64_shedding_200k_steps"]["forces"]["strouhal"] * 1.5, 4)
    lat.close()
    log(f"cylinder shedding run: {out['fp64_shedding_200k_steps']}")
    return out


def liner_bench(steps=20000, seed=2025):
    """The acoustic liner duct, 8192 x 256, tau = 0.51: Gaussian pulse (1e-3 rho0,
    half-width 20, x = 1024) at rest, rigid top wall, impedance wall (mass-spring-
    damper, R = rho0 c0, resonance 1/40 per step, seeded mass M) on x in [3072, 5120] of the
    bottom wall, convective ends with sponges; 20,000 steps in FP64 (bit-identical
    to the oracle) and in both FP32-storage modes.  Reported: time per run, and the
    FP32 modes' largest error, relative to the pulse's p' amplitude, in the wall
    pressure (three wall nodes over time and the whole wall at the end) and in the
    transmitted pressure at x = 6144."""
    from lbc_amd import batch as B
    from lbc_amd import cases as CS_
    from lbc_amd import native
    # seeded reactance slope: the mass drawn uniform in [1, 16] rho0 c0 steps (the
    # serving requests' range) until the ODE is stable under the explicit step
    rng = np.random.default_rng(seed)
    M = float(rng.uniform(1.0, 16.0))
    while not CS_.ade_stable(CS_.LinerSpec(R=1.0, f_res=1.0 / 40.0, M=M).impedance()):
        M = float(rng.uniform(1.0, 16.0))
    imp = CS_.LinerSpec(R=1.0, f_res=1.0 / 40.0, M=M).impedance()
    probes = [(3584, 0), (4096, 0), (4608, 0), (6144, 128)]
    out = {"workload": f"liner_duct_8192x256_tau0.51_{steps}_steps",
           "impedance": {"R_over_rho0c0": 1.0, "f_res_per_step": 1.0 / 40.0,
                         "M_over_rho0c0": round(M, 4), "seed": seed}}
    runs = {}
    for name, prec in (("fp64", native.F64), ("fp32_fp32collision", native.F32_SHIFTED),
                       ("fp32_fp64collision", native.F32_WIDE)):
        c = B.liner_duct(8192, 256, steps, imp, tau=0.51, amp=1e-3, x0=1024.0,
                         halfwidth=20.0, liner=(3072, 5120), sponge=192, probes=probes)
        b = B.Batch(prec, budget=B.case_bytes(c, prec) + (64 << 20))
        b.set_pairs(True)     # two steps per pass, bit-identical to one step per launch
        i = b.add(c)
        b.synchronize()
        ms = b.time(steps)
        hist = b.probes(i)
        rho, _, _ = b.moments(i)
        mass, energy = b.totals(i)
        b.close()
        runs[name] = (hist, rho)
        out[name] = {"s": round(ms / 1e3, 3), "MLUPS": round(8192 * 256 * steps / ms / 1e3, 1),
                     "mass": mass, "kinetic_energy": energy}
    cs2 = 1.0 / 3.0
    scale = 1e-3 * cs2                       # the pulse's initial p' amplitude
    h64, r64 = runs["fp64"]
    wall64 = cs2 * (r64[0, 3072:5121] - 1.0)
    for name in ("fp32_fp32collision", "fp32_fp64collision"):
        h, r = runs[name]
        out[name]["wall_p_series_err"] = float(f"{cs2 * np.abs(h[:, :3] - h64[:, :3]).max() / scale:.3e}")
        out[name]["wall_p_final_err"] = float(
            f"{np.abs(cs2 * (r[0, 3072:5121] - 1.0) - wall64).max() / scale:.3e}")
        out[name]["transmitted_err"] = float(f"{cs2 * np.abs(h[:, 3] - h64[:, 3]).max() / scale:.3e}")
        for k in ("mass", "kinetic_energy"):
            out[name][k + "_rel_diff"] = float(
                f"{abs(out[name][k] - out['fp64'][k]) / abs(out['fp64'][k]):.3e}")
    out["transmitted_peak_over_incident"] = float(
        f"{cs2 * np.abs(h64[:, 3] - 1.0).max() / (scale / 2):.4f}")
    log(f"liner duct: {out}")
    return out


def characteristic_velocity(c):
    """Velocity scale of a batched case: the wall speed (Couette, cavity), the
    Taylor-Green amplitude, or for the force-driven channel the smaller of the
    steady centre-line speed F H^2 / (8 nu) and the speed F t reached from rest."""
    if c.kind in ("couette", "cavity"):
        return c.top_velocity[0]
    if c.kind == "tgv":
        return c.u0
    nu = (c.tau - 0.5) / 3.0
    H = c.Ny - 2
    return min(c.Fx * H * H / (8.0 * nu), c.Fx * c.steps)


def load_traffic(tag):
    """HBM bytes per launch of the fused kernel from the committed pmc summary."""
    p = os.path.join(ROOT, "profiles", f"pmc_{tag}.json")
    if not os.path.exists(p):
        return None
    try:
        with open(p) as fh:
            d = json.load(fh)
        return float(d["hbm_bytes_per_launch"])
    except Exception:
        return None


def load_issue(tag):
    """VALU / LDS issue counters of the fused kernel from the committed rocprofv3
    SQ-counter summary (profiles/sq_<tag>.json), or None."""
    p = os.path.join(ROOT, "profiles", f"sq_{tag}.json")
    if not os.path.exists(p):
        return None
    try:
        with open(p) as fh:
            return json.load(fh)
    except Exception:
        return None


def host_cpu():
    """Model, sockets and physical cores of this host (lscpu), and the CPUs this
    process may run on."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except Exception:
        pass
    def num(key):
        try:
            return int(info.get(key, ""))
        except ValueError:
            return None
    sockets, cps, tpc = num("Socket(s)"), num("Core(s) per socket"), num("Thread(s) per core")
    try:
        allowed = len(os.sched_getaffinity(0))
    except Exception:
        allowed = os.cpu_count()
    return {"model": info.get("Model name"), "sockets": sockets,
            "physical_cores": sockets * cps if sockets and cps else None,
            "threads_per_core": tpc, "logical_cpus": num("CPU(s)"),
            "cpus_allowed": allowed}


def cpu_baseline(N, min_seconds=10.0, one_thread_seconds=8.0):
    """The reference library's main_lbm on a bounded sample of the TGV workload
    (oracle/cpu_bench.py, a child process per point so that the OpenMP binding takes
    effect): the box's OpenMP thread count (OMP_NUM_THREADS, the host share this
    process is given) bound close to one thread per physical core, and one thread."""
    import subprocess
    threads = int(os.environ.get("OMP_NUM_THREADS", 0) or 0) or (os.cpu_count() or 1)
    host = host_cpu()
    if host["physical_cores"]:
        threads = min(threads, host["physical_cores"])

    def point(nthreads, seconds):
        env = dict(os.environ, OMP_NUM_THREADS=str(nthreads), OMP_PROC_BIND="close",
                   OMP_PLACES="cores")
        r = subprocess.run([sys.executable, "-m", "oracle.cpu_bench", "--size", str(N),
                            "--seconds", str(seconds)], cwd=ROOT, env=env,
                           capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            raise RuntimeError(f"cpu_bench failed ({r.returncode}): {r.stderr[-500:]}")
        return json.loads(r.stdout.strip().splitlines()[-1])

    full = point(threads, min_seconds)
    one = point(1, one_thread_seconds)
    how = ("built from lbc's own src/ with its flags (gcc -O3 -fopenmp) by oracle/Makefile in "
           "the build container and carried to this host with the repository snapshot"
           if full["kind"] == "reference" else "this project's C restatement of the step")
    return {"value": round(full["MLUPS"], 2), "unit": "MLUPS", "cores": threads,
            "kind": full["kind"],
            "sample": f"main_lbm FP64 periodic Taylor-Green {N}x{N}, {full['steps']} steps "
                      f"({full['seconds']:.1f} s) on {threads} OpenMP threads "
                      f"(OMP_PROC_BIND=close OMP_PLACES=cores); one thread: "
                      f"{one['steps']} steps ({one['seconds']:.1f} s)",
            "one_thread_MLUPS": round(one["MLUPS"], 2),
            "scaling_vs_one_thread": round(full["MLUPS"] / one["MLUPS"], 2),
            "binary": f"{full['binary']} ({how})",
            "binding": "OMP_PROC_BIND=close OMP_PLACES=cores",
            "host": host}


def _section(name, fn, *args, **kw):
    """A secondary workload whose failure (a host-side error) is reported in its
    place instead of ending the run before the headline line is printed."""
    try:
        return fn(*args, **kw)
    except Exception as e:   # reported, not fatal: the headline number stands
        log(f"{name} failed: {type(e).__name__}: {e}")
        return {"error": f"{type(e).__name__}: {e}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--size", type=int, default=16384)
    ap.add_argument("--precision", choices=("f64", "f32", "f32wide"), default="f64",
                    help="storage/arithmetic of the headline lattice (default f64, the "
                         "reference's arithmetic; f32 = FP32 storage and collision, f32wide = "
                         "FP32 storage, FP64 collision) -- for profiling one kernel")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--cpu-size", type=int, default=2048)
    ap.add_argument("--batched", type=int, default=1024, metavar="N",
                    help="lattices of the batched workload (0: skip)")
    ap.add_argument("--batched-steps", type=int, default=5000)
    ap.add_argument("--batched-sample", type=int, default=48,
                    help="lattices timed one by one for the sequential FP64 baseline")
    ap.add_argument("--batched-single", action="store_true",
                    help="batched workload at one step per pass instead of pair mode")
    ap.add_argument("--serving", type=int, default=128, metavar="N",
                    help="requests per arrival rate of the serving sweep (0: skip)")
    ap.add_argument("--dropin", type=int, default=100, metavar="STEPS",
                    help="main_lbm calls of the drop-in benchmark (0: skip)")
    ap.add_argument("--cylinder-steps", type=int, default=50000)
    ap.add_argument("--cylinder-only", action="store_true",
                    help="only run the channel past a cylinder (FP64, FP32 storage, FP64 from "
                         "rest; no 200,000-step shedding run) and print its JSON -- for "
                         "profiling its kernels")
    ap.add_argument("--no-sweep", dest="sweep", action="store_false",
                    help="skip the Taylor-Green roofline sweep 1024^2 .. 16384^2")
    ap.add_argument("--liner", type=int, default=20000, metavar="STEPS",
                    help="steps of the 8192 x 256 liner duct in FP64 and FP32 (0: skip)")
    ap.add_argument("--calibrate-copy", type=float, default=0.0, metavar="GIB",
                    help="only run one plain device-to-device copy of GIB GiB (torch) and "
                         "exit: the known-size reference for scaling rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE readings")
    args = ap.parse_args()
    if args.calibrate_copy > 0:
        import torch
        n = int(args.calibrate_copy * (1 << 30)) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibration_copy_bytes": 8 * n}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU "
              "slab decomposition is a design, not built", file=sys.stderr)
        sys.exit(2)

    from lbc_amd import native
    N = args.size
    cells = N * N
    log(f"device: {native.device_name()}")
    if args.cylinder_only:
        print(json.dumps(cylinder_bench(args.cylinder_steps, shedding=False)), flush=True)
        return

    prec, bpc, ptag, pname = {
        "f64": (native.F64, 144, "fp64", "FP64"),
        "f32": (native.F32_SHIFTED, 72, "fp32", "FP32-storage"),
        "f32wide": (native.F32_WIDE, 72, "fp32wide", "FP32-storage FP64-collision")}[args.precision]
    lat = tgv_lattice(N, prec)
    wall, dev_ms = time_lattice(lat, args.steps, args.warmup)
    # the dominant kernel and its launch time: once the lattice runs in gather form a
    # run of D*m steps is exactly m D-step launches, so the difference of two such
    # runs over the difference of m is one launch, from HIP events on the lattice's
    # stream
    steps_per_launch = lat.launch_depth()
    temporal = steps_per_launch > 1
    if temporal:
        m_a, m_b = 40, 10
        t_a = lat.time_steps(steps_per_launch * m_a)
        t_b = lat.time_steps(steps_per_launch * m_b)
        launch_ms = (t_a - t_b) / (m_a - m_b)
    else:
        launch_ms = dev_ms / args.steps
    kernel = (FP64_KERNELS.get(steps_per_launch, f"{steps_per_launch} steps per launch")
              if prec == native.F64 else f"{pname} gather form, {steps_per_launch} steps per launch")
    lat.close()
    mlups = cells * args.steps / wall / 1e6
    # roofline: the least HBM traffic of one launch is one read and one write of the
    # lattice (bpc B per cell), whatever its steps per launch; per-step algorithmic
    # bytes (bpc per cell-update) go in secondary, since a D-step launch moves only
    # a D-th of them
    achieved = float(bpc) * cells / (launch_ms * 1e-3) / 1e9
    per_step_alg = achieved * steps_per_launch
    ptag_launch = f"tgv{N}_{ptag}" + (f"_{steps_per_launch}step" if temporal else "")
    traffic = load_traffic(ptag_launch)
    issue = load_issue(ptag_launch)
    log(f"{pname} {N}^2: {wall / args.steps * 1e3:.3f} ms/step wall, "
        f"{dev_ms / args.steps:.3f} ms/step events, {mlups:.0f} MLUPS; {kernel}: "
        f"{launch_ms:.3f} ms per launch, one pass {achieved:.0f} GB/s "
        f"({achieved / HBM_PEAK_GBS:.1%} of HBM peak); per-step algorithmic {per_step_alg:.0f} GB/s")

    secondary = {}
    if not args.no_secondary:
        # the FP32 runs time whole launches of their own depth (eight steps for FP32
        # collision), the timed steps rounded up to a multiple of it
        lat = tgv_lattice(N, native.F32_SHIFTED)
        spl32 = lat.launch_depth()
        n32 = -(-args.steps // spl32) * spl32
        w32, d32 = time_lattice(lat, n32, args.warmup)
        lat.close()
        m32 = cells * n32 / w32 / 1e6
        secondary["fp32_shifted_tgv"] = {
            "workload": f"taylor_green_periodic_{N}x{N}_fp32_storage_fp32_collision",
            "MLUPS": round(m32, 1), "steps": n32,
            "one_pass_GB_per_s": round(72.0 * cells / (d32 / n32 * spl32 * 1e-3) / 1e9, 1),
            "per_step_alg_GB_per_s": round(72.0 * cells / (d32 / n32 * 1e-3) / 1e9, 1),
            "steps_per_launch": spl32,
            "vs_fp64": round(m32 / mlups, 3)}
        log(f"FP32-shifted {N}^2: {m32:.0f} MLUPS ({m32 / mlups:.2f}x FP64)")
        lat = tgv_lattice(N, native.F32_WIDE)
        splw = lat.launch_depth()
        nw = -(-args.steps // splw) * splw
        ww, dw = time_lattice(lat, nw, args.warmup)
        lat.close()
        mw = cells * nw / ww / 1e6
        secondary["fp32_wide_tgv"] = {
            "workload": f"taylor_green_periodic_{N}x{N}_fp32_storage_fp64_collision",
            "MLUPS": round(mw, 1), "steps": nw,
            "one_pass_GB_per_s": round(72.0 * cells / (dw / nw * splw * 1e-3) / 1e9, 1),
            "per_step_alg_GB_per_s": round(72.0 * cells / (dw / nw * 1e-3) / 1e9, 1),
            "steps_per_launch": splw,
            "vs_fp64": round(mw / mlups, 3)}
        log(f"FP32 storage, FP64 collision {N}^2: {mw:.0f} MLUPS ({mw / mlups:.2f}x FP64)")
        secondary["fp32_accuracy"] = _section("fp32_accuracy", fp32_accuracy)
        log(f"FP32 storage accuracy: {secondary['fp32_accuracy']}")
        copy = copy_ceiling()
        hbm = traffic / (launch_ms * 1e-3) / 1e9 if traffic else None
        secondary["copy_ceiling"] = {
            "GB_per_s": round(copy, 1),
            "fp64_kernel_hbm_frac_of_copy": round(hbm / copy, 4) if hbm else None}
        log(f"copy ceiling {copy:.0f} GB/s"
            + (f"; the FP64 kernel's measured traffic runs at {hbm / copy:.1%} of it" if hbm else ""))
        if args.sweep:
            secondary["roofline_sweep"] = _section("roofline_sweep", roofline_sweep)
        secondary["tgv256"] = _section("tgv256", tgv256_bench)
        log(f"TGV 256^2: {secondary['tgv256']}")
        secondary["cylinder"] = _section("cylinder", cylinder_bench, args.cylinder_steps)
    if args.batched > 0:
        secondary["batched"] = _section("batched", batched_bench, args.batched,
                                        args.batched_steps, args.batched_sample,
                                        pairs=not args.batched_single)
    if args.serving > 0:
        secondary["serving"] = _section("serving", serving_bench, n=args.serving)
    if args.dropin > 0:
        secondary["dropin"] = _section("dropin", dropin_bench, args.dropin)
    if args.liner > 0:
        secondary["liner_duct"] = _section("liner_duct", liner_bench, args.liner)

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(args.cpu_size)
            log(f"cpu baseline: {cpu}")
        except Exception as e:   # reported, not fatal: the GPU number stands
            log(f"cpu baseline failed: {e}")

    secondary["headline_kernel"] = {
        "kernel": kernel, "steps_per_launch": steps_per_launch,
        "launch_ms": round(launch_ms, 4),
        "one_pass_bytes_per_launch": bpc * cells,
        "per_step_algorithmic_bytes_per_launch": bpc * cells * steps_per_launch,
        "per_step_alg_GB_per_s": round(per_step_alg, 1),
        "hbm_bytes_per_launch": traffic,
        "hbm_GB_per_s": round(traffic / (launch_ms * 1e-3) / 1e9, 1) if traffic else None,
        "issue": issue,
        "note": f"{steps_per_launch} steps per HBM pass: the roofline line counts one read "
                "and one write of the lattice per launch; the launch also runs "
                f"{steps_per_launch} collisions per cell, and 'issue' gives the VALU and LDS "
                "counters of the same launch" if temporal else ""}
    out = {
        "metric": f"MLUPS ({pname} fused D2Q9 BGK step) + % HBM roofline",
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f64": "f64", "f32": "f32", "f32wide": "f32 storage, f64 arithmetic"}[args.precision],
        "data": "synthetic (Taylor-Green vortex generated on the device)",
        "config": {"workload": f"taylor_green_periodic_{N}x{N}_{ptag}", "Nx": N, "Ny": N,
                   "tau": 0.8, "u0": 0.01, "lattice": "D2Q9 BGK", "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "limiter": (issue or {}).get("limiter")},
        "cpu_baseline": cpu,
        "secondary": secondary,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
