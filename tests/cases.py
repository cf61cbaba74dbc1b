"""Seeded lattice scenarios shared by the oracle, drop-in and HIP parity tests.

Every scenario is a plain description: sizes, solid mask, signed distance phi,
per-node omega, initial f / f_new / rho / u / v (deliberately NOT consistent with
each other, so the lagged-moment and stale-entry semantics of the reference step
are exercised), a boundary table and periodicity flags.  Boundary values change
every step (``vals(step)``), as cases/nit/nit.c:303 and cases/testbc/testbc.c:170
do.

Boundary tables use the reference's function names; the periodicity flags follow
the drivers' rule (a flag is set when periodic_x / periodic_y is in the table,
cases/nit/nit.c:186-192).
"""
from __future__ import annotations

import numpy as np

CX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
CY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
W = np.array([4.0 / 9.0] + [1.0 / 9.0] * 4 + [1.0 / 36.0] * 4)


def feq(rho, u, v):
    """Second-order equilibrium, [Ny, Nx, 9]."""
    cu = CX[None, None, :] * u[..., None] + CY[None, None, :] * v[..., None]
    usq = (u * u + v * v)[..., None]
    return W * rho[..., None] * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq)


class Scenario:
    def __init__(self, name, Nx, Ny, bcs, seed=0, use_ibb=0, Fx=0.0, Fy=0.0):
        self.name, self.Nx, self.Ny = name, Nx, Ny
        self.rng = np.random.default_rng(seed)
        self.bc_names = bcs            # list of (name, index)
        self.use_ibb = use_ibb
        self.Fx, self.Fy = Fx, Fy
        self.mask = np.zeros((Ny, Nx), dtype=np.int32)
        self.phi = np.zeros((Ny, Nx))
        self.omega = np.full((Ny, Nx), 1.0 / 0.8)
        names = [b[0] for b in bcs]
        self.per_x = int("periodic_x" in names)
        self.per_y = int("periodic_y" in names)
        self._val_seed = seed + 1000

    # ---- geometry helpers ---------------------------------------------------
    def solid_rows(self, *rows):
        for j in rows:
            self.mask[j, :] = 1
            self.phi[j, :] = -1e30
        return self

    def solid_cols(self, *cols):
        for i in cols:
            self.mask[:, i] = 1
            self.phi[:, i] = -1e30
        return self

    def disk(self, xc, yc, r):
        j, i = np.mgrid[0:self.Ny, 0:self.Nx]
        d = np.sqrt((i - xc) ** 2 + (j - yc) ** 2) - r
        inside = d <= 0
        self.mask[inside] = 1
        keep = self.phi > -1e29
        self.phi[keep] = d[keep]
        return self

    def random_omega(self, lo=1.0 / 1.2, hi=1.0 / 0.55):
        self.omega = self.rng.uniform(lo, hi, (self.Ny, self.Nx))
        return self

    # ---- state --------------------------------------------------------------
    def initial_state(self, amp=0.02, noise=1e-3):
        """(f, f_new, rho, u, v): f near equilibrium, everything else perturbed."""
        r = self.rng
        Ny, Nx = self.Ny, self.Nx
        rho = 1.0 + amp * r.standard_normal((Ny, Nx))
        u = amp * r.standard_normal((Ny, Nx))
        v = amp * r.standard_normal((Ny, Nx))
        f = feq(rho, u, v) * (1.0 + noise * r.standard_normal((Ny, Nx, 9)))
        f_new = feq(rho, u, v) * (1.0 + noise * r.standard_normal((Ny, Nx, 9)))
        # lagged arrays not equal to the moments of f (as after a driver IC)
        rho_a = rho * (1.0 + noise * r.standard_normal((Ny, Nx)))
        u_a = u + noise * amp * r.standard_normal((Ny, Nx))
        v_a = v + noise * amp * r.standard_normal((Ny, Nx))
        return (np.ascontiguousarray(f), np.ascontiguousarray(f_new), rho_a, u_a, v_a)

    def vals(self, step):
        """Per-step boundary value arrays: list of (name, index, val1, val2)."""
        r = np.random.default_rng(self._val_seed + step)
        out = []
        for name, index in self.bc_names:
            along_col = any(s in name for s in ("inlet", "outlet"))
            n = self.Ny if along_col else self.Nx
            if name.startswith("pressure"):
                v1 = 1.0 + 0.01 * r.standard_normal(n)
                v2 = np.zeros(n)
            elif name.startswith("velocity"):
                v1 = 0.03 * r.standard_normal(n)
                v2 = 0.03 * r.standard_normal(n)
            else:
                v1 = v2 = None
            out.append((name, index, v1, v2))
        return out


def all_scenarios(scale=1):
    """The parity battery.  scale multiplies the lattice sizes."""
    s = scale
    sc = []
    sc.append(Scenario("tgv_periodic", 24 * s, 16 * s, [("periodic_x", 0), ("periodic_y", 0)], 1))
    sc.append(Scenario("poiseuille_bb_force", 20 * s, 13 * s, [("periodic_x", 0)], 2,
                       Fx=2e-5).solid_rows(0, 13 * s - 1))
    sc.append(Scenario("couette_top", 18 * s, 12 * s,
                       [("periodic_x", 0), ("velocity_top", 12 * s - 1)], 3).solid_rows(0))
    sc.append(Scenario("ldc_walls", 16 * s, 16 * s, [("velocity_top", 16 * s - 1)], 4)
              .solid_rows(0).solid_cols(0, 16 * s - 1))
    sc.append(Scenario("cylinder_ibb", 40 * s, 20 * s,
                       [("periodic_y", 0), ("velocity_inlet", 0),
                        ("convective_outlet", 40 * s - 1)], 5, use_ibb=1)
              .disk(10.3 * s, 9.7 * s, 3.4 * s))
    sc.append(Scenario("channel_regularized_ibb", 36 * s, 18 * s,
                       [("velocity_inlet_regularized", 0),
                        ("pressure_outlet_regularized", 36 * s - 1)], 6, use_ibb=1)
              .solid_rows(0, 18 * s - 1).disk(12.2 * s, 8.6 * s, 3.1 * s))
    sc.append(Scenario("testbc_regularized", 20 * s, 20 * s,
                       [("velocity_inlet_regularized", 0),
                        ("velocity_outlet_regularized", 20 * s - 1),
                        ("velocity_bottom_regularized", 0),
                        ("velocity_top_regularized", 20 * s - 1)], 7).random_omega())
    sc.append(Scenario("nit_like", 30 * s, 9 * s,
                       [("periodic_y", 0), ("pressure_inlet", 0),
                        ("velocity_outlet_regularized", 30 * s - 1)], 8))
    sc.append(Scenario("zouhe_box", 17 * s, 15 * s,
                       [("velocity_inlet", 0), ("velocity_outlet", 17 * s - 1),
                        ("velocity_bottom", 0), ("velocity_top", 15 * s - 1)], 9))
    sc.append(Scenario("pressure_zouhe", 22 * s, 12 * s,
                       [("pressure_inlet", 0), ("pressure_outlet", 22 * s - 1)], 10)
              .solid_rows(0, 12 * s - 1))
    sc.append(Scenario("convective_all", 19 * s, 14 * s,
                       [("convective_inlet", 0), ("convective_outlet", 19 * s - 1),
                        ("convective_bottom", 0), ("convective_top", 14 * s - 1)], 11))
    sc.append(Scenario("pressure_reg_rows", 21 * s, 13 * s,
                       [("periodic_x", 0), ("pressure_bottom_regularized", 0),
                        ("pressure_top_regularized", 13 * s - 1)], 12))
    sc.append(Scenario("inner_index_bcs", 26 * s, 14 * s,
                       [("periodic_y", 0), ("pressure_inlet_regularized", 3),
                        ("velocity_outlet", 26 * s - 4),
                        ("velocity_bottom_regularized", 0),
                        ("pressure_top_regularized", 14 * s - 1)], 13).random_omega())
    sc.append(Scenario("late_periodic_overwrite", 16 * s, 15 * s,
                       [("velocity_bottom", 0), ("velocity_top", 15 * s - 1),
                        ("periodic_x", 0)], 14, Fx=1e-5, Fy=-2e-5).disk(8.1 * s, 7.2 * s, 2.3 * s))
    sc.append(Scenario("bb_periodic_both_force", 23 * s, 17 * s,
                       [("periodic_x", 0), ("periodic_y", 0)], 15, Fx=3e-5, Fy=1e-5)
              .disk(11.5 * s, 8.5 * s, 4.2 * s).random_omega())
    sc.append(Scenario("ibb_periodic_walls", 27 * s, 15 * s,
                       [("periodic_x", 0), ("periodic_y", 0)], 16, use_ibb=1, Fx=1e-5)
              .solid_rows(0).disk(13.3 * s, 7.4 * s, 3.6 * s))
    return sc
