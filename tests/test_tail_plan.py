"""The launch plan of the work after streaming (lbmx_tail_plan), checked on the CPU.

lbmx_step applies the boundary table in launches whose entries run concurrently,
and puts the stale-edge bookkeeping (which rewrites S and the lagged moments at
the non-periodic domain edges) into the last launch when it can.  That is only
correct if the entries of one launch commute and none of the edge launch's
entries reads S or the lagged moments at an edge node.  Both properties are
checked here with the oracle's own boundary routines (orc_apply_bc) on random
states: every launch applied in reverse entry order must give the table-order
result bit for bit, and perturbing f / rho / u / v at the edge nodes must not
change what the edge launch's entries write.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from cases import all_scenarios
from lbc_amd import native

ROW_TYPES = {"velocity_top", "velocity_bottom", "velocity_top_regularized",
             "velocity_bottom_regularized", "pressure_top_regularized",
             "pressure_bottom_regularized", "convective_top", "convective_bottom"}


def values(name, Nx, Ny, rng):
    n = Nx if name in ROW_TYPES else Ny
    if name.startswith("pressure"):
        return 1.0 + 0.01 * rng.standard_normal(n), None
    return 0.02 * rng.standard_normal(n), 0.02 * rng.standard_normal(n)


def apply(Nx, Ny, entries, f, fn, rho, u, v, mask):
    for name, index, v1, v2 in entries:
        tbl = O.make_bcs([(name, index, v1, v2)])
        O.ORACLE.orc_apply_bc(Nx, Ny, tbl, O._p(f), O._p(fn), O._p(rho), O._p(u), O._p(v),
                              O._p(mask))


def edge_mask(Nx, Ny, per_x, per_y, mask):
    e = np.zeros((Ny, Nx), dtype=bool)
    if not per_y:
        e[0, :] = e[-1, :] = True
    if not per_x:
        e[:, 0] = e[:, -1] = True
    return e & (mask == 0)


def check_plan(Nx, Ny, per_x, per_y, mask, bcs, seed=0):
    """Asserts the plan's two properties; returns (nlaunch, launch_of, edge_launch)."""
    n, launch_of, edge_launch = native.tail_plan(Nx, Ny, per_x, per_y, mask, bcs)
    rng = np.random.default_rng(seed)
    kept = [b for b in range(len(bcs)) if launch_of[b] >= 0]
    # leading periodic entries of a wrapping axis are the only ones left out
    for b in range(len(bcs)):
        if launch_of[b] < 0:
            assert all(launch_of[c] < 0 for c in range(b))
            assert bcs[b][0] == ("periodic_x" if per_x else "") or \
                bcs[b][0] == ("periodic_y" if per_y else "")
    # table order kept, launches numbered 0, 1, ... without gaps
    seq = [launch_of[b] for b in kept]
    assert seq == sorted(seq)
    assert set(seq) | ({edge_launch} if edge_launch >= 0 else set()) == set(range(n))
    assert (edge_launch >= 0) == (not (per_x and per_y))
    if seq and edge_launch >= 0:
        assert edge_launch >= seq[-1]

    vals = [values(name, Nx, Ny, rng) for name, _ in bcs]
    entries = [(bcs[b][0], bcs[b][1], *vals[b]) for b in range(len(bcs))]
    f = rng.random((Ny, Nx, 9)) + 0.05
    fn0 = rng.random((Ny, Nx, 9)) + 0.05
    rho = 1.0 + 0.05 * rng.standard_normal((Ny, Nx))
    u = 0.03 * rng.standard_normal((Ny, Nx))
    v = 0.03 * rng.standard_normal((Ny, Nx))
    m = np.ascontiguousarray(mask, dtype=np.int32)

    ref = fn0.copy()
    apply(Nx, Ny, [entries[b] for b in kept], f, ref, rho, u, v, m)
    got = fn0.copy()
    for launch in range(n):
        members = [entries[b] for b in kept if launch_of[b] == launch]
        apply(Nx, Ny, members[::-1], f, got, rho, u, v, m)
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), "launch entries do not commute"

    if edge_launch >= 0:
        before = fn0.copy()
        apply(Nx, Ny, [entries[b] for b in kept if launch_of[b] < edge_launch], f, before, rho,
              u, v, m)
        members = [entries[b] for b in kept if launch_of[b] == edge_launch]
        e = edge_mask(Nx, Ny, per_x, per_y, mask)
        a = before.copy()
        apply(Nx, Ny, members, f, a, rho, u, v, m)
        f2, rho2, u2, v2 = f.copy(), rho.copy(), u.copy(), v.copy()
        f2[e] += 0.25
        rho2[e] += 0.1
        u2[e] += 0.01
        v2[e] -= 0.01
        b2 = before.copy()
        apply(Nx, Ny, members, f2, b2, rho2, u2, v2, m)
        assert np.array_equal(a.view(np.uint64), b2.view(np.uint64)), \
            "an entry of the edge launch reads S at an edge node"
    return n, launch_of, edge_launch


@pytest.mark.parametrize("sc", all_scenarios(1), ids=lambda s: s.name)
def test_scenario_plans(sc):
    check_plan(sc.Nx, sc.Ny, sc.per_x, sc.per_y, sc.mask, sc.bc_names, seed=sc.Nx)


def test_channel_is_one_launch():
    """Cylinder channel (solid walls, regularized inlet and outlet): both columns and
    the edge job in a single launch."""
    Nx, Ny = 64, 24
    mask = np.zeros((Ny, Nx), dtype=np.int32)
    mask[0, :] = mask[-1, :] = 1
    j, i = np.mgrid[0:Ny, 0:Nx]
    mask[(i - 16) ** 2 + (j - 12) ** 2 <= 16] = 1
    bcs = [("velocity_inlet_regularized", 0), ("pressure_outlet_regularized", Nx - 1)]
    n, launch_of, edge = check_plan(Nx, Ny, 0, 0, mask, bcs)
    assert (n, launch_of, edge) == (1, [0, 0], 0)


def test_corners_split_launches():
    """All four sides without solid walls: each column meets each row at a corner,
    so a row entry cannot share a launch with a column entry."""
    Nx, Ny = 20, 20
    bcs = [("velocity_inlet_regularized", 0), ("velocity_outlet_regularized", Nx - 1),
           ("velocity_bottom_regularized", 0), ("velocity_top_regularized", Ny - 1)]
    n, launch_of, edge = check_plan(Nx, Ny, 0, 0, np.zeros((Ny, Nx), np.int32), bcs)
    assert launch_of[0] == launch_of[1] and launch_of[2] != launch_of[1]
    assert n >= 2


BC_CHOICES = [
    ("velocity_inlet", "col0"), ("pressure_inlet", "col0"), ("velocity_outlet", "colN"),
    ("pressure_outlet", "colN"), ("velocity_top", "rowN"), ("velocity_bottom", "row0"),
    ("velocity_inlet_regularized", "col0"), ("velocity_outlet_regularized", "colN"),
    ("velocity_top_regularized", "rowN"), ("velocity_bottom_regularized", "row0"),
    ("pressure_inlet_regularized", "col0"), ("pressure_outlet_regularized", "colN"),
    ("pressure_top_regularized", "rowN"), ("pressure_bottom_regularized", "row0"),
    ("convective_inlet", "col0"), ("convective_outlet", "colN"), ("convective_top", "rowN"),
    ("convective_bottom", "row0"), ("periodic_x", "p"), ("periodic_y", "p"),
]


@pytest.mark.parametrize("seed", range(40))
def test_random_tables(seed):
    rng = np.random.default_rng(seed)
    Nx, Ny = int(rng.integers(8, 30)), int(rng.integers(8, 30))
    bcs = []
    for _ in range(int(rng.integers(1, 7))):
        name, where = BC_CHOICES[int(rng.integers(len(BC_CHOICES)))]
        off = int(rng.integers(0, 3))
        index = {"col0": off, "colN": Nx - 1 - off, "row0": off, "rowN": Ny - 1 - off,
                 "p": 0}[where]
        bcs.append((name, index))
    names = [b[0] for b in bcs]
    per_x, per_y = int("periodic_x" in names), int("periodic_y" in names)
    mask = np.zeros((Ny, Nx), dtype=np.int32)
    if rng.random() < 0.5:
        mask[0, :] = 1
    if rng.random() < 0.3:
        mask[:, -1] = 1
    if rng.random() < 0.5:
        j, i = np.mgrid[0:Ny, 0:Nx]
        mask[(i - Nx / 2) ** 2 + (j - Ny / 2) ** 2 <= (min(Nx, Ny) / 5) ** 2] = 1
    check_plan(Nx, Ny, per_x, per_y, mask, bcs, seed=seed)


def test_invalid_arguments():
    with pytest.raises(native.LbmxError):
        native.tail_plan(2, 10, 0, 0, None, [("velocity_inlet", 0)])
    with pytest.raises(native.LbmxError):
        native.tail_plan(10, 10, 0, 0, None, [("velocity_inlet", 12)])
