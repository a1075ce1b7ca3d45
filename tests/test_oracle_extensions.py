"""Checks for the oracle functions that no reference golden covers: the
diagonal and non-default-orientation boundary rules, the moving wall, the
adaptive time step and the red-black solver.  Each check comes from what the
physics or the mathematics fixes, not from the oracle's own formulas:
wall invariants, closed forms, and a manufactured Poisson problem."""
import math

import numpy as np
import pytest

from oracle import pyoracle as O
from stroemung_amd import presets
from tests.gridgen import mixed_grid

N_, S_ = (0, -1), (0, 1)
E_, W_ = (1, 0), (-1, 0)
NORTH_T = (1, 2, 8)     # N, NE, NW
SOUTH_T = (4, 5, 6)     # SE, S, SW
EAST_T = (2, 3, 4)      # NE, E, SE
WEST_T = (6, 7, 8)      # SW, W, NW


def _sim(ct, vel, seed=0, **kw):
    rng = np.random.default_rng(seed)
    nx, ny = ct.shape
    base = dict(delx=0.1, dely=0.1, delt=0.01, gamma=0.9, reynolds=100.0,
                sor_absolute_epsilon=1e-10, max_iterations=100000, omega=1.7)
    base.update(kw)
    s = O.Sim(ct, vel, u=rng.uniform(-1, 1, (nx, ny)), v=rng.uniform(-1, 1, (nx, ny)),
              p=rng.uniform(-1, 1, (nx, ny)), **base)
    assert s.call("rebuild_boundary_list") == 0
    return s


@pytest.mark.parametrize("seed", range(6))
def test_setbcond_wall_invariants(seed):
    """After SETBCOND, at every fluid/wall interface of a NoSlip or Inflow
    cell the normal face velocity equals the wall's, and for straight edges
    the ghost/fluid average of the tangential velocity equals it (negated
    ghost); for a moving wall the average equals the wall velocity."""
    ct, vel = mixed_grid(100 + seed, 40, 30)
    s = _sim(ct, vel, seed)
    u0, v0 = s.u.copy(), s.v.copy()
    assert s.call("set_boundary_u_and_v") == 0
    u, v, edge = s.u, s.v, s.edge
    # The ghost is formed from the fluid value at the time the cell is
    # visited: N, S and E neighbours still hold their pre-pass values then
    # (their only writers come later in (x, y) order); the W neighbour's v
    # may already have been set by the wall cell at (x-1, y+1).
    checked = 0
    for x in range(ct.shape[0]):
        for y in range(ct.shape[1]):
            t, e = int(ct[x, y]), int(edge[x, y])
            if t in (O.FLUID, O.OUTFLOW) or e <= 0:
                continue
            bu, bv = (vel[x, y] if t in (O.INFLOW, O.MOVING_WALL) else (0.0, 0.0))
            # normal faces
            if e in NORTH_T:
                assert v[x, y - 1] == bv
            if e in WEST_T:
                assert u[x - 1, y] == bu
            if e in EAST_T:
                assert u[x, y] == bu
            if e in SOUTH_T and e != 5:
                assert v[x, y] == bv
            if e == 5:
                assert v[x, y] == bv
            # tangential ghosts of straight edges
            mv = t == O.MOVING_WALL
            if e == 1:   # North: fluid above
                got = u[x, y] + u0[x, y - 1]
                assert got == pytest.approx(2 * bu, abs=1e-15) if mv else got == 0.0
            if e == 5:
                got = u[x, y] + u0[x, y + 1]
                assert got == pytest.approx(2 * bu, abs=1e-15) if mv else got == 0.0
            if e == 3:
                got = v[x, y] + v0[x + 1, y]
                assert got == pytest.approx(2 * bv, abs=1e-15) if mv else got == 0.0
            if e == 7:
                got = v[x, y] + v[x - 1, y]
                assert got == pytest.approx(2 * bv, abs=1e-15) if mv else got == 0.0
            checked += 1
    assert checked > 50


def test_setbcond_outflow_zero_gradient():
    """An outflow cell copies the fluid values across each of its fluid faces."""
    for seed in range(4):
        ct, vel = mixed_grid(200 + seed, 36, 28)
        s = _sim(ct, vel, seed)
        u0, v0 = s.u.copy(), s.v.copy()
        assert s.call("set_boundary_u_and_v") == 0
        for x in range(ct.shape[0]):
            for y in range(ct.shape[1]):
                if ct[x, y] != O.OUTFLOW or s.edge[x, y] <= 0:
                    continue
                e = int(s.edge[x, y])
                # u from the N/S fluid neighbour for N/S kinds, else E/W; v mirrored
                src_u = {1: N_, 2: N_, 3: E_, 4: E_, 5: S_, 6: W_, 7: W_, 8: N_}[e]
                src_v = {1: N_, 2: E_, 3: E_, 4: S_, 5: S_, 6: S_, 7: W_, 8: W_}[e]
                assert s.u[x, y] == u0[x + src_u[0], y + src_u[1]]
                if e in (7, 8):   # v[W] may have been set by the wall below-left first
                    assert s.v[x, y] == s.v[x - 1, y]
                else:
                    assert s.v[x, y] == v0[x + src_v[0], y + src_v[1]]


def test_ghost_pressure_is_neighbour_mean():
    ct, vel = mixed_grid(7, 40, 30)
    s = _sim(ct, vel)
    assert s.call("copy_pressure_to_boundaries") == 0
    p = s.p
    src = {1: [N_], 2: [N_, E_], 3: [E_], 4: [S_, E_], 5: [S_], 6: [S_, W_], 7: [W_], 8: [N_, W_]}
    n = 0
    for x in range(ct.shape[0]):
        for y in range(ct.shape[1]):
            e = int(s.edge[x, y])
            if e <= 0:
                continue
            vals = [p[x + dx, y + dy] for dx, dy in src[e]]
            assert p[x, y] == sum(vals) / len(vals)
            n += 1
    assert n > 50


def test_max_uv_definition():
    ct, vel = mixed_grid(9, 30, 22)
    s = _sim(ct, vel)
    um, vm = s.max_uv()
    fl = ct == O.FLUID
    cand_u = list(np.abs(s.u[fl]))
    cand_v = list(np.abs(s.v[fl]))
    for x, y in zip(*np.nonzero(fl)):
        if x > 0 and not fl[x - 1, y]:
            cand_u.append(abs(s.u[x - 1, y]))
        if y > 0 and not fl[x, y - 1]:
            cand_v.append(abs(s.v[x, y - 1]))
    assert um == max(cand_u) and vm == max(cand_v)


def test_comp_delt_closed_form():
    ct, vel = presets.lid_cavity(8, 8)
    s = _sim(ct, vel, delx=0.5, dely=0.25, reynolds=10.0, tau=0.5, delt=123.0)
    diff = (10.0 / 2.0) / (1 / 0.25 + 1 / 0.0625)   # = 0.25
    assert diff == 0.25
    f = O.lib().orc_comp_delt
    assert f(s.s, 0.0, 0.0) == 0.5 * 0.25              # diffusive limit
    assert f(s.s, 4.0, 0.0) == 0.5 * (0.5 / 4.0)       # dx/umax = 0.125 < 0.25
    assert f(s.s, 0.1, 2.0) == 0.5 * (0.25 / 2.0)      # dy/vmax = 0.125
    s.set(tau=0.0)
    assert f(s.s, 4.0, 2.0) == 123.0                   # fixed step


def _neumann_laplacian(p, fluid, dx, dy):
    """5-point Laplacian with the mirror (zero-gradient) boundary: a missing
    neighbour takes the value of the cell itself."""
    nx, ny = p.shape
    out = np.zeros_like(p)
    for x in range(1, nx - 1):
        for y in range(1, ny - 1):
            if not fluid[x, y]:
                continue
            def nb(a, b):
                return p[a, b] if fluid[a, b] else p[x, y]
            out[x, y] = ((nb(x + 1, y) - 2 * p[x, y] + nb(x - 1, y)) / dx ** 2
                         + (nb(x, y + 1) - 2 * p[x, y] + nb(x, y - 1)) / dy ** 2)
    return out


@pytest.mark.parametrize("solver", [O.SOLVER_RB, O.SOLVER_LEX])
def test_manufactured_neumann_poisson(solver):
    """rhs = L p* for a smooth p* on a box with straight walls; the solver must
    return p* up to a constant (pure Neumann problem)."""
    ct, vel = presets.lid_cavity(24, 16)
    nx, ny = ct.shape
    dx, dy = 1 / 24, 1 / 16
    fl = ct == O.FLUID
    X, Y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    pstar = np.cos(math.pi * (X - 0.5) / 24) * np.cos(2 * math.pi * (Y - 0.5) / 16) + 0.3 * X / 24
    rhs = _neumann_laplacian(pstar, fl, dx, dy)
    s = _sim(ct, vel, delx=dx, dely=dy, omega=1.7, solver=solver, check_every=1,
             sor_absolute_epsilon=1e-13, masked_stop=1)
    s.p[...] = 0.0
    s.rhs[...] = rhs
    it, n2 = s.solve()
    assert it < 100000
    got = s.p[fl] - s.p[fl].mean()
    want = pstar[fl] - pstar[fl].mean()
    assert np.abs(got - want).max() < 1e-9 * np.abs(want).max()


def test_red_black_and_lexicographic_same_fixed_point():
    ct, vel = mixed_grid(31, 34, 26)
    fl = ct == O.FLUID
    rng = np.random.default_rng(3)
    rhs = rng.standard_normal(ct.shape)
    rhs[fl] -= rhs[fl].mean()           # compatible with the pure Neumann problem
    sols = []
    for solver in (O.SOLVER_RB, O.SOLVER_LEX):
        s = _sim(ct, vel, solver=solver, check_every=4, sor_absolute_epsilon=1e-12,
                 masked_stop=1, omega=1.8)
        s.p[...] = 0.0
        s.rhs[...] = np.where(fl, rhs, 0.0)
        s.solve()
        sols.append(s.p[fl] - s.p[fl].mean())
    scale = np.abs(sols[0]).max()
    assert np.abs(sols[0] - sols[1]).max() < 1e-8 * scale


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_threaded_red_black_sweeps_equal_the_serial_ones(threads):
    """The multi-core CPU comparison line: each half-sweep split over threads
    gives the serial red-black sweeps bit for bit (a half-sweep's cells read
    only the other colour)."""
    ct, vel = mixed_grid(17, 45, 38)
    rng = np.random.default_rng(5)
    runs = []
    for mt in (False, True):
        s = _sim(ct, vel, solver=O.SOLVER_RB, omega=1.7)
        s.p[...] = rng.standard_normal(ct.shape) if not runs else runs[0][1]
        s.rhs[...] = rng.standard_normal(ct.shape) if not runs else runs[0][2]
        p0, r0 = s.p.copy(), s.rhs.copy()
        assert s.call("rebuild_boundary_list") == 0
        s.set(rhs_mean=0.25)
        if mt:
            s.call("rb_sweeps_mt", 9, threads)
        else:
            s.call("rb_sweeps", 9)
        runs.append((s.p.copy(), p0, r0))
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    assert not np.array_equal(runs[0][0], runs[0][1])    # the sweeps changed p
