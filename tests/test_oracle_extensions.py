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


# ------------------------------------------------ diagonal kinds and restore list

NE_, SW_, NW_ = 2, 6, 8


def _wall_velocity(t, vel, x, y):
    return tuple(vel[x, y]) if t in (O.INFLOW, O.MOVING_WALL) else (0.0, 0.0)


def test_setbcond_diagonal_tangential_ghosts():
    """The tangential ghost of each diagonal kind (src/grid/mod.rs:446-488,
    544-586): NorthEast v[b] = -v[E], SouthWest u[b] = -u[S], NorthWest
    u[b] = -u[N] and v[b] = -v[W] for NoSlip and Inflow (the sum of ghost and
    fluid value is exactly 0); 2 * wall velocity for the moving wall.  The
    N, S and E sources hold their pre-pass values when the cell is visited;
    v[W] may already have been set by the wall cell at (x-1, y+1).  A rule
    reading any other neighbour fails: the fields are random."""
    seen = {}
    for seed in range(8):
        ct, vel = mixed_grid(300 + seed, 44, 34, n_rect=10)
        s = _sim(ct, vel, seed)
        u0, v0 = s.u.copy(), s.v.copy()
        assert s.call("set_boundary_u_and_v") == 0
        u, v, edge = s.u, s.v, s.edge
        for x in range(ct.shape[0]):
            for y in range(ct.shape[1]):
                t, e = int(ct[x, y]), int(edge[x, y])
                if t in (O.FLUID, O.OUTFLOW) or e not in (NE_, SW_, NW_):
                    continue
                bu, bv = _wall_velocity(t, vel, x, y)
                mv = t == O.MOVING_WALL
                sums = []
                if e == NE_:
                    sums.append((v[x, y] + v0[x + 1, y], bv))
                if e == SW_:
                    sums.append((u[x, y] + u0[x, y + 1], bu))
                if e == NW_:
                    sums.append((u[x, y] + u0[x, y - 1], bu))
                    sums.append((v[x, y] + v[x - 1, y], bv))
                for got, w in sums:
                    if mv:
                        assert got == pytest.approx(2 * w, abs=1e-15), (seed, x, y, e)
                    else:
                        assert got == 0.0, (seed, x, y, e, t)
                seen[(e, t)] = seen.get((e, t), 0) + 1
    for e in (NE_, SW_, NW_):
        for t in (O.NOSLIP, O.INFLOW, O.MOVING_WALL):
            assert seen.get((e, t), 0) > 0, ("kind never exercised", e, t)


def test_restore_list_net_effect_after_adap_uv():
    """src/simulation.rs:287-322 with the list of src/grid/mod.rs:595-648.
    After ADAP_UV every cell of [0, nx-2] x [0, ny-2] that is not a boundary
    cell holds the plain ADAP_UV value (fluid faces written by SETBCOND are
    not reset).  A boundary cell holds what SETBCOND left in it, except that
    a cell whose kind has a West side (SW, W, NW) receives u of its west
    fluid face and one with a North side (N, NE, NW) v of its north fluid
    face, as SETBCOND left them (the list carries the face's value under the
    boundary cell's own index).  A restore written to the face, or the
    wrong face's value, fails."""
    checked = {}
    for seed in range(6):
        ct, vel = mixed_grid(400 + seed, 42, 31, n_rect=10)
        s = _sim(ct, vel, seed, delt=0.0173, delx=0.31, dely=0.23)
        rng = np.random.default_rng(seed + 50)
        s.f[...] = rng.uniform(-1, 1, ct.shape)
        s.g[...] = rng.uniform(-1, 1, ct.shape)
        assert s.call("set_boundary_u_and_v") == 0
        us, vs = s.u.copy(), s.v.copy()
        f, g, p = s.f.copy(), s.g.copy(), s.p.copy()
        s.call("set_u_and_v")
        dt = s.s.delt
        eu, ev = us.copy(), vs.copy()
        eu[:-1, :-1] = f[:-1, :-1] - ((dt / 0.31) * (p[1:, :-1] - p[:-1, :-1]))
        ev[:-1, :-1] = g[:-1, :-1] - ((dt / 0.23) * (p[:-1, 1:] - p[:-1, :-1]))
        for x in range(ct.shape[0]):
            for y in range(ct.shape[1]):
                if ct[x, y] == O.FLUID:
                    continue
                e = int(s.edge[x, y])
                eu[x, y] = us[x - 1, y] if e in (SW_, 7, NW_) else us[x, y]
                ev[x, y] = vs[x, y - 1] if e in (1, NE_, NW_) else vs[x, y]
                checked[e] = checked.get(e, 0) + 1
        np.testing.assert_array_equal(s.u.view(np.uint64), eu.view(np.uint64))
        np.testing.assert_array_equal(s.v.view(np.uint64), ev.view(np.uint64))
    assert all(checked.get(e, 0) > 0 for e in range(9)), checked


def _ranges_brute(ct, u, v, p):
    """The reference's folds (src/grid/mod.rs:237-268): over fluid cells,
    accumulators start at (f64::MAX, 0.0); the speed range takes sqrt at the
    end."""
    fl = ct == O.FLUID
    big = np.finfo(np.float64).max
    if fl.any():
        sp = u[fl] * u[fl] + v[fl] * v[fl]
        pr = (min(big, float(p[fl].min())), max(0.0, float(p[fl].max())))
        sr = (math.sqrt(min(big, float(sp.min()))), math.sqrt(max(0.0, float(sp.max()))))
    else:
        pr, sr = (big, 0.0), (math.sqrt(big), 0.0)
    return pr, sr


@pytest.mark.parametrize("case", ["mixed", "all_negative_p", "all_positive_p", "no_fluid",
                                  "one_fluid_cell"])
def test_ranges_brute_force(case):
    """calculate_pressure_range / calculate_speed_range against a numpy
    brute force, including an all-negative p (max stays at the 0.0 start)
    and an all-positive p (min comes from the data, not f64::MAX)."""
    ct, vel = mixed_grid(77, 36, 28)
    if case == "no_fluid":
        ct = np.full_like(ct, O.NOSLIP)
    if case == "one_fluid_cell":
        ct = np.full_like(ct, O.NOSLIP)
        ct[10, 10] = O.FLUID
    rng = np.random.default_rng(9)
    nx, ny = ct.shape
    s = O.Sim(ct, vel, delx=0.1, dely=0.1, delt=0.01, gamma=0.9, reynolds=100.0,
              sor_absolute_epsilon=1e-3, max_iterations=10, omega=1.7,
              u=rng.uniform(-2, 2, (nx, ny)), v=rng.uniform(-2, 2, (nx, ny)),
              p=rng.uniform(-5, 5, (nx, ny)))
    if case == "all_negative_p":
        s.p[...] = -rng.uniform(1, 5, (nx, ny))
    if case == "all_positive_p":
        s.p[...] = rng.uniform(1, 5, (nx, ny))
    s.call("calculate_pressure_range")
    s.call("calculate_speed_range")
    pr, sr = _ranges_brute(ct, s.u, s.v, s.p)
    assert tuple(s.s.pressure_range) == pr
    assert tuple(s.s.speed_range) == sr
    if case == "all_negative_p":
        assert pr[1] == 0.0 and pr[0] < -1.0
    if case == "all_positive_p":
        assert pr[0] >= 1.0
