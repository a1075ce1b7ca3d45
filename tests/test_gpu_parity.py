"""Parity of the HIP path (through the C ABI) with the reference's goldens and
the CPU oracle.  Bit-exact for the stencil, boundary and velocity-update
stages and for the reference-compatible lexicographic solve; red-black
sweeps are bit-exact against the oracle's red-black sweeps."""
import sys

import numpy as np
import pytest

from oracle import pyoracle as O
from stroemung_amd import presets
from stroemung_amd.simulation import Context, Simulation, make_params
from tests.gridgen import mixed_grid, random_fields
from tests.refdata import arr, assert_bits_equal, from_json, golden, preset, sim_kwargs

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _pair(ct, vel, fields=None, **kw):
    """A one-channel GPU context and an oracle simulation on the same inputs."""
    kw.setdefault("solver", O.SOLVER_LEX)
    solver = kw["solver"]
    okw = dict(kw)
    okw["solver"] = solver
    osim = O.Sim(ct, vel, **okw)
    p = make_params(kw["delx"], kw["dely"], kw["delt"], kw["gamma"], kw["reynolds"],
                    kw["sor_absolute_epsilon"], kw["max_iterations"], kw["omega"],
                    kw.get("initial_norm_squared"), 0, 0.0, kw.get("tau", 0.0),
                    "rb" if solver == O.SOLVER_RB else "lex", kw.get("check_every", 1))
    ctx = Context(0)
    ch = ctx.add_grid(p, ct, vel)
    if fields is not None:
        u, v, pp, f, g, rhs = fields
        for name, a in zip(("u", "v", "p", "f", "g", "rhs"), fields):
            getattr(osim, name)[...] = a
        ctx.upload(ch, u=u, v=v, p=pp, f=f, g=g, rhs=rhs)
    return ctx, ch, osim


def _cmp(ctx, ch, osim, names, what):
    got = ctx.download(ch)
    for n in names:
        assert_bits_equal(got[n], getattr(osim, n), "%s %s" % (what, n))


# ------------------------------------------------------------ reference goldens

def test_tick_trajectory_4x3_matches_reference_snapshots():
    say("test: 4x3 trajectory vs reference snapshots")
    T = golden("tick_4x3.json")
    setup = T["setup"]
    ct, vel = preset(setup["preset"], setup["size"])
    dx, dy = setup["cell_size"]
    sim = Simulation(setup["size"], (dx, dy), setup["delt"], setup["gamma"], setup["reynolds"],
                     setup["sor_absolute_epsilon"], setup["max_iterations"], setup["omega"],
                     ct, vel)
    assert sim.initial_norm_squared == 0.0

    def check(snap, what, with_fgr=True):
        if with_fgr:
            assert_bits_equal(sim.f, arr(snap["f"]), what + " f")
            assert_bits_equal(sim.g, arr(snap["g"]), what + " g")
            assert_bits_equal(sim.rhs, arr(snap["rhs"]), what + " rhs")
        g = snap["sim"]["grid"]
        assert_bits_equal(sim.pressure, arr(g["pressure"]), what + " p")
        assert_bits_equal(sim.u, arr(g["u"]), what + " u")
        assert_bits_equal(sim.v, arr(g["v"]), what + " v")
        assert sim.iterations == snap["sim"]["iterations"]
        assert sim.time == snap["sim"]["time"]

    it, n2 = sim.run_simulation_tick()
    assert (it, n2) == (T["after_1"]["sor_iterations"], T["after_1"]["norm_squared"])
    check(T["after_1"], "tick1")
    for _ in range(100):
        it, n2 = sim.run_simulation_tick()
    assert (it, n2) == (T["after_101"]["sor_iterations"], T["after_101"]["norm_squared"])
    check(T["after_101"], "tick101")
    for _ in range(100):
        sim.run_simulation_tick()
    check(T["after_201"], "tick201", with_fgr=False)


def test_construct_from_nast2d_data():
    say("test: construct from NaSt2D-derived JSON")
    d = golden("simulation_io.json")["small_simulation_with_boundaries"]
    sim = Simulation.from_reader(d["input"])
    assert sim.initial_norm_squared == 899.9547140394143
    exp = d["expected"]["grid"]
    assert_bits_equal(sim.pressure, arr(exp["pressure"]), "p")
    assert_bits_equal(sim.u, arr(exp["u"]), "u")


# ------------------------------------------------------------ oracle, full ticks

C1 = dict(delx=0.1, dely=0.1, delt=0.005, gamma=0.9, reynolds=100.0,
          sor_absolute_epsilon=1e-3, max_iterations=100, omega=1.7)


def test_c1_square_obstacle_lex_100_ticks_bit_exact():
    say("test: C1 square obstacle, 100 ticks, lexicographic")
    ct, vel = presets.square_obstacle()
    ctx, ch, osim = _pair(ct, vel, **C1)
    ctx.construct()
    osim.construct()
    assert ctx.state(ch)["initial_norm_squared"] == osim.s.initial_norm_squared
    for step in range(100):
        its, nrm = ctx.tick()
        oit, on2 = osim.tick()
        assert (its[ch], nrm[ch]) == (oit, on2), "tick %d" % step
        if step % 20 == 19 or step == 0:
            _cmp(ctx, ch, osim, ("u", "v", "p", "f", "g", "rhs"), "tick %d" % step)
    st = ctx.state(ch)
    assert st["speed_range"] == list(osim.s.speed_range)
    assert st["pressure_range"] == list(osim.s.pressure_range)
    assert st["time"] == osim.s.time


# Grid sizes (ring included) for each form of the lexicographic kernel: in
# LDS (at most 4 608 cells) swept by one wave with a lane per row (at most 64
# interior rows), by one wave with a lane per cell of a diagonal (at most 64
# interior columns), or by the workgroup (both sides over 64); and in global
# memory (more than 4 608 cells) with long and with short diagonals.
@pytest.mark.parametrize("seed,nx,ny", [(1, 37, 29), (2, 37, 29), (6, 30, 125), (3, 67, 68),
                                        (4, 92, 72), (5, 302, 22)])
def test_mixed_boundaries_lex_ticks_bit_exact(seed, nx, ny):
    say("test: mixed boundary types, lexicographic ticks, seed", seed, nx, ny)
    ct, vel = mixed_grid(seed, nx, ny)
    kw = dict(C1, max_iterations=30, sor_absolute_epsilon=1e-9)
    ctx, ch, osim = _pair(ct, vel, **kw)
    ctx.construct()
    osim.construct()
    for step in range(8):
        its, nrm = ctx.tick()
        oit, on2 = osim.tick()
        assert (its[ch], nrm[ch]) == (oit, on2), "tick %d" % step
        _cmp(ctx, ch, osim, ("u", "v", "p", "f", "g", "rhs"), "tick %d" % step)


def test_adaptive_delt_lid_cavity_bit_exact():
    say("test: adaptive delt on a lid cavity")
    ct, vel = presets.lid_cavity(40, 30)
    kw = dict(delx=1 / 40, dely=1 / 30, delt=0.0, gamma=0.9, reynolds=1000.0,
              sor_absolute_epsilon=1e-4, max_iterations=50, omega=1.7, tau=0.5)
    ctx, ch, osim = _pair(ct, vel, **kw)
    ctx.construct()
    osim.construct()
    for step in range(6):
        its, nrm = ctx.tick()
        oit, on2 = osim.tick()
        assert ctx.state(ch)["delt"] == osim.s.delt
        assert (its[ch], nrm[ch]) == (oit, on2)
        _cmp(ctx, ch, osim, ("u", "v", "p"), "tick %d" % step)


# ------------------------------------------------------------ stages, random data

@pytest.mark.parametrize("seed,nx,ny", [(3, 70, 45), (4, 131, 200), (5, 1026, 1026)])
def test_stages_bit_exact_random_fields(seed, nx, ny):
    say("test: stages on random fields", seed, nx, ny)
    if nx == 1026:
        ct, vel = presets.lid_cavity(nx - 2, ny - 2)
    else:
        ct, vel = mixed_grid(seed, nx, ny)
    fields = random_fields(seed, (nx, ny))
    kw = dict(delx=0.37, dely=0.21, delt=0.013, gamma=0.8, reynolds=250.0,
              sor_absolute_epsilon=1e-3, max_iterations=10, omega=1.6)
    ctx, ch, osim = _pair(ct, vel, fields=fields, **kw)
    assert osim.call("rebuild_boundary_list") == 0
    ctx.stage("set_boundary_u_and_v")
    assert osim.call("set_boundary_u_and_v") == 0
    _cmp(ctx, ch, osim, ("u", "v"), "SETBCOND")
    ctx.stage("calculate_f_and_g")
    osim.call("calculate_f_and_g")
    _cmp(ctx, ch, osim, ("f", "g"), "COMP_FG")
    ctx.stage("calculate_rhs")
    osim.call("calculate_rhs")
    _cmp(ctx, ch, osim, ("rhs",), "COMP_RHS")
    # a fresh pressure, then the velocity update with the restore list
    p = np.random.default_rng(seed + 100).uniform(-3, 3, (nx, ny))
    ctx.upload(ch, p=p)
    osim.p[...] = p
    ctx.stage("set_u_and_v")
    osim.call("set_u_and_v")
    _cmp(ctx, ch, osim, ("u", "v"), "ADAP_UV")
    ctx.stage("copy_pressure_to_boundaries")
    osim.call("copy_pressure_to_boundaries")
    _cmp(ctx, ch, osim, ("p",), "ghost p")


N_VARIANTS = 10


@pytest.mark.parametrize("variant", range(N_VARIANTS))
@pytest.mark.parametrize("seed,nx,ny", [(6, 50, 33), (7, 300, 140)])
def test_rb_sweeps_bit_exact(seed, nx, ny, variant):
    say("test: red-black sweeps", seed, nx, ny, "variant", variant)
    ct, vel = mixed_grid(seed, nx, ny)
    fields = random_fields(seed, (nx, ny))
    kw = dict(delx=0.3, dely=0.2, delt=0.01, gamma=0.9, reynolds=100.0,
              sor_absolute_epsilon=1e-6, max_iterations=50, omega=1.7, solver=O.SOLVER_RB,
              check_every=0)
    ctx, ch, osim = _pair(ct, vel, fields=fields, **kw)
    assert ctx.solver_variant(variant)[0] == variant
    assert osim.call("rebuild_boundary_list") == 0
    ctx.stage("rb_sweeps", 7)
    # the sweeps use rhs - mean; the GPU's mean is a tree sum, the oracle's a
    # sequential one: equal to rounding, and the sweeps are compared on the
    # GPU's value
    mean = ctx.state(ch)["rhs_mean"]
    assert mean == pytest.approx(O.lib().orc_rhs_mean(osim.s), rel=1e-12, abs=1e-15)
    osim.set(rhs_mean=mean)
    osim.call("rb_sweeps", 7)
    _cmp(ctx, ch, osim, ("p",), "7 sweeps")
    ctx.stage("rb_sweeps", 12)
    osim.call("rb_sweeps", 12)
    _cmp(ctx, ch, osim, ("p",), "19 sweeps")


# Grid sides (boundary ring included) on either side of the sizes where the
# blocked tiling goes from one tile per axis to two: a single tile's region
# starts outside the channel on both sides (x: side <= region - 2; y, in
# aligned pairs: side <= region - 3), two tiles use the edge-side room.  The
# regions are 32, 64 or 96 cells in x and 64, 96 or 128 in y.
@pytest.mark.parametrize("variant", range(N_VARIANTS))
@pytest.mark.parametrize("nx,ny", [(30, 125), (31, 126), (62, 61), (63, 62), (94, 93),
                                   (95, 94)])
def test_rb_sweeps_bit_exact_at_tile_edges(nx, ny, variant):
    say("test: red-black sweeps at tile edges", nx, ny, "variant", variant)
    seed = 20 + nx
    ct, vel = mixed_grid(seed, nx, ny)
    fields = random_fields(seed, (nx, ny))
    kw = dict(delx=0.3, dely=0.2, delt=0.01, gamma=0.9, reynolds=100.0,
              sor_absolute_epsilon=1e-6, max_iterations=50, omega=1.7, solver=O.SOLVER_RB,
              check_every=0)
    ctx, ch, osim = _pair(ct, vel, fields=fields, **kw)
    assert ctx.solver_variant(variant)[0] == variant
    assert osim.call("rebuild_boundary_list") == 0
    ctx.stage("rb_sweeps", 13)
    osim.set(rhs_mean=ctx.state(ch)["rhs_mean"])
    osim.call("rb_sweeps", 13)
    _cmp(ctx, ch, osim, ("p",), "13 sweeps")


def test_rb_solve_matches_oracle_iterations():
    say("test: red-black solve, iterations vs oracle")
    ct, vel = presets.lid_cavity(48, 48)
    kw = dict(delx=1 / 48, dely=1 / 48, delt=1e-3, gamma=0.9, reynolds=1000.0,
              sor_absolute_epsilon=1e-8, max_iterations=5000, omega=1.8,
              solver=O.SOLVER_RB, check_every=4)
    ctx, ch, osim = _pair(ct, vel, **kw)
    ctx.construct()
    osim.construct()
    for step in range(3):
        its, nrm = ctx.tick()
        oit, on2 = osim.tick()
        assert abs(its[ch] - oit) <= 4, (its[ch], oit)
        st = ctx.state(ch)
        assert st["residual_squared"] <= st["tolerance_squared"]
        got = ctx.download(ch)
        scale = np.abs(osim.p).max() + 1e-300
        assert np.abs(got["p"] - osim.p).max() / scale < 1e-9


def test_batch_equals_single_channels():
    say("test: batched channels equal single runs")
    grids = [mixed_grid(11, 40, 25), presets.lid_cavity(33, 47), presets.square_obstacle(60, 18)]
    kw = dict(delx=0.05, dely=0.05, delt=2e-3, gamma=0.9, reynolds=200.0,
              sor_absolute_epsilon=1e-6, max_iterations=400, omega=1.7)
    params = make_params(kw["delx"], kw["dely"], kw["delt"], kw["gamma"], kw["reynolds"],
                         kw["sor_absolute_epsilon"], kw["max_iterations"], kw["omega"],
                         solver="rb", check_every=2)
    batch = Context(0)
    for ct, vel in grids:
        batch.add_grid(params, ct, vel)
    batch.construct()
    singles = []
    for ct, vel in grids:
        c = Context(0)
        c.add_grid(params, ct, vel)
        c.construct()
        singles.append(c)
    for step in range(4):
        its, nrm = batch.tick()
        for k, c in enumerate(singles):
            sit, snrm = c.tick()
            assert (its[k], nrm[k]) == (sit[0], snrm[0]), (step, k)
    for k, c in enumerate(singles):
        a, b = batch.download(k), c.download(0)
        for n in a:
            assert_bits_equal(a[n], b[n], "channel %d %s" % (k, n))


def test_compacted_launches_equal_single_channels():
    """Two large channels with a loose tolerance stop long before a small one
    with a tight tolerance, so later launches cover only the small channel's
    tiles (compaction); every channel must still equal its own single run, bit
    for bit, iterations included."""
    say("test: compacted blocked launches equal single runs")
    # omega = 1 (Gauss-Seidel) and a tight tolerance keep the small channel
    # sweeping long after the large ones have met their loose tolerance
    grids = [(presets.lid_cavity(120, 90), 0.3, 1.8), (mixed_grid(21, 110, 96), 0.3, 1.8),
             (presets.lid_cavity(40, 30), 1e-10, 1.0)]
    def params(eps, omega):
        return make_params(1.0 / 90, 1.0 / 90, 2e-3, 0.9, 300.0, eps, 50000, omega,
                           solver="rb", check_every=4)
    batch = Context(0)
    for (ct, vel), eps, om in grids:
        batch.add_grid(params(eps, om), ct, vel)
    batch.construct()
    singles = []
    for (ct, vel), eps, om in grids:
        c = Context(0)
        c.add_grid(params(eps, om), ct, vel)
        c.construct()
        singles.append(c)
    for step in range(3):
        its, nrm = batch.tick()
        assert its[2] > 4 * max(its[0], its[1]), its     # the small channel runs on alone
        for k, c in enumerate(singles):
            sit, snrm = c.tick()
            assert (its[k], nrm[k]) == (sit[0], snrm[0]), (step, k)
    for k, c in enumerate(singles):
        a, b = batch.download(k), c.download(0)
        for n in a:
            assert_bits_equal(a[n], b[n], "channel %d %s" % (k, n))
    batch.close()
    for c in singles:
        c.close()


def test_whole_tick_autotune_picks_a_permitted_variant_and_keeps_the_bound():
    say("test: whole-tick autotune on an uneven batch")
    from stroemung_amd import workloads
    ws = workloads.random_channels(count=6, seed=3, imax=(24, 64), jmax=(16, 32), eps=1e-5)
    ctx = Context(0)
    for ct, vel, prm in ws:
        ctx.add_grid(prm, ct, vel)
    ctx.construct()
    ctx.tick()
    t0 = [ctx.state(c)["time"] for c in range(len(ws))]
    v, rate = ctx.autotune_ticks()
    assert 0 <= v < N_VARIANTS and rate > 0
    assert ctx.solver_variant()[0] == v
    for c in range(len(ws)):   # one real tick per variant
        assert ctx.state(c)["time"] > t0[c]
    for _ in range(2):
        its, _ = ctx.tick()
        for c in range(len(ws)):
            s = ctx.state(c)
            assert s["residual_squared"] <= s["tolerance_squared"] and not s["exhausted"]
            assert s["sor_iterations"] == its[c]
    ctx.close()


# ---------------------------------------------- trajectories at equal residual

def _rel(a, b, mask):
    return float(np.linalg.norm((a - b)[mask]) / np.linalg.norm(b[mask]))


TRAJ = [
    ("square obstacle channel 100x20", lambda: presets.square_obstacle(), 20,
     dict(delx=0.1, dely=0.1, delt=0.005, gamma=0.9, reynolds=100.0, omega=1.7)),
    ("lid cavity 64x64, adaptive step", lambda: presets.lid_cavity(64, 64), 10,
     dict(delx=1 / 64, dely=1 / 64, delt=0.0, tau=0.5, gamma=0.9, reynolds=1000.0, omega=1.8)),
]


@pytest.mark.parametrize("name,grid,steps,kw", TRAJ, ids=[t[0] for t in TRAJ])
def test_rb_trajectory_within_1e6_of_converged_oracle(name, grid, steps, kw):
    """u, v and p after N ticks of the GPU red-black path at eps = 1e-6 match
    the FP64 oracle solved to 1e-12 within 1e-6 relative L2 (p up to its
    fluid mean; the pressure problem is pure Neumann)."""
    say("test: trajectory", name)
    ct, vel = grid()
    fl = ct == O.FLUID
    params = make_params(kw["delx"], kw["dely"], kw["delt"], kw["gamma"], kw["reynolds"], 1e-6,
                         200000, kw["omega"], tau=kw.get("tau", 0.0), solver="rb",
                         check_every=8)
    ctx = Context(0)
    ch = ctx.add_grid(params, ct, vel)
    ctx.construct()
    gpu_its = [ctx.tick()[0][ch] for _ in range(steps)]
    got = ctx.download(ch)
    truth = O.Sim(ct, vel, sor_absolute_epsilon=1e-12, max_iterations=400000,
                  solver=O.SOLVER_RB, check_every=8, **kw).construct()
    for _ in range(steps):
        it, _ = truth.tick()
        assert it < 400000
    assert max(gpu_its) < 200000
    pm = lambda p: np.where(fl, p - p[fl].mean(), 0.0)  # noqa: E731
    errs = (_rel(got["u"], truth.u, fl), _rel(got["v"], truth.v, fl),
            _rel(pm(got["p"]), pm(truth.p), fl))
    say("  rel L2 vs 1e-12 oracle: u %.2e v %.2e p %.2e; sweeps/tick %s" % (errs + (gpu_its,)))
    assert max(errs) < 1e-6


def test_check_every_takes_the_sweep_counts_of_the_variants():
    """params.check_every = K pins the blocked solver to a K-sweep variant;
    a K that no variant has is refused when the grid is added."""
    from stroemung_amd import _lib
    say("test: check_every values")
    ct, vel = presets.lid_cavity(40, 40)
    for k in (1, 2, 4, 6, 8):
        prm = make_params(1 / 40, 1 / 40, 1e-3, 0.9, 100.0, 1e-6, 100, 1.7, solver="rb",
                          check_every=k)
        ctx = Context(0)
        ch = ctx.add_grid(prm, ct, vel)
        ctx.construct()
        assert ctx.solver_variant()[1] == k
        ctx.tick()
        ctx.close()
    for k in (3, 5, 16):
        prm = make_params(1 / 40, 1 / 40, 1e-3, 0.9, 100.0, 1e-6, 100, 1.7, solver="rb",
                          check_every=k)
        ctx = Context(0)
        with pytest.raises(_lib.StmError, match="check_every"):
            ctx.add_grid(prm, ct, vel)
        ctx.close()


@pytest.mark.parametrize("solver", ["rb", "lex"])
def test_nan_residual_ends_the_solve_and_is_reported(solver):
    """A NaN in p makes the residual NaN: the tick completes, reports
    STM_ERR_NAN, and the red-black solve stops at its first residual test
    instead of sweeping to itermax."""
    from stroemung_amd import _lib
    say("test: NaN residual", solver)
    n = 40
    ct, vel = presets.lid_cavity(n, n)
    itermax = 200000 if solver == "rb" else 60
    prm = make_params(1 / n, 1 / n, 1e-3, 0.9, 100.0, 1e-6, itermax, 1.7, solver=solver,
                      check_every=0)
    ctx = Context(0)
    ch = ctx.add_grid(prm, ct, vel)
    ctx.construct()
    ctx.tick()
    p = ctx.download(ch, ("p",))["p"]
    p[n // 2, n // 2] = np.nan
    ctx.upload(ch, p=p)
    with pytest.raises(_lib.NanError, match="not a number"):
        ctx.tick()
    st = ctx.state(ch)
    assert np.isnan(st["norm_squared"])
    if solver == "rb":
        assert st["sor_iterations"] <= 16, st["sor_iterations"]
    else:   # the reference's loop: no test passes, so it runs to itermax
        assert st["sor_iterations"] >= itermax - 1
    ctx.close()
