"""State I/O, checkpoint / resume and grid edits on the GPU path.

- to_json() reproduces the reference's serialize and deserialize snapshots
  (src/simulation.rs:408-447);
- a run stopped after N ticks and resumed (in memory, or through the JSON
  form) gives the same bits as N + M uninterrupted ticks;
- a grid edit between ticks follows draw_cells (src/lib.rs:38-78): the
  changed cells restart from rest, the boundary list is rebuilt, and the next
  ticks match the oracle rebuilt on the same cells; a too-thin edit is
  refused with the reference's error and changes nothing;
- new scalars through stm_set_params take effect, the check interval
  included.
"""
import json
import sys

import numpy as np
import pytest

from oracle import pyoracle as O
from stroemung_amd import io as sio
from stroemung_amd import presets
from stroemung_amd._lib import InvalidGridError
from stroemung_amd.simulation import Context, Simulation, make_params
from tests.refdata import assert_bits_equal, golden, preset

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


FIELDS = ("u", "v", "p", "f", "g", "rhs")


def _same_json(got, want, path="$"):
    """Equal JSON values, floats by bit pattern, object keys in the same order."""
    if isinstance(want, dict):
        assert isinstance(got, dict), path
        assert list(got.keys()) == list(want.keys()), (path, list(got.keys()), list(want.keys()))
        for k in want:
            _same_json(got[k], want[k], path + "." + k)
    elif isinstance(want, list):
        assert isinstance(got, list) and len(got) == len(want), path
        for i, (a, b) in enumerate(zip(got, want)):
            _same_json(a, b, "%s[%d]" % (path, i))
    elif isinstance(want, float) or isinstance(got, float):
        assert np.float64(got).view(np.uint64) == np.float64(want).view(np.uint64), \
            (path, got, want)
    else:
        assert got == want, (path, got, want)


# ------------------------------------------------------------ serde snapshots

def test_to_json_matches_the_serialize_snapshot():
    """src/simulation.rs:422-447: a 5x7 empty grid, constructed, serialised."""
    say("test: serialize snapshot")
    d = golden("simulation_io.json")["serialize"]
    s = d["setup"]
    ct, vel = preset(s["preset"], s["size"])
    sim = Simulation(s["size"], s["cell_size"], s["delt"], s["gamma"], s["reynolds"],
                     s["sor_absolute_epsilon"], s["max_iterations"], s["omega"], ct, vel)
    text = sio.dumps(sim.to_json())
    _same_json(json.loads(text), d["expected"])
    sim.ctx.close()


@pytest.mark.parametrize("name", ["simple_simulation", "small_simulation_with_boundaries"])
def test_from_reader_to_json_matches_the_deserialize_snapshots(name):
    """src/simulation.rs:408-420: each test-data file read with from_reader
    (floats parsed as serde_json parses them), constructed, serialised."""
    say("test: deserialize snapshot", name)
    d = golden("simulation_io.json")[name]
    sim = Simulation.from_reader(d["input"])
    _same_json(json.loads(sio.dumps(sim.to_json())), d["expected"])
    sim.ctx.close()


# ------------------------------------------------------------ checkpoint / resume

C1 = dict(delx=0.1, dely=0.1, delt=0.005, gamma=0.9, reynolds=100.0,
          sor_absolute_epsilon=1e-3, max_iterations=100, omega=1.7)


def _sim(solver, tau=0.0):
    ct, vel = presets.square_obstacle(60, 20)
    kw = dict(C1)
    if solver == "rb":
        kw.update(sor_absolute_epsilon=1e-7, max_iterations=20000)
    return Simulation((62, 22), (kw["delx"], kw["dely"]), kw["delt"], kw["gamma"],
                      kw["reynolds"], kw["sor_absolute_epsilon"], kw["max_iterations"],
                      kw["omega"], ct, vel, tau=tau, solver=solver,
                      check_every=4 if solver == "rb" else 1)


@pytest.mark.parametrize("solver,tau", [("lex", 0.0), ("rb", 0.0), ("rb", 0.5)])
def test_resume_in_memory_equals_the_uninterrupted_run(solver, tau):
    """N ticks; the fields and scalars read back; a new context given those
    scalars, the fields uploaded and constructed (the resume sequence a host
    runs: construction keeps the carried initial_norm_squared and forms the
    max |u|, |v| the next adaptive step reads); M more ticks: the same bits
    and iteration counts as N + M ticks in one run.  (The pressure range is
    the one construction forms from the loaded p, as in the reference.)"""
    say("test: in-memory resume", solver, tau)
    n, m = 7, 6
    a = _sim(solver, tau)
    for _ in range(n):
        a.run_simulation_tick()
    st = a.state
    fld = a.ctx.download(a.channel)
    prm = make_params(0.1, 0.1, st["delt"], C1["gamma"], C1["reynolds"],
                      a.sor_absolute_epsilon, a.max_iterations, C1["omega"],
                      st["initial_norm_squared"] if st["has_initial_norm_squared"] else None,
                      st["iterations"], st["time"], tau, solver, a.check_every)
    b = Context(0)
    ch = b.add_grid(prm, a.cell_type, a.cell_velocity)
    b.upload(ch, **fld)
    b.construct()
    assert b.state(ch)["initial_norm_squared"] == st["initial_norm_squared"]
    ra, rb = [], []
    for _ in range(m):
        ra.append(a.run_simulation_tick())
        its, nrm = b.tick()
        rb.append((its[ch], nrm[ch]))
    assert ra == rb
    got, want = b.download(ch), a.ctx.download(a.channel)
    for k in FIELDS:
        assert_bits_equal(got[k], want[k], "resumed %s" % k)
    sa, sb = a.state, b.state(ch)
    for k in ("time", "delt", "iterations", "initial_norm_squared", "speed_range"):
        assert sa[k] == sb[k], k
    b.close()
    a.ctx.close()


@pytest.mark.parametrize("solver,tau", [("lex", 0.0), ("rb", 0.5)])
def test_resume_through_json_equals_the_uninterrupted_run(solver, tau):
    """N ticks, to_json, from_reader with correctly rounded parsing, M more
    ticks: the same bits as the uninterrupted run (f, g and rhs are not
    serialised; construction recomputes them and a tick reads only values it
    recomputes).  With the reference's own float parsing a 17-digit value
    can come back one ulp off: the loaded fields are within one ulp."""
    say("test: JSON resume", solver, tau)
    n, m = 6, 5
    a = _sim(solver, tau)
    for _ in range(n):
        a.run_simulation_tick()
    text = sio.dumps(a.to_json())
    ext = dict(tau=tau, solver=solver, check_every=a.check_every)
    b = Simulation.from_reader(text, exact=True, **ext)
    for name in ("u", "v", "pressure"):
        assert_bits_equal(getattr(b, name), getattr(a, name), "loaded " + name)
    # the reference's parse: at most one ulp from the saved values
    c = Simulation.from_reader(text, **ext)
    for name in ("u", "v", "pressure"):
        x, y = getattr(c, name), getattr(a, name)
        assert (np.sign(x) == np.sign(y)).all(), name
        assert np.abs(x.view(np.int64) - y.view(np.int64)).max() <= 1, name
    c.ctx.close()
    for _ in range(m):
        assert a.run_simulation_tick() == b.run_simulation_tick()
    for name in ("u", "v", "pressure"):
        assert_bits_equal(getattr(b, name), getattr(a, name), "resumed " + name)
    assert (b.time, b.iterations, b.delt) == (a.time, a.iterations, a.delt)
    for s in (a, b):
        s.ctx.close()


# ------------------------------------------------------------ grid edits

def _draw(ct, x0, y0, kind):
    """draw_cells (src/lib.rs:38-78) on a copy of the types: the 2x2 square
    at (x0, y0), outer ring untouched; returns (new types, changed mask)."""
    t = ct.copy()
    for x, y in ((x0, y0), (x0 + 1, y0), (x0, y0 + 1), (x0 + 1, y0 + 1)):
        if 0 < x < ct.shape[0] - 1 and 0 < y < ct.shape[1] - 1:
            t[x, y] = kind
    return t, t != ct


def _oracle_edit(osim, new, changed):
    osim.u[changed] = 0.0
    osim.v[changed] = 0.0
    osim.p[changed] = 0.0
    osim.cell_type[...] = new
    assert osim.call("rebuild_boundary_list") == 0


def test_edit_between_ticks_matches_the_oracle():
    """Add a 2x2 obstacle to a running channel, later erase part of the
    original obstacle: every tick stays bit-exact with the oracle edited the
    same way (lexicographic solver, every field and the iteration counts)."""
    say("test: grid edits vs oracle")
    ct, vel = presets.square_obstacle(60, 20)
    ctx = Context(0)
    prm = make_params(C1["delx"], C1["dely"], C1["delt"], C1["gamma"], C1["reynolds"],
                      C1["sor_absolute_epsilon"], C1["max_iterations"], C1["omega"])
    ch = ctx.add_grid(prm, ct, vel)
    ctx.construct()
    osim = O.Sim(ct, vel, **C1).construct()
    cur = ct.copy()
    edits = [(3, (40, 8), O.NOSLIP), (6, (18, 9), O.FLUID)]
    for step in range(9):
        for at, (x0, y0), kind in edits:
            if step == at:
                new, changed = _draw(cur, x0, y0, kind)
                assert changed.any()
                ctx.set_cells(ch, new, vel)
                _oracle_edit(osim, new, changed)
                cur = new
                got = ctx.download(ch, ("u", "v", "p"))
                for k in ("u", "v", "p"):
                    assert_bits_equal(got[k], getattr(osim, k), "after edit %s" % k)
        its, nrm = ctx.tick()
        assert (its[ch], nrm[ch]) == osim.tick(), step
        got = ctx.download(ch)
        for k in FIELDS:
            assert_bits_equal(got[k], getattr(osim, k), "tick %d %s" % (step, k))
    ctx.close()


def test_too_thin_edit_is_refused_and_changes_nothing():
    """A 1-cell-wide wall between fluid cells is a BoundaryTooThinError
    (src/grid/mod.rs:57-58, raised by rebuild_boundary_list); draw_cells then
    restores the old cells and fields, so the next tick equals the tick of
    an unedited run."""
    say("test: too-thin edit")
    ct, vel = presets.square_obstacle(60, 20)
    prm = make_params(C1["delx"], C1["dely"], C1["delt"], C1["gamma"], C1["reynolds"],
                      C1["sor_absolute_epsilon"], C1["max_iterations"], C1["omega"])
    runs = []
    for edit in (False, True):
        ctx = Context(0)
        ch = ctx.add_grid(prm, ct, vel)
        ctx.construct()
        for _ in range(3):
            ctx.tick()
        if edit:
            bad = ct.copy()
            bad[40, 5:15] = O.NOSLIP          # one cell thick: fluid west and east
            before = ctx.download(ch)
            with pytest.raises(InvalidGridError) as ei:
                ctx.set_cells(ch, bad, vel)
            assert "BoundaryTooThinError: A cell `Boundary(NoSlip)` at `(40, 5)` has fluid " \
                   "on opposing sides." in str(ei.value)
            after = ctx.download(ch)
            for k in FIELDS:
                assert_bits_equal(after[k], before[k], "refused edit %s" % k)
        its, nrm = ctx.tick()
        runs.append(((its[ch], nrm[ch]), ctx.download(ch)))
        ctx.close()
    assert runs[0][0] == runs[1][0]
    for k in FIELDS:
        assert_bits_equal(runs[1][1][k], runs[0][1][k], "next tick %s" % k)


# ------------------------------------------------------------ new scalars

def test_set_params_changes_the_check_interval_and_the_scalars():
    """stm_set_params after construction: a new check_every moves the blocked
    solver to a variant with that many sweeps per launch, and the ticks then
    equal those of a context created with the new scalars."""
    say("test: set_params")
    ct, vel = presets.lid_cavity(48, 40)

    def prm(k, omega):
        return make_params(1 / 48, 1 / 40, 1e-3, 0.9, 400.0, 1e-7, 50000, omega, solver="rb",
                           check_every=k)
    a = Context(0)
    ca = a.add_grid(prm(4, 1.7), ct, vel)
    a.construct()
    for _ in range(2):
        a.tick()
    assert a.solver_variant()[1] == 4
    st = a.state(ca)
    fld = a.download(ca)
    new = prm(8, 1.8)
    new.time, new.iterations = st["time"], st["iterations"]
    a.set_params(ca, new)
    assert a.solver_variant()[1] == 8
    b = Context(0)
    cb = b.add_grid(new, ct, vel)
    b.upload(cb, **fld)
    for _ in range(2):
        ia, na = a.tick()
        ib, nb = b.tick()
        assert (ia[ca], na[ca]) == (ib[cb], nb[cb])
        assert ia[ca] % 8 == 0 or a.state(ca)["exhausted"]
    ga, gb = a.download(ca), b.download(cb)
    for k in FIELDS:
        assert_bits_equal(ga[k], gb[k], "after new params %s" % k)
    assert a.state(ca)["time"] == b.state(cb)["time"]
    a.close()
    b.close()
