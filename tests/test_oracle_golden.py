"""Pin the CPU oracle against every golden vector the reference holds for the
time step: its unit-test vectors, its boundary-classification tests and its
insta snapshots (tests/golden/, written by make_golden.py)."""
import numpy as np
import pytest

from oracle import pyoracle as O
from stroemung_amd import io
from tests.refdata import arr, assert_bits_equal, from_json, golden, preset, sim_kwargs

MV = golden("math_vectors.json")


def _v(a):
    return O._dp(O.view3(a))


def test_du2dx():
    for u, dx, gam, exp in MV["du2dx"]:
        assert O.lib().orc_du2dx(_v(u), dx, gam) == exp


def test_duvdx():
    for u, v, dx, gam, exp in MV["duvdx"]:
        a, b = O.view3(u), O.view3(v)
        assert O.lib().orc_duvdx(O._dp(a), O._dp(b), dx, gam) == exp


def test_duvdy():
    for u, v, dy, gam, exp in MV["duvdy"]:
        a, b = O.view3(u), O.view3(v)
        assert O.lib().orc_duvdy(O._dp(a), O._dp(b), dy, gam) == exp


def test_dv2dy():
    for v, dy, gam, exp in MV["dv2dy"]:
        assert O.lib().orc_dv2dy(_v(v), dy, gam) == exp


def test_laplacian():
    for e, dx, dy, exp in MV["laplacian"]:
        assert O.lib().orc_laplacian(_v(e), dx, dy) == exp


@pytest.mark.parametrize("name", ["calculate_f", "calculate_g"])
def test_calculate_f_g(name):
    fn = getattr(O.lib(), "orc_" + name)
    for u, v, dx, dy, dt, gam, re, exp in MV[name]:
        a, b = O.view3(u), O.view3(v)
        assert fn(O._dp(a), O._dp(b), dx, dy, dt, gam, re) == exp


def test_residual_closed_form():
    # r = laplacian(p) - rhs with the reference's grouping; a quadratic field
    # p = x^2 + 3y^2 on unit cells has laplacian 2 + 6 = 8 exactly.
    p = np.array([[(x - 1) ** 2 + 3 * (y - 1) ** 2 for y in range(3)] for x in range(3)], float)
    assert O.lib().orc_residual(O._dp(O.view3(p)), 1.0, 1.0, 8.0) == 0.0
    assert O.lib().orc_residual(O._dp(O.view3(p)), 1.0, 1.0, 0.5) == 7.5


# ------------------------------------------------------ boundary classification

G = golden("grid_io.json")
KIND = {None: 0, "North": 1, "NorthEast": 2, "East": 3, "SouthEast": 4, "South": 5,
        "SouthWest": 6, "West": 7, "NorthWest": 8}
OFF = {"north_neighbor": (0, -1), "south_neighbor": (0, 1), "east_neighbor": (1, 0),
       "west_neighbor": (-1, 0)}


def test_rebuild_boundary_list():
    ex = G["rebuild_boundary_list"]
    nx, ny = ex["size"]
    for case in ex["examples"]:
        t = np.zeros((nx, ny), np.uint8)
        for x, y in case["cells"]:
            t[x, y] = O.NOSLIP
        edge, blist, fluid = O.classify(t)
        assert [tuple(divmod(int(i), ny)) for i in blist] == [tuple(c) for c in case["cells"]]
        for (x, y), e in zip(case["cells"], case["edges"]):
            if e is None:
                assert edge[x, y] == 0
                continue
            kind, nbrs = e
            assert edge[x, y] == KIND[kind]
            # the neighbour indices the reference stores follow from the kind
            for name, (nxp, nyp) in nbrs.items():
                dx, dy = OFF[name]
                assert (x + dx, y + dy) == (nxp, nyp)
        assert fluid == nx * ny - len(case["cells"])


def test_thin_boundary():
    ex = G["thin_boundary"]
    for cells in ex["examples"]:
        t = np.zeros(ex["size"], np.uint8)
        for x, y in cells:
            t[x, y] = O.NOSLIP
        with pytest.raises(ValueError, match="TooThin"):
            O.classify(t)


def test_deserialize_boundaries():
    d = G["small_grid_with_boundaries"]
    ct, _ = io.decode_cells(io.load_reference_json(d["input"])["cell_type"])
    edge, blist, fluid = O.classify(ct)
    ny = ct.shape[1]
    got = [[int(i) // ny, int(i) % ny, io.EDGE_NAMES[edge[divmod(int(i), ny)]]] for i in blist]
    assert got == d["boundary_list"]
    assert fluid == 2.0


# ------------------------------------------------------------ construction

S = golden("simulation_io.json")


def _oracle_from_json(sim_json):
    ct, vel, u, v, p, kw = from_json(sim_json)
    return O.Sim(ct, vel, u=u, v=v, p=p, **kw).construct()


def test_initial_norm_from_nast2d_data():
    # deserialize-2.snap:17 -- F/G without boundary conditions, RHS, residual.
    case = S["small_simulation_with_boundaries"]
    sim = _oracle_from_json(case["input"])
    assert sim.s.initial_norm_squared == 899.9547140394143
    assert case["expected"]["initial_norm_squared"] == 899.9547140394143


def test_initial_norm_zero_fields():
    for key in ("simple_simulation",):
        sim = _oracle_from_json(S[key]["input"])
        assert sim.s.initial_norm_squared == S[key]["expected"]["initial_norm_squared"] == 0.0
    ser = S["serialize"]
    ct, vel = preset(ser["setup"]["preset"], ser["setup"]["size"])
    sim = O.Sim(ct, vel, **sim_kwargs(ser["setup"])).construct()
    assert sim.s.initial_norm_squared == ser["expected"]["initial_norm_squared"]


# ------------------------------------------------------------ tick trajectory

T = golden("tick_4x3.json")


def _check_state(sim, snapshot, what):
    g = snapshot["grid"]
    assert_bits_equal(sim.p, arr(g["pressure"]), what + " p")
    assert_bits_equal(sim.u, arr(g["u"]), what + " u")
    assert_bits_equal(sim.v, arr(g["v"]), what + " v")
    assert sim.s.iterations == snapshot["iterations"]
    assert sim.s.time == snapshot["time"]
    assert sim.s.initial_norm_squared == snapshot["initial_norm_squared"]


def test_tick_trajectory_4x3():
    setup = T["setup"]
    ct, vel = preset(setup["preset"], setup["size"])
    sim = O.Sim(ct, vel, **sim_kwargs(setup)).construct()
    it, n2 = sim.tick()
    a1 = T["after_1"]
    assert (it, n2) == (a1["sor_iterations"], a1["norm_squared"])
    assert_bits_equal(sim.f, arr(a1["f"]), "tick1 f")
    assert_bits_equal(sim.g, arr(a1["g"]), "tick1 g")
    assert_bits_equal(sim.rhs, arr(a1["rhs"]), "tick1 rhs")
    _check_state(sim, a1["sim"], "tick1")
    for _ in range(100):
        it, n2 = sim.tick()
    a101 = T["after_101"]
    assert (it, n2) == (a101["sor_iterations"], a101["norm_squared"])
    assert_bits_equal(sim.f, arr(a101["f"]), "tick101 f")
    assert_bits_equal(sim.g, arr(a101["g"]), "tick101 g")
    assert_bits_equal(sim.rhs, arr(a101["rhs"]), "tick101 rhs")
    _check_state(sim, a101["sim"], "tick101")
    for _ in range(100):
        sim.tick()
    _check_state(sim, T["after_201"]["sim"], "tick201")
