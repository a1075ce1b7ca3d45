"""Write the golden fixtures under tests/golden/ from the reference's own data.

Run in a container that has the reference checkout (read-only) at
/root/reference:   python tests/golden/make_golden.py

Only data is taken: the insta snapshot bodies (JSON) and the JSON test inputs,
plus the numeric test vectors and asserted values of the reference's unit
tests, transcribed by hand below with their file:line.  All floats are
shortest round-trip decimals, so json.loads recovers the exact f64 bits
(including -0.0) and json.dump writes them back unchanged.
"""
import json
import os
import sys

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def snap(path):
    """Body of an insta snapshot: the text after the second '---' line."""
    with open(os.path.join(REF, path)) as fh:
        text = fh.read()
    body = text.split("---\n", 2)[2]
    return body


def snap_json(path):
    return json.loads(snap(path))


def load(path):
    """A JSON input file of the reference's tests, kept as text: the reference
    parses it with serde_json, whose float parsing the host reader emulates."""
    with open(os.path.join(REF, path)) as fh:
        return fh.read()


S = "src/snapshots/stroemung__simulation__tests__"

# Asserted values of the tick test, src/simulation.rs:597-598 and 605-606.
TICK_ASSERTS = {
    "after_1": {"sor_iterations": 100, "norm_squared": 562901.7447199143},
    "after_101": {"sor_iterations": 1, "norm_squared": 3.8344148218167323e-20},
}

# Setup of the tick test, src/simulation.rs:573-588 (grid: presets::simple_inflow([4,3])).
TICK_SETUP = {
    "size": [4, 3], "cell_size": [0.1, 0.2], "delt": 0.005, "gamma": 0.9,
    "reynolds": 100.0, "sor_absolute_epsilon": 0.001, "max_iterations": 100,
    "omega": 1.7, "preset": "simple_inflow",
}


def tick_fixture():
    return {
        "setup": TICK_SETUP,
        # src/simulation.rs:590-594: f, g, rhs, then the whole simulation.
        "after_1": {
            "f": snap_json(S + "simulation_tick.snap"),
            "g": snap_json(S + "simulation_tick-2.snap"),
            "rhs": snap_json(S + "simulation_tick-3.snap"),
            "sim": snap_json(S + "simulation_tick-4.snap"),
            **TICK_ASSERTS["after_1"],
        },
        # src/simulation.rs:602-610
        "after_101": {
            "f": snap_json(S + "simulation_tick-5.snap"),
            "g": snap_json(S + "simulation_tick-6.snap"),
            "rhs": snap_json(S + "simulation_tick-7.snap"),
            "sim": snap_json(S + "simulation_tick-8.snap"),
            **TICK_ASSERTS["after_101"],
        },
        # src/simulation.rs:612-617
        "after_201": {"sim": snap_json(S + "simulation_tick-9.snap")},
    }


def deserialize_fixture():
    return {
        # src/simulation.rs:408-420: inputs and their snapshots, in order.
        "simple_simulation": {
            "input": load("src/test_data/simple_simulation.json"),
            "expected": snap_json(S + "deserialize.snap"),
        },
        "small_simulation_with_boundaries": {
            "input": load("src/test_data/small_simulation_with_boundaries.json"),
            "expected": snap_json(S + "deserialize-2.snap"),
        },
        # src/simulation.rs:422-447 (presets::empty([5,7]), delt 1.4, gamma 1.7)
        "serialize": {
            "setup": {"size": [5, 7], "cell_size": [1.0, 2.0], "delt": 1.4, "gamma": 1.7,
                      "reynolds": 100.0, "sor_absolute_epsilon": 0.001,
                      "max_iterations": 100, "omega": 1.7, "preset": "empty"},
            "expected": snap_json(S + "serialize.snap"),
        },
    }


def parse_boundary_display(text):
    """BoundaryList's Display (src/grid/mod.rs:71-83) -> sorted list of (idx, edge)."""
    out = []
    section = None
    for line in text.splitlines():
        line = line.strip()
        if line == "Sorted Boundary List:":
            section = "sorted"
            continue
        if line == "Boundaries:":
            section = "set"
            continue
        if section != "sorted" or not line:
            continue
        # ((x, y), None) | ((x, y), Some(Kind { ... }))
        inner = line[1:-1]
        xy, rest = inner.split("), ", 1)
        x, y = (int(t) for t in xy.lstrip("(").split(","))
        kind = None if rest == "None" else rest[len("Some("):].split(" ", 1)[0]
        out.append([x, y, kind])
    return out


def grid_fixture():
    G = "src/grid/snapshots/stroemung__grid__tests__"
    return {
        # src/grid/mod.rs:803-811
        "simple_grid": {"input": load("src/test_data/simple_grid.json"),
                        "expected": snap_json(G + "deserialize.snap")},
        # src/grid/mod.rs:813-823
        "small_grid_with_boundaries": {
            "input": load("src/test_data/small_grid_with_boundaries.json"),
            "expected": snap_json(G + "deserialize_boundaries.snap"),
            "boundary_list": parse_boundary_display(snap(G + "deserialize_boundaries-2.snap")),
        },
        # src/grid/mod.rs:825-830 (presets::empty([2,3]))
        "serialize": {"size": [2, 3], "expected": snap_json(G + "serialize.snap")},
        # tests/test_file_parsing.rs:15-21
        "small_data_out": {"input": load("tests/test_data/small_data.out.json"),
                           "expected": snap_json("tests/snapshots/test_file_parsing__deserialize.snap")},
        # src/grid/mod.rs:683-705: both layouts must fail with BoundaryTooThinError.
        "thin_boundary": {"size": [3, 3],
                          "examples": [[[1, 0], [1, 1], [1, 2]], [[0, 1], [1, 1], [2, 1]]]},
        # src/grid/mod.rs:707-801: NoSlip cells -> expected sorted list with edge kinds
        # and neighbour indices.
        "rebuild_boundary_list": {"size": [3, 3], "examples": [
            {"cells": [[0, 0], [0, 1], [0, 2], [1, 0], [1, 2], [2, 0], [2, 1], [2, 2]],
             "edges": [None, ["East", {"east_neighbor": [1, 1]}], None,
                       ["South", {"south_neighbor": [1, 1]}],
                       ["North", {"north_neighbor": [1, 1]}], None,
                       ["West", {"west_neighbor": [1, 1]}], None]},
            {"cells": [[0, 0], [0, 2], [2, 0], [2, 2]],
             "edges": [["SouthEast", {"south_neighbor": [0, 1], "east_neighbor": [1, 0]}],
                       ["NorthEast", {"north_neighbor": [0, 1], "east_neighbor": [1, 2]}],
                       ["SouthWest", {"south_neighbor": [2, 1], "west_neighbor": [1, 0]}],
                       ["NorthWest", {"north_neighbor": [2, 1], "west_neighbor": [1, 2]}]]},
        ]},
    }


# Numeric vectors of the reference's unit tests, transcribed by hand.
# A 3x3 view is given row by row: view[(a, b)] = e[x-1+a][y-1+b].
_A = [[1., 2., 3.], [4., 5., 6.], [7., 8., 9.]]
_B = [[8., 9., 10.], [11., 12., 13.], [14., 15., 16.]]
_A2 = [[1., 2., 3.], [4., 5., -6.], [-7., 8., 9.]]
_B2 = [[8., 9., 10.], [11., -12., 13.], [14., 15., -16.]]
_C = [[1., 4., 1.], [1., 5., 3.], [2., 1., 1.]]
_R = lambda a, b, c: [[0., a, 0.], [0., b, 0.], [0., c, 0.]]   # noqa: E731
_Y = lambda a, b, c: [[0., 0., 0.], [a, b, c], [0., 0., 0.]]   # noqa: E731

MATH_VECTORS = {
    # src/math.rs:195-232 (u, delx, gamma, expected)
    "du2dx": [[_R(1, 2, 3), 1., 1.7, 3.15], [_R(-1, 2, 3), 1., 1.7, 5.15],
              [_R(1, 1, 1), 1., 1.7, 0.], [_R(1, 2, 3), 1., 1.3, 3.35],
              [_R(1, 2, 3), 1.5, 1.7, 2.1], [_R(10, 20, 30), 1.5, 1.7, 210.]],
    # src/math.rs:241-270 (u, v, delx, gamma, expected)
    "duvdx": [[_A, _B, 1., 1.7, 40.35], [_A2, _B2, 1., 1.7, -53.1],
              [_A, _B, 1.6, 1.7, 25.218750], [_A, _B, 1., 1.5, 41.25]],
    # src/math.rs:282-311 (u, v, dely, gamma, expected)
    "duvdy": [[_A, _B, 1., 1.7, 17.15], [_A2, _B2, 1., 1.7, -32.35],
              [_A, _B, 1.6, 1.7, 10.718749999999998], [_A, _B, 1., 1.5, 17.25]],
    # src/math.rs:323-360 (v, dely, gamma, expected)
    "dv2dy": [[_Y(1, 2, 3), 1., 1.7, 3.15], [_Y(-1, 2, 3), 1., 1.7, 5.15],
              [_Y(1, 1, 1), 1., 1.7, 0.], [_Y(1, 2, 3), 1., 1.3, 3.35],
              [_Y(1, 2, 3), 1.5, 1.7, 2.1], [_Y(10, 20, 30), 1.5, 1.7, 210.]],
    # src/math.rs:369-398 (e, delx, dely, expected)
    "laplacian": [[_C, 1., 1., -11.],
                  [[[-1., -1., -1.], [-1., -5., -1.], [-1., -1., -1.]], 1., 1., 16.],
                  [_C, 1.6, 1.3, -5.503420857988164], [_C, 1.3, 0.7, -15.20347784084048]],
    # src/simulation.rs:452-493 (u, v, delx, dely, delt, gamma, reynolds, expected)
    "calculate_f": [[_A, _B, 1., 1., 0.005, 1.7, 100., 4.802500],
                    [_A2, _B2, 1., 1., 0.006, 1.7, 10., 5.052800],
                    [_A, _B, 1.6, 1., 0.007, 1.7, 14., 4.782168750],
                    [_A, _B, 1., 1.6, 0.003, 1.5, 400., 4.89790625]],
    # src/simulation.rs:513-554
    "calculate_g": [[_A, _B, 1., 1., 0.005, 1.7, 100., 11.6825],
                    [_A2, _B2, 1., 1., 0.006, 1.7, 10., -11.5014],
                    [_A, _B, 1.6, 1., 0.007, 1.7, 14., 11.66141875],
                    [_A, _B, 1., 1.6, 0.003, 1.5, 400., 11.83265625]],
}


def nast2d_fixture():
    """python/test_data: a NaSt2D .out file (raw bytes, as a list of ints)
    with the two outputs the reference's test_generate_test_data.py expects
    from it."""
    d = "python/test_data/small_data.out"
    with open(os.path.join(REF, d), "rb") as fh:
        raw = list(fh.read())
    with open(os.path.join(REF, d + "_expected.json")) as fh:
        parsed = json.load(fh)
    with open(os.path.join(REF, d + "_rust_expected.json")) as fh:
        grid = json.load(fh)
    return {"out_bytes": raw, "parsed": parsed, "rust_grid": grid}


def main():
    if not os.path.isdir(REF):
        sys.exit("reference checkout not found at " + REF)
    fixtures = {
        "math_vectors.json": MATH_VECTORS,
        "tick_4x3.json": tick_fixture(),
        "simulation_io.json": deserialize_fixture(),
        "grid_io.json": grid_fixture(),
        "nast2d_small.json": nast2d_fixture(),
    }
    for name, data in fixtures.items():
        with open(os.path.join(HERE, name), "w") as fh:
            json.dump(data, fh, separators=(",", ":"))
            fh.write("\n")
        print("wrote", name)


if __name__ == "__main__":
    main()
