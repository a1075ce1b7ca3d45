"""Helpers that turn the golden fixtures into simulation inputs."""
import json
import os

import numpy as np

from stroemung_amd import io, presets

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(name):
    with open(os.path.join(GOLDEN, name)) as fh:
        return json.load(fh)


def preset(name, size):
    return getattr(presets, name)(tuple(size))


def sim_kwargs(setup):
    dx, dy = setup["cell_size"]
    return dict(delx=dx, dely=dy, delt=setup["delt"], gamma=setup["gamma"],
                reynolds=setup["reynolds"], sor_absolute_epsilon=setup["sor_absolute_epsilon"],
                max_iterations=setup["max_iterations"], omega=setup["omega"])


def from_json(sim_json):
    """Reference `UnfinalizedSimulation` JSON (text, parsed as serde_json does,
    or an already parsed dict) -> (cell_type, cell_vel, u, v, p, kwargs)."""
    if isinstance(sim_json, str):
        sim_json = io.load_reference_json(sim_json)
    g = sim_json["grid"]
    ct, vel = io.decode_cells(g["cell_type"])
    u, v, p = (io.decode_array(g[k]) for k in ("u", "v", "pressure"))
    kw = sim_kwargs(sim_json)
    kw["initial_norm_squared"] = sim_json.get("initial_norm_squared")
    kw["iterations"] = sim_json["iterations"]
    kw["time"] = sim_json["time"]
    return ct, vel, u, v, p, kw


def arr(snap):
    return io.decode_array(snap)


def bits(a):
    """Exact bit patterns (so -0.0 != 0.0 and NaNs compare by payload)."""
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def assert_bits_equal(a, b, what=""):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    ba, bb = bits(a), bits(b)
    if not np.array_equal(ba, bb):
        bad = np.argwhere(ba != bb)
        i = tuple(bad[0])
        raise AssertionError("%s: %d cells differ, first at %s: %r vs %r"
                             % (what, len(bad), i, a[i], b[i]))
