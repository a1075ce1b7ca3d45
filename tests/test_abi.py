"""The C-ABI library: it loads, exports every function include/stroemung_hip.h
declares, fails cleanly without a device, and its host-side boundary
classification (no device needed) agrees with the oracle and the reference's
classification tests."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from oracle import pyoracle as O
from stroemung_amd import _lib, presets
from stroemung_amd.simulation import rebuild_boundary_list
from tests.refdata import golden

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "stroemung_hip.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(stm_[a-z_0-9]+)\s*\(", text)))


def test_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared()
    assert len(names) >= 25
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == set(_lib.SIGNATURES), "ctypes table out of sync with the header"


def test_no_device_is_a_clean_error():
    n = C.c_int(-1)
    rc = _lib.lib().stm_device_count(C.byref(n))
    if rc == 0 and n.value > 0:
        pytest.skip("a GPU is present")
    assert rc == -1 and n.value == 0
    ctx = C.c_void_p()
    rc = _lib.lib().stm_create(0, C.byref(ctx))
    assert rc == -1
    assert not ctx.value
    assert b"device" in _lib.lib().stm_last_error(None).lower()
    with pytest.raises(_lib.NoDeviceError):
        _lib.check(rc)


def test_invalid_arguments():
    L = _lib.lib()
    assert L.stm_create(0, None) == -6
    assert L.stm_construct(None) == -6
    assert L.stm_rebuild_boundary_list(0, 0, None, None, None, None, None, None, 0) == -6


def _random_grid(seed, nx, ny):
    """Odd seeds: scattered cells (mostly too thin); even seeds: rectangles of
    side >= 2 (valid unless they nearly touch)."""
    rng = np.random.default_rng(seed)
    t = np.zeros((nx, ny), np.uint8)
    if seed % 2:
        t[rng.random((nx, ny)) < 0.3] = 1
    else:
        for _ in range(rng.integers(1, 4)):
            w, h = rng.integers(2, 4, size=2)
            x0, y0 = rng.integers(0, nx - 1), rng.integers(0, ny - 1)
            t[x0:x0 + w, y0:y0 + h] = 1
    t[t == 1] = rng.integers(1, 5, size=int((t == 1).sum()))
    return t


def test_classification_matches_oracle_random():
    seen_ok = seen_thin = 0
    for seed in range(400):
        t = _random_grid(seed, 7 + seed % 5, 5 + seed % 7)
        try:
            e1, b1, f1 = O.classify(t)
        except ValueError:
            with pytest.raises(_lib.InvalidGridError, match="BoundaryTooThinError"):
                rebuild_boundary_list(t)
            seen_thin += 1
            continue
        e2, b2, f2 = rebuild_boundary_list(t)
        assert np.array_equal(e1, e2)
        assert [divmod(int(i), t.shape[1]) for i in b1] == b2
        assert f1 == f2
        seen_ok += 1
    assert seen_ok > 20 and seen_thin > 20


def test_classification_reference_cases():
    g = golden("grid_io.json")
    for case in g["rebuild_boundary_list"]["examples"]:
        t = np.zeros((3, 3), np.uint8)
        for x, y in case["cells"]:
            t[x, y] = 1
        edge, blist, fluid = rebuild_boundary_list(t)
        assert blist == [tuple(c) for c in case["cells"]]
    for cells in g["thin_boundary"]["examples"]:
        t = np.zeros((3, 3), np.uint8)
        for x, y in cells:
            t[x, y] = 1
        with pytest.raises(_lib.InvalidGridError):
            rebuild_boundary_list(t)


def test_presets_classify():
    for t, _ in (presets.simple_inflow((4, 3)), presets.obstacle((100, 20)),
                 presets.square_obstacle(), presets.lid_cavity(32, 16)):
        rebuild_boundary_list(t)
    t, _ = presets.square_obstacle()
    _, _, fluid = rebuild_boundary_list(t)
    assert fluid == 100 * 20 - 16


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus N > 1 exits with a message, before
    touching a device."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_bench_cpu_baseline_legs_run_on_a_small_grid():
    """The CPU legs of bench.py (oracle timing, no GPU) on a 48x48 cavity:
    the fields the JSON line carries, positive rates, and the threaded
    red-black leg labelled as not the reference algorithm."""
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sys.path.insert(0, root)
    import bench
    base = bench.cpu_baseline(48, 5000, 5)
    assert base["value"] > 0 and base["unit"] == "cell-updates/s" and base["cores"] == 1
    assert base["kind"] == "port" and "48x48" in base["sample"]
    par = bench.cpu_parallel(48, 5000, 5)
    assert par["value"] > 0 and par["cores"] >= 1 and "not the reference" in par["kind"]
