"""The GUI image (src/visualization.rs:7-105, render_simulation).

CPU: the oracle's restatement against closed forms -- the ends of the range
are pure blue and pure red, boundary cells have the fixed colours, a
degenerate range gives the NaN hue's colour -- and against the textbook HSL
conversion (Python's colorsys) at every hue.  GPU: the device image is
bit-identical to the oracle's on the same fields and ranges.
"""
import colorsys

import numpy as np
import pytest

from oracle import pyoracle as O
from stroemung_amd import presets


def _sim(ct, p=None, u=None, v=None):
    return O.Sim(ct, delx=0.1, dely=0.1, delt=0.01, gamma=0.9, reynolds=100.0,
                 sor_absolute_epsilon=1e-3, max_iterations=10, omega=1.7, p=p, u=u, v=v)


def _u8(c):
    v = c * 255.0
    return 0 if not v > 0 else (255 if v >= 255 else int(v))


def test_range_ends_and_boundary_colours():
    ct = np.zeros((5, 3), np.uint8)
    ct[0, :] = 1                          # a wall column
    p = np.zeros((5, 3))
    p[1, 1], p[2, 1] = -2.0, 6.0          # the range ends
    s = _sim(ct, p=p)
    s.s.pressure_range[0], s.s.pressure_range[1] = -2.0, 6.0
    img = s.render(0)
    assert list(img[1, 1]) == [0, 0, 255, 255]      # low end: hue 240, blue
    assert list(img[1, 2]) == [255, 0, 0, 255]      # high end: hue 0, red
    assert list(img[0, 0]) == [127, 0, 0, 255]      # boundary, pressure view
    s.s.speed_range[0], s.s.speed_range[1] = 0.0, 1.0
    assert list(s.render(1)[0, 0]) == [127, 127, 127, 255]   # boundary, speed view


def test_degenerate_range_gives_the_nan_hue_colour():
    ct = np.zeros((3, 3), np.uint8)
    s = _sim(ct)
    s.s.speed_range[0] = s.s.speed_range[1] = 0.0   # zero flow: 0 * 240 / 0
    img = s.render(1)
    assert np.all(img.reshape(-1, 4) == [255, 0, 0, 255])


def test_hues_follow_the_textbook_hsl_conversion():
    n = 721
    ct = np.zeros((n, 1), np.uint8)
    p = np.linspace(0.0, 240.0, n).reshape(n, 1)
    s = _sim(ct, p=p)
    s.s.pressure_range[0], s.s.pressure_range[1] = 0.0, 240.0
    img = s.render(0)
    for x in range(n):
        hue = 240.0 - p[x, 0]
        r, g, b = colorsys.hls_to_rgb(hue / 360.0, 0.5, 1.0)
        want = [_u8(r), _u8(g), _u8(b)]
        got = [int(c) for c in img[0, x, :3]]
        assert all(abs(a - b_) <= 1 for a, b_ in zip(got, want)), (hue, got, want)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_gpu_image_is_bit_identical_to_the_oracle(mode):
    from stroemung_amd.simulation import Context, make_params
    from tests.gridgen import mixed_grid
    ct, vel = mixed_grid(3, 40, 27)
    prm = make_params(0.05, 0.07, 0.002, 0.9, 200.0, 1e-3, 50, 1.7, tau=0.5, solver="rb")
    ctx = Context(0)
    ctx.add_grid(prm, ct, vel)
    ctx.construct()
    # zero fields first: degenerate speed range
    img0 = ctx.render(0, "speed" if mode else "pressure")
    for _ in range(5):
        ctx.tick()
    st = ctx.state(0)
    fld = ctx.download(0, ("u", "v", "p"))
    img = ctx.render(0, "speed" if mode else "pressure")
    s = _sim(ct, p=fld["p"], u=fld["u"], v=fld["v"])
    s.s.pressure_range[0], s.s.pressure_range[1] = st["pressure_range"]
    s.s.speed_range[0], s.s.speed_range[1] = st["speed_range"]
    np.testing.assert_array_equal(img, s.render(mode))
    assert img0.shape == img.shape == (27, 40, 4)
    ctx.close()
