"""Initial grids: the reference presets and the benchmark workloads.

The reference presets follow src/grid/presets.rs:8-87.  The workloads add a
square-obstacle channel, a lid-driven cavity and random-obstacle channels
drawn with a SplitMix64 generator, so the same seed gives the same layouts in
any language.
"""
import math

import numpy as np

from .io import FLUID, INFLOW, MOVING_WALL, NOSLIP, OUTFLOW


def empty(size):
    """src/grid/presets.rs:8-17: all fluid."""
    nx, ny = size
    return np.zeros((nx, ny), np.uint8), np.zeros((nx, ny, 2))


def _channel(size, inflow=(1.0, 0.0)):
    nx, ny = size
    t = np.zeros((nx, ny), np.uint8)
    vel = np.zeros((nx, ny, 2))
    t[:, 0] = NOSLIP
    t[:, ny - 1] = NOSLIP
    t[0, 1:ny - 1] = INFLOW
    vel[0, 1:ny - 1] = inflow
    t[nx - 1, 1:ny - 1] = OUTFLOW
    return t, vel


def simple_inflow(size):
    """src/grid/presets.rs:19-40: no-slip top/bottom, inflow (1,0) west, outflow east."""
    return _channel(size)


def draw_circle(t, x, y, radius):
    """src/grid/presets.rs:42-62 (half-open spans, saturating at 0)."""
    nx, ny = t.shape
    r = int(radius)
    for xi in range(max(x - r, 0), x + r):
        if xi >= nx:
            continue
        for yi in range(max(y - r, 0), y + r):
            if yi >= ny:
                continue
            d = math.sqrt(float((xi - x) ** 2 + (yi - y) ** 2))
            if d < radius:
                t[xi, yi] = NOSLIP


def obstacle(size):
    """src/grid/presets.rs:64-87: channel plus a circle of radius 5 at (20, NY/2)."""
    t, vel = _channel(size)
    draw_circle(t, 20, size[1] // 2, 5.0)
    return t, vel


def square_obstacle(imax=100, jmax=20, side=4, centre=20):
    """Channel of imax x jmax interior cells with a side x side NoSlip square
    whose half-open span [centre - side/2, centre + side/2) is centred on
    cell `centre` in x and on the channel axis in y."""
    nx, ny = imax + 2, jmax + 2
    t, vel = _channel((nx, ny))
    h = side // 2
    cy = ny // 2
    t[centre - h:centre + h, cy - h:cy + h] = NOSLIP
    return t, vel


def lid_cavity(imax, jmax, lid=1.0):
    """Closed box; the y = 0 row (drawn on top) is a wall moving with u = lid."""
    nx, ny = imax + 2, jmax + 2
    t = np.full((nx, ny), NOSLIP, np.uint8)
    vel = np.zeros((nx, ny, 2))
    t[1:nx - 1, 1:ny - 1] = FLUID
    t[1:nx - 1, 0] = MOVING_WALL
    vel[1:nx - 1, 0] = (lid, 0.0)
    return t, vel


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014): the seeded generator of the workloads."""

    MASK = (1 << 64) - 1

    def __init__(self, seed):
        self.state = seed & self.MASK

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & self.MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self):
        """Double in [0, 1) from the top 53 bits."""
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def randint(self, lo, hi):
        """Integer in [lo, hi]."""
        return lo + self.next_u64() % (hi - lo + 1)


def _thin(t):
    """True if some boundary cell has fluid on opposite sides (the reference's
    BoundaryTooThinError, src/grid/mod.rs:297-331)."""
    b = t != FLUID
    fl = ~b
    nx, ny = t.shape
    pad = np.zeros((nx + 2, ny + 2), bool)
    pad[1:-1, 1:-1] = fl
    w, e = pad[:-2, 1:-1], pad[2:, 1:-1]
    n, s = pad[1:-1, :-2], pad[1:-1, 2:]
    return bool(np.any(b & ((w & e) | (n & s))))


def random_obstacles(t, rng, count, margin_lo, margin_hi, min_side=2, max_frac=0.25):
    """Place `count` axis-aligned NoSlip rectangles with sides >= min_side in
    the interior columns [margin_lo, nx - margin_hi).  Returns False if the
    layout makes a too-thin boundary (the caller redraws)."""
    nx, ny = t.shape
    for _ in range(count):
        w = rng.randint(min_side, max(min_side, int((nx - 2) * max_frac)))
        h = rng.randint(min_side, max(min_side, int((ny - 2) * max_frac)))
        x0 = rng.randint(margin_lo, max(margin_lo, nx - margin_hi - w))
        y0 = rng.randint(2, max(2, ny - 2 - h - 1))
        t[x0:x0 + w, y0:y0 + h] = NOSLIP
    return not _thin(t)


def random_channel(rng, imax, jmax):
    """Channel (inflow west, outflow east) with 1-4 random rectangles; redrawn
    until no boundary cell is too thin."""
    while True:
        t, vel = _channel((imax + 2, jmax + 2))
        k = rng.randint(1, 4)
        if random_obstacles(t, rng, k, margin_lo=4, margin_hi=4):
            return t, vel


def random_cavity(rng, imax, jmax, lid=1.0):
    """Lid cavity with 1-4 random rectangles."""
    while True:
        t, vel = lid_cavity(imax, jmax, lid)
        k = rng.randint(1, 4)
        if random_obstacles(t, rng, k, margin_lo=3, margin_hi=3):
            return t, vel


def log_uniform_int(rng, lo, hi):
    return int(round(math.exp(math.log(lo) + rng.uniform() * (math.log(hi) - math.log(lo)))))
