"""Seeded grids with every boundary type and edge kind, for parity tests."""
import numpy as np

from stroemung_amd import presets
from stroemung_amd.io import FLUID, INFLOW, MOVING_WALL, NOSLIP, OUTFLOW


def mixed_grid(seed, nx, ny, n_rect=6):
    """Ring and rectangles of random boundary types (NoSlip, Inflow, Outflow,
    moving wall) with random velocities; rectangles of side >= 2 so the layout
    has every edge kind.  Redrawn until no boundary cell is too thin."""
    rng = np.random.default_rng(seed)
    while True:
        t = np.zeros((nx, ny), np.uint8)
        t[0, :] = t[-1, :] = t[:, 0] = t[:, -1] = NOSLIP
        for _ in range(n_rect):
            w = int(rng.integers(2, max(3, nx // 5)))
            h = int(rng.integers(2, max(3, ny // 5)))
            x0 = int(rng.integers(2, nx - w - 2))
            y0 = int(rng.integers(2, ny - h - 2))
            t[x0:x0 + w, y0:y0 + h] = NOSLIP
        if not presets._thin(t):
            break
    b = t != FLUID
    kinds = rng.choice([NOSLIP, INFLOW, OUTFLOW, MOVING_WALL], size=int(b.sum()),
                       p=[0.4, 0.2, 0.2, 0.2])
    t[b] = kinds
    vel = np.zeros((nx, ny, 2))
    vel[b] = rng.uniform(-1.5, 1.5, size=(int(b.sum()), 2))
    return t, vel


def random_fields(seed, shape, scale=1.0):
    rng = np.random.default_rng(seed)
    return [rng.uniform(-scale, scale, size=shape) for _ in range(6)]
