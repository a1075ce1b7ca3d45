"""The benchmark workloads as concrete inputs.

square_channel  -- flow past a 4x4 square (imax 100, jmax 20, 102x22 cells with the ring)
cavity          -- lid-driven cavity, n x n cells on the unit square
random_channels -- many independent channels of uneven size with random obstacles
random_cavities -- lid cavities with random obstacles (training data for the network)
fixed_channels  -- many channels of one size with random obstacles (learned + batched)

All random draws come from SplitMix64 (presets.SplitMix64) with the stated seed, so
the same seed gives the same layouts and parameters.
"""
from . import presets
from .simulation import make_params


def square_channel_params(solver="lex", eps=1e-3, itermax=100, check_every=1):
    """100x20 interior, xlength 10, ylength 2, Re 100, gamma 0.9, omega 1.7,
    delt 0.005 fixed (the reference's default step)."""
    return make_params(0.1, 0.1, 0.005, 0.9, 100.0, eps, itermax, 1.7, solver=solver,
                       check_every=check_every)


def cavity(n=1024, eps=1e-6, omega=1.8, tau=0.5, reynolds=1000.0, check_every=0,
           itermax=10_000_000):
    ct, vel = presets.lid_cavity(n, n)
    prm = make_params(1.0 / n, 1.0 / n, 0.0, 0.9, reynolds, eps, itermax, omega, tau=tau,
                      solver="rb", check_every=check_every)
    return ct, vel, prm


def random_channels(count=256, seed=0, imax=(64, 512), jmax=(32, 256), re=(50.0, 1000.0),
                    eps=1e-5, omega=1.7, tau=0.5, check_every=0, itermax=1_000_000):
    """`count` channels: imax and jmax drawn independently log-uniform in
    their ranges; 1-4 NoSlip rectangles (sides >= 2, kept off the inflow and
    outflow columns, redrawn if a boundary cell would have fluid on opposite
    sides); Re uniform; inflow u = 1 on the west, outflow east, no-slip top
    and bottom; the channel is 1 unit high (dy = 1/jmax, dx = dy); adaptive
    step with tau; red-black SOR to a relative residual eps."""
    rng = presets.SplitMix64(seed)
    out = []
    for _ in range(count):
        ni = presets.log_uniform_int(rng, *imax)
        nj = presets.log_uniform_int(rng, *jmax)
        r = re[0] + rng.uniform() * (re[1] - re[0])
        ct, vel = presets.random_channel(rng, ni, nj)
        h = 1.0 / nj
        prm = make_params(h, h, 0.0, 0.9, r, eps, itermax, omega, tau=tau, solver="rb",
                          check_every=check_every)
        out.append((ct, vel, prm))
    return out


def random_cavities(count, n=256, seed=0, re=(50.0, 1000.0), eps=1e-6, omega=1.8, tau=0.5,
                    check_every=0, itermax=1_000_000):
    """Lid cavities n x n with 1-4 random rectangles and Re uniform."""
    rng = presets.SplitMix64(seed)
    out = []
    for _ in range(count):
        r = re[0] + rng.uniform() * (re[1] - re[0])
        ct, vel = presets.random_cavity(rng, n, n)
        prm = make_params(1.0 / n, 1.0 / n, 0.0, 0.9, r, eps, itermax, omega, tau=tau,
                          solver="rb", check_every=check_every)
        out.append((ct, vel, prm))
    return out


def fixed_channels(count=64, seed=1, imax=512, jmax=256, re=(100.0, 1000.0), eps=1e-7,
                   omega=1.7, tau=0.5, check_every=0, itermax=10_000_000):
    """`count` channels of imax x jmax interior cells with 1-4 random NoSlip
    rectangles (as random_channels), Re uniform in `re`, inflow u = 1 west,
    outflow east, 1 unit high, red-black SOR to relative residual eps."""
    rng = presets.SplitMix64(seed)
    out = []
    for _ in range(count):
        r = re[0] + rng.uniform() * (re[1] - re[0])
        ct, vel = presets.random_channel(rng, imax, jmax)
        h = 1.0 / jmax
        prm = make_params(h, h, 0.0, 0.9, r, eps, itermax, omega, tau=tau, solver="rb",
                          check_every=check_every)
        out.append((ct, vel, prm))
    return out
