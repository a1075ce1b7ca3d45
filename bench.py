"""Benchmark: cell-updates/s of the NaSt2D time step at a fixed Poisson residual.

Default workload: lid-driven cavity, 1024 x 1024 cells on the unit square,
Re = 1000, lid u = 1, tau = 0.5 (adaptive step), gamma = 0.9, red-black SOR
with omega = 1.8 to a fluid-masked residual of 1e-6 relative to rhs, FP64,
zero initial fields.  A step is one full tick (boundary conditions, F/G,
RHS, pressure solve to the residual, velocity update, next step size).

    python bench.py [--gpus 1] [--steps K] [--warmup W]
    python bench.py --profile        # fixed sweeps + one 256 MiB copy, for rocprofv3
    python bench.py --workload channels           # 256 uneven channels, plain
    python bench.py --workload learned_cavities   # 32 cavities 256x256: plain vs learned
    python bench.py --workload learned_cavities --no-graphs   # cycles launched, not replayed
    python bench.py --workload learned_channels   # 64 channels 512x256: plain vs learned
    python bench.py --workload learned_channels_500   # the same, 500 learned steps from rest
    python bench.py --workload learned_uneven     # the 256 uneven channels: plain vs learned
    python bench.py --workload sweeps --n 4096    # the blocked kernel alone (HBM-resident at 4096)
    python bench.py --workload sync               # relaunch vs graph vs resident-kernel waits
    python bench.py --workload c1                 # square obstacle, reference lexicographic solve

Prints one JSON line.  `value` = interior cells * K / wall seconds of the K
timed ticks, inputs resident in HBM.  The roofline entry is for the dominant
kernel (the blocked red-black SOR launch), timed with HIP events on the
library's stream inside the timed region; the CPU baseline is the oracle's
lexicographic SOR (the reference algorithm, one core) timed on a bounded
sample of the same grid.  The default line also carries "learned": the
plain solve against the learned two-grid solve on 32 random-obstacle lid
cavities of 256 x 256 (the network trained first, on the GPU, from pairs the
solver produces on other layouts), both to the same residual; and
"learned_same_grid": that network's learned solve on the benchmark's own grid,
the same ticks timed as for `value`, each solve to the same residual bound.
Two more summaries ride along: "channels", the 256 uneven channels advanced
together (two ticks), and "learned_channels", plain against learned on the 64
channels of 512 x 256 at eps 1e-7 (a shorter training than
--workload learned_channels).  --no-channels / --no-learned leave them out.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s at fixed Poisson residual: plain GPU SOR vs learned+batched"
HBM_PEAK_GBS = 8000.0
# bytes per interior cell of one red-black sweep streamed once per sweep
# (each half-sweep reads own p, other p, own rhs and writes own p at 4 B/cell,
# plus 1 B flags): the rate a kernel without temporal blocking would have to
# stream for the same progress -- reported as "effective_sweep_gbs", not as
# the roofline's achieved rate
SWEEP_BYTES_PER_CELL = 33.0
# stage kernels: algorithmic bytes per cell (f64 8 B, flags 1 B, each value once)
# (fg_rhs: u, v, flags read, f, g, rhs written; adap_ranges: f, g, p, flags read,
# u, v written -- the ranges add no traffic)
STAGE_BYTES_PER_CELL = {"fg": 33.0, "rhs": 24.0, "fg_rhs": 41.0, "adap_ranges": 41.0,
                        "ranges": 25.0}
COPY_PROBE_BYTES = 256 << 20
# template arguments (R, HY, NS) of each blocked-solver variant, after K
VARIANT_SHAPES = {0: "4, 32, 16", 1: "4, 32, 16", 2: "4, 64, 16", 3: "4, 32, 16",
                  4: "4, 64, 16", 5: "6, 64, 16", 6: "6, 64, 16", 7: "2, 64, 16",
                  8: "12, 64, 8", 9: "6, 48, 16"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1024, help="interior cells per side")
    ap.add_argument("--eps", type=float, default=1e-6)
    ap.add_argument("--check-every", type=int, default=0,
                    help="sweeps per blocked launch (1/2/4/6/8); 0 = autotuned")
    ap.add_argument("--sample", type=int, default=16, help="time every n-th launch")
    ap.add_argument("--cpu-sweeps", type=int, default=150)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile", action="store_true",
                    help="short fixed run for hardware counters, no JSON line")
    ap.add_argument("--variant", type=int, default=-1,
                    help="with --profile or --workload sweeps: this blocked-solver variant "
                         "instead of the autotuner's choice (counter collection changes the "
                         "timings the tuner sees)")
    ap.add_argument("--workload", choices=("cavity", "channels", "learned_cavities",
                                           "learned_channels", "learned_channels_500",
                                           "learned_uneven", "net", "sweeps", "sync", "c1"),
                    default="cavity",
                    help="cavity: the 1024x1024 lid cavity (default); channels: 256 "
                         "independent channels of uneven size, seed 0; learned_*: plain vs "
                         "learned solve; sync: launch, graph and resident-kernel wait "
                         "costs")
    ap.add_argument("--channels", type=int, default=256)
    ap.add_argument("--no-graphs", action="store_true",
                    help="learned_cavities: the learned ticks launch every cycle instead of "
                         "replaying hipGraphs (for kernel traces)")
    ap.add_argument("--no-channels", action="store_true",
                    help="leave the 256-channel summary out of the default line")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned comparison of the default run")
    return ap.parse_args()


def host_info():
    """nproc and CPU model of the machine the CPU baseline ran on."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def cpu_parallel(n, sweeps_per_step, nsweeps):
    """Not the reference algorithm: the oracle's red-black sweeps with each
    half-sweep split over the host threads (OMP_NUM_THREADS, else every core;
    bit-identical to the serial sweeps), timed for `nsweeps` sweeps on the
    same grid and extrapolated to the GPU's red-black sweeps per step, plus one
    tick's other stages on one core."""
    import numpy as np

    from oracle import pyoracle as O
    from stroemung_amd import presets

    threads = int(os.environ.get("OMP_NUM_THREADS", 0)) or os.cpu_count()
    ct, vel = presets.lid_cavity(n, n)
    sim = O.Sim(ct, vel, delx=1.0 / n, dely=1.0 / n, delt=1e-4, gamma=0.9, reynolds=1000.0,
                sor_absolute_epsilon=0.0, max_iterations=1, omega=1.8, solver=O.SOLVER_RB)
    sim.u[...] = np.random.default_rng(0).uniform(-0.1, 0.1, sim.u.shape)
    assert sim.call("rebuild_boundary_list") == 0
    t0 = time.perf_counter()
    sim.call("set_boundary_u_and_v")
    sim.call("calculate_f_and_g")
    sim.call("calculate_rhs")
    sim.set(rhs_mean=O.lib().orc_rhs_mean(sim.s))
    t1 = time.perf_counter()
    sim.call("rb_sweeps_mt", nsweeps, threads)
    t2 = time.perf_counter()
    sim.call("set_u_and_v")
    t3 = time.perf_counter()
    t_stages = (t1 - t0) + (t3 - t2)
    t_sweep = (t2 - t1) / nsweeps
    return {
        "value": n * n / (t_stages + sweeps_per_step * t_sweep),
        "unit": "cell-updates/s",
        "cores": threads,
        "kind": "port, red-black (not the reference algorithm)",
        "host": host_info(),
        "sample": ("oracle red-black sweeps on %d threads on the %dx%d cavity: %d sweeps "
                   "timed (%.3f ms/sweep) plus one tick's other stages on one core (%.1f ms); "
                   "extrapolated to %.0f sweeps per step as measured on the GPU"
                   % (threads, n, n, nsweeps, t_sweep * 1e3, t_stages * 1e3, sweeps_per_step)),
    }


def cpu_baseline(n, sweeps_per_step, nsweeps):
    """Oracle (plain-C port of the reference, one core) on the same grid:
    one tick's non-SOR stages plus `nsweeps` lexicographic SOR sweeps,
    extrapolated to the GPU's measured sweeps per step."""
    import numpy as np

    from oracle import pyoracle as O
    from stroemung_amd import presets

    ct, vel = presets.lid_cavity(n, n)
    kw = dict(delx=1.0 / n, dely=1.0 / n, delt=1e-4, gamma=0.9, reynolds=1000.0,
              sor_absolute_epsilon=0.0, max_iterations=nsweeps, omega=1.8,
              masked_stop=1)
    sim = O.Sim(ct, vel, **kw)
    sim.u[...] = np.random.default_rng(0).uniform(-0.1, 0.1, sim.u.shape)
    assert sim.call("rebuild_boundary_list") == 0
    t0 = time.perf_counter()
    sim.call("set_boundary_u_and_v")
    sim.call("calculate_f_and_g")
    sim.call("calculate_rhs")
    t1 = time.perf_counter()
    sim.solve()
    t2 = time.perf_counter()
    sim.call("set_u_and_v")
    t3 = time.perf_counter()
    t_stages = (t1 - t0) + (t3 - t2)
    t_sweep = (t2 - t1) / nsweeps
    step_s = t_stages + sweeps_per_step * t_sweep
    return {
        "value": n * n / step_s,
        "unit": "cell-updates/s",
        "cores": 1,
        "kind": "port",
        "host": host_info(),
        "sample": ("oracle lexicographic SOR (reference algorithm, masked residual test each "
                   "sweep) on the %dx%d cavity: %d sweeps timed (%.3f ms/sweep) plus one "
                   "tick's BC/FG/RHS/ADAP_UV (%.1f ms); extrapolated to %.0f sweeps per "
                   "step as measured on the GPU" % (n, n, nsweeps, t_sweep * 1e3,
                                                   t_stages * 1e3, sweeps_per_step)),
    }


def make_context(a):
    from stroemung_amd import presets
    from stroemung_amd.simulation import Context, make_params

    n = a.n
    ct, vel = presets.lid_cavity(n, n)
    prm = make_params(1.0 / n, 1.0 / n, 0.0, 0.9, 1000.0, a.eps, 10_000_000, 1.8, tau=0.5,
                      solver="rb", check_every=a.check_every)
    ctx = Context(0)
    ch = ctx.add_grid(prm, ct, vel)
    ctx.construct()
    return ctx, ch


PROFILE_TICK_SWEEPS = 96


def profile_run(a):
    from stroemung_amd.simulation import make_params
    """One full tick; then 8 ticks whose solves stop at PROFILE_TICK_SWEEPS
    sweeps (every stage kernel 8 times, on the benchmark's grid and fields);
    the tuned variant's 64 blocked launches of fixed sweeps; one copy of
    COPY_PROBE_BYTES (the byte-counter reference)."""
    ctx, ch = make_context(a)
    ctx.tick()
    n = a.n
    ctx.set_params(ch, make_params(1.0 / n, 1.0 / n, 0.0, 0.9, 1000.0, a.eps,
                                   PROFILE_TICK_SWEEPS, 1.8, tau=0.5, solver="rb",
                                   check_every=a.check_every))
    for _ in range(8):
        ctx.tick()
    if a.variant >= 0:
        ctx.solver_variant(a.variant)
    else:
        ctx.autotune(8)
    variant, k, tx, ty = ctx.solver_variant()
    ctx.stage("rb_sweeps", 64 * k)
    gbs = ctx.stream_copy(COPY_PROBE_BYTES, 1)
    ctx.synchronize()
    print("profile run: 1 + 8 ticks (%d sweeps each after the first); variant %d (K=%d, "
          "tile %dx%d), %d sweeps; copy probe %d bytes, %.0f GB/s"
          % (PROFILE_TICK_SWEEPS, variant, k, tx, ty, 64 * k, COPY_PROBE_BYTES, gbs), flush=True)


def sweeps_run(a):
    """The blocked red-black kernel alone: fixed sweeps on the n x n lid
    cavity with random p and rhs (no tick), timed with HIP events on the
    library's stream.  At n = 4096 the solve's working set (p, second p
    buffer, rhs, flags: 420 MB) exceeds the 256 MB Infinity Cache, so the
    kernel runs from HBM; n = 1024 (26 MB) is cache-resident."""
    import numpy as np

    n = a.n
    ctx, ch = make_context(a)
    rng = np.random.default_rng(0)
    ctx.upload(ch, p=rng.standard_normal((n + 2, n + 2)),
               rhs=rng.standard_normal((n + 2, n + 2)))
    if a.variant >= 0:
        ctx.solver_variant(a.variant)
        tuned = (a.variant, None)
    else:
        tuned = ctx.autotune(8)
    variant, k, tx, ty = ctx.solver_variant()
    launches = 64
    ctx.stage("rb_sweeps", 8 * k)                     # warm-up
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.stage("rb_sweeps", launches * k)
    ctx.synchronize()
    t1 = time.perf_counter()
    # wall time per working launch: the host polls, the launch that takes
    # the stop decision and the queued launches after it are included, so
    # this slightly understates the kernel (the rocprofv3 trace gives it alone)
    avg_s = (t1 - t0) / launches
    unique, loaded = ctx.solver_bytes()
    achieved = unique / avg_s / 1e9
    effective = SWEEP_BYTES_PER_CELL * k * n * n / avg_s / 1e9
    working_set = (3 * 8 + 1) * (n + 2) * (n + 2)
    ctx.close()
    out = {
        "metric": "blocked red-black SOR launch, algorithmic bytes per second",
        "value": achieved, "unit": "GB/s", "n_gpus": 1, "steps": launches, "warmup": 8,
        "ms_per_step": avg_s * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "%d K-sweep launches on the %dx%d lid cavity, random p and rhs"
                               % (launches, n, n),
                   "variant": {"id": variant, "sweeps_per_launch": k, "tile": [tx, ty],
                               "autotune_us_per_sweep": tuned[1]},
                   "wall_us_per_sweep": (t1 - t0) * 1e6 / (launches * k),
                   "timing": "wall time of the sweeps stage / working launches"},
        "roofline": {"kernel": "k_rb_strip<%d, %s>" % (k, VARIANT_SHAPES.get(variant, "?")),
                     "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_launch": unique, "loaded_bytes_per_launch": loaded,
                     "halo_factor": loaded / unique, "effective_sweep_gbs": effective,
                     "avg_launch_us": avg_s * 1e6, "working_set_bytes": working_set,
                     "cache_resident": working_set < MALL_BYTES},
    }
    print(json.dumps(out), flush=True)


def sync_run(a):
    """What keeping the blocked solver resident between its K-sweep steps
    would cost, against relaunching it: 1000 steps of the blocked kernel's
    launch shape on the 1026 x 1026 grid (256 workgroups of 768 threads, each
    writing a 64 x 64 FP64 tile), as separate launches, replayed from a
    hipGraph, and inside one resident kernel that waits at a grid-wide
    barrier or for its four neighbours' flags; HIP events on the library's
    stream."""
    from stroemung_amd.simulation import Context

    steps = 1000
    ctx = Context(0)
    us = ctx.sync_benchmark(steps)
    ctx.close()
    out = {
        "metric": "cost per step of the blocked kernel's launch shape",
        "value": us["launch"], "unit": "us", "n_gpus": 1, "steps": steps, "warmup": 10,
        "ms_per_step": us["launch"] * 1e-3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "256 workgroups x 768 threads writing 64x64 FP64 tiles, "
                               "%d steps" % steps},
        "us_per_step": us,
    }
    print(json.dumps(out), flush=True)


def c1_cpu_baseline(steps, got):
    """Oracle (plain-C port of the reference, one core) on the square-obstacle
    channel for the same `steps` lexicographic ticks; its final u, v and p
    are also the check that the GPU's `got` fields are bit-identical."""
    import numpy as np

    from oracle import pyoracle as O
    from stroemung_amd import presets

    O.set_threads(1)
    ct, vel = presets.square_obstacle()
    sim = O.Sim(ct, vel, solver=O.SOLVER_LEX, **C1_PARAMS).construct()
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.tick()
    el = time.perf_counter() - t0
    same = all(np.array_equal(np.asarray(got[k]).view(np.uint64),
                              np.asarray(getattr(sim, k)).view(np.uint64)) for k in ("u", "v", "p"))
    interior = (ct.shape[0] - 2) * (ct.shape[1] - 2)
    return {"value": interior * steps / el, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "host": host_info(), "ms_per_step": el * 1e3 / steps,
            "sample": "the whole workload: %d lexicographic ticks of the oracle" % steps}, same


# the square-obstacle channel with the reference's settings: 100 x 20 interior
# cells, xlength 10, ylength 2, Re 100, gamma 0.9, omega 1.7, eps 1e-3, 100
# iterations, fixed step 0.005
C1_PARAMS = dict(delx=0.1, dely=0.1, delt=0.005, gamma=0.9, reynolds=100.0,
                 sor_absolute_epsilon=1e-3, max_iterations=100, omega=1.7)


def c1_run(a):
    """The reference's own CPU-runnable case: flow past a square obstacle,
    100 x 20 cells, 100 ticks of the reference's lexicographic solve (the
    same bits as the reference).  One channel on the GPU, then 256 copies
    advanced together (one workgroup each), with the oracle's one-core time on
    the same ticks beside them.  The solve is a serial chain (each cell waits
    for its W and N neighbours, then the residual is summed in order), so one
    channel is latency-bound; the GPU's gain is in many channels at once."""
    from stroemung_amd import presets, workloads
    from stroemung_amd.simulation import Context

    steps, batch = 100, 256
    ct, vel = presets.square_obstacle()
    interior = (ct.shape[0] - 2) * (ct.shape[1] - 2)
    prm = workloads.square_channel_params()
    timings = {}
    got = None
    for n in (1, batch):
        ctx = Context(0)
        for _ in range(n):
            ctx.add_grid(prm, ct, vel)
        ctx.construct()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.tick()
        ctx.synchronize()
        timings[n] = time.perf_counter() - t0
        if n == 1:
            got = ctx.download(0)
        ctx.close()
    value = interior * steps / timings[1]
    out = {
        "metric": "cell-updates/s at fixed Poisson residual, reference lexicographic solve",
        "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": timings[1] * 1e3 / steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "flow past a square obstacle, 100x20 cells, Re 100, omega 1.7, "
                               "eps 1e-3, 100 iterations per tick, fixed step 0.005, %d ticks, "
                               "lexicographic SOR (reference algorithm)" % steps},
        "batched": {"channels": batch, "ms_per_step": timings[batch] * 1e3 / steps,
                    "cell_updates_per_s": batch * interior * steps / timings[batch]},
    }
    if not a.no_cpu:
        cpu, same = c1_cpu_baseline(steps, got)
        out["cpu_baseline"] = cpu
        out["bit_identical_to_oracle"] = same
    print(json.dumps(out), flush=True)


def channels_run(a):
    """256 channels (imax 64-512, jmax 32-256 log-uniform, 1-4 random obstacles,
    Re 50-1000, eps 1e-5, seed 0) advanced together, one blocked launch per K
    sweeps over every running channel."""
    from stroemung_amd import workloads
    from stroemung_amd.simulation import Context

    ws = workloads.random_channels(count=a.channels, seed=0, eps=1e-5,
                                   check_every=a.check_every)
    ctx = Context(0)
    for ct, vel, prm in ws:
        ctx.add_grid(prm, ct, vel)
    ctx.construct()
    cells = [(ct.shape[0] - 2) * (ct.shape[1] - 2) for ct, _, _ in ws]
    for _ in range(a.warmup):
        ctx.tick()
    # whole-tick tuning: most of a tick runs with few channels left
    tuned = ctx.autotune_ticks()
    variant, ksweeps, tile_x, tile_y = ctx.solver_variant()
    ctx.synchronize()
    ctx.reset_stats()
    ctx.profile(a.sample)
    swept = 0.0
    per_tick = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        its, _ = ctx.tick()
        swept += sum(c * i for c, i in zip(cells, its))
        per_tick.append(its)
    ctx.synchronize()
    t1 = time.perf_counter()
    ctx.profile(0)
    elapsed = t1 - t0
    total = sum(cells)
    # batches of queued launches timed as a whole, as for the cavity in main()
    ms, timed, _ = ctx.kernel_stats("rb_tb_batch")
    sms, stimed, launches = ctx.kernel_stats("rb_tb")
    if not timed:
        ms, timed = sms, stimed
    avg_s = ms / 1e3 / max(timed, 1)
    # every launch moves at most each cell of every channel once (25 B); launches
    # after compaction move only the running channels', so this is an upper bound
    unique, loaded = ctx.solver_bytes()
    achieved = unique / avg_s / 1e9 if timed else None
    effective = SWEEP_BYTES_PER_CELL * swept / (avg_s * launches) / 1e9 if timed else None
    flat = [i for t in per_tick for i in t]
    out = {
        "metric": METRIC,
        "value": total * a.steps / elapsed,
        "unit": "cell-updates/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": elapsed * 1e3 / a.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {
            "workload": "%d channels, imax 64-512 x jmax 32-256 log-uniform, 1-4 random "
                        "obstacles, Re 50-1000, inflow u=1, red-black SOR omega=1.7 to relative "
                        "residual 1e-5, tau=0.5, fp64, seed 0" % len(ws),
            "interior_cells": total,
            "variant": {"id": variant, "sweeps_per_launch": ksweeps, "tile": [tile_x, tile_y],
                        "autotune": "one tick per variant",
                        "tuning_cell_sweeps_per_s": tuned[1]},
            "sweeps_per_channel_tick": {"mean": sum(flat) / len(flat), "max": max(flat),
                                        "min": min(flat)},
            "cell_sweeps_per_s": swept / elapsed,
        },
        "roofline": {
            "kernel": "k_rb_strip<%d, %s>" % (ksweeps, VARIANT_SHAPES.get(variant, "?")),
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": None,
            "bytes_per_launch": unique, "loaded_bytes_per_launch": loaded,
            "halo_factor": loaded / unique,
            "model": "25 B per cell of every channel per launch (an upper bound: launches "
                     "after compaction cover only the running channels)",
            "effective_sweep_gbs": effective,
            "avg_launch_us": avg_s * 1e6, "launches": launches,
        },
    }
    ctx.close()
    print(json.dumps(out), flush=True)


MFMA_BF16_PEAK_TFLOPS = 2500.0
MALL_BYTES = 256 << 20          # MI355X Infinity Cache


def train_cavity_net(log=None):
    """The network of the learned cavity comparisons: trained on the pairs
    the solver produces on 64 random-obstacle lid cavities of 256 x 256
    (seed 0), coarsening 4.  Returns (params, info)."""
    from stroemung_amd import learned, workloads
    train = workloads.random_cavities(64, n=256, seed=0, eps=1e-6)
    return learned.train_net(train, coarsen=4, capture_ticks=48, points=4, capture_sweeps=32,
                             steps=3000, batch=32, dagger_rounds=2, dagger_ticks=16,
                             dagger_steps=1500, cycles=32, sweeps=16, step="momentum",
                             log=log or (lambda *x: None))


def learned_cavities(ticks=10, log=None, trained=None, graphs=True):
    """32 random-obstacle lid cavities of 256 x 256 cells, Re 50-1000, eps
    1e-6: network trained on 64 seed-0 layouts, compared on seed-100 layouts."""
    from stroemung_amd import learned, workloads
    log = log or (lambda *x: None)
    params, info = trained or train_cavity_net(log)
    test = workloads.random_cavities(32, n=256, seed=100, eps=1e-6)
    res = learned.compare(test, params, coarsen=4, cycles=32, sweeps=16, warm=20, ticks=ticks,
                          step="momentum", graphs=graphs, log=log)
    res["training"] = info
    res["workload"] = ("32 lid cavities 256x256 with 1-4 random obstacles, Re 50-1000, "
                       "red-black SOR omega=1.8 to relative residual 1e-6, tau=0.5, fp64; "
                       "trained on 64 seed-0 layouts (20 480 pairs), measured on seed 100 "
                       "after 20 plain ticks")
    return res


def learned_same_grid(a, params, plain_value, plain_sweeps, log=None):
    """The learned solve on the benchmark's own grid (the n x n lid cavity),
    with the network trained on 256 x 256 cavities (fully convolutional, so it
    runs on an (n/16) x (n/16) frame at coarsening 16): the same warm-up tick and
    tuner as the plain run, then the same ticks timed, each solve to the same
    residual bound."""
    from stroemung_amd.learned import PressureNet
    n, f = a.n, 16
    ctx, ch = make_context(a)
    net = PressureNet(ctx, n // f, n // f, 1, seed=1)
    try:
        net.set_params(params)
        for _ in range(a.warmup):
            ctx.tick()
        ctx.autotune(8)
        ctx.set_graphs(True)
        # 48 sweeps per cycle: fastest of 24-128 on this grid with K = 6 launches
        # (fewer cycles, so fewer network forwards, until the extra sweeps cost more)
        ctx.set_learned(net, cycles=256, sweeps=48, step="momentum", coarsen=f)
        ctx.synchronize()
        sweeps, bound = 0, True
        t0 = time.perf_counter()
        for _ in range(a.steps):
            its, _ = ctx.tick()
            sweeps += its[ch]
            st = ctx.state(ch)
            bound = bound and st["residual_squared"] <= st["tolerance_squared"] \
                and not st["exhausted"]
        ctx.synchronize()
        el = time.perf_counter() - t0
        ctx.set_learned(None)
    finally:
        net.close()
        ctx.close()
    value = n * n * a.steps / el
    if log:
        log("same grid: learned %.2f ms/tick, %.0f sweeps/step, %.1fx the plain run"
            % (el * 1e3 / a.steps, sweeps / a.steps, value / plain_value))
    return {"workload": "the %dx%d lid cavity above, same ticks, same residual bound; network "
                        "trained on 64 random-obstacle 256x256 cavities (coarsening 4), run "
                        "here on a %dx%d frame at coarsening %d" % (n, n, n // f, n // f, f),
            "cell_updates_per_s": value, "ms_per_tick": el * 1e3 / a.steps,
            "sweeps_per_step": sweeps / a.steps, "plain_sweeps_per_step": plain_sweeps,
            "speedup_vs_plain": value / plain_value, "residual_bound_met": bound,
            "cycles": 256, "sweeps_per_cycle": 48, "step": "momentum", "coarsen": f,
            "graphs": True}


def learned_channels(ticks, log=None):
    """64 channels 512 x 256, 1-4 random obstacles, Re 100-1000, eps 1e-7:
    network trained on seed-1001 layouts, compared on seed 1."""
    from stroemung_amd import learned, workloads
    log = log or (lambda *x: None)
    train = workloads.fixed_channels(64, seed=1001, eps=1e-7)
    params, info = learned.train_net(train, coarsen=8, capture_ticks=40, points=4,
                                     capture_sweeps=32, steps=3000, batch=32, dagger_rounds=2,
                                     dagger_ticks=24, dagger_steps=1500, cycles=64, sweeps=32,
                                     step="momentum", log=log)
    test = workloads.fixed_channels(64, seed=1, eps=1e-7)
    res = learned.compare(test, params, coarsen=8, cycles=64, sweeps=32, warm=10, ticks=ticks,
                          step="momentum", log=log)
    res["training"] = info
    res["workload"] = ("64 channels 512x256 with 1-4 random obstacles, Re 100-1000, inflow u=1, "
                       "red-black SOR omega=1.7 to relative residual 1e-7, tau=0.5, fp64; "
                       "trained on seed 1001, measured on seed 1 after 10 plain ticks")
    return res


def channels_summary(a):
    """The 256 uneven channels (as --workload channels) in brief for the
    default line: one warm-up tick, the whole-tick tuner, two timed ticks."""
    from stroemung_amd import workloads
    from stroemung_amd.simulation import Context

    ws = workloads.random_channels(count=256, seed=0, eps=1e-5, check_every=a.check_every)
    ctx = Context(0)
    try:
        for ct, vel, prm in ws:
            ctx.add_grid(prm, ct, vel)
        ctx.construct()
        cells = [(ct.shape[0] - 2) * (ct.shape[1] - 2) for ct, _, _ in ws]
        ctx.tick()
        ctx.autotune_ticks()          # as --workload channels: the whole-tick tuner
        variant, k, tx, ty = ctx.solver_variant()
        ctx.synchronize()
        its_all, bound = [], True
        t0 = time.perf_counter()
        for _ in range(2):
            its, _ = ctx.tick()
            its_all += its
        ctx.synchronize()
        el = time.perf_counter() - t0
        for c in range(len(ws)):
            st = ctx.state(c)
            bound = bound and st["residual_squared"] <= st["tolerance_squared"] \
                and not st["exhausted"]
    finally:
        ctx.close()
    return {"workload": "256 channels, imax 64-512 x jmax 32-256 log-uniform, 1-4 random "
                        "obstacles, Re 50-1000, eps 1e-5, seed 0 (as --workload channels): "
                        "1 warm-up tick, the whole-tick tuner, 2 timed ticks; variant %d (K=%d)"
                        % (variant, k),
            "cell_updates_per_s": sum(cells) * 2 / el, "ms_per_tick": el * 1e3 / 2,
            "interior_cells": sum(cells),
            "sweeps_per_channel_tick": {"mean": sum(its_all) / len(its_all),
                                        "max": max(its_all), "min": min(its_all)},
            "residual_bound_met": bound}


def learned_channels_summary(log=None):
    """The 64 channels 512 x 256 at eps 1e-7 (as --workload learned_channels)
    in brief for the default line: a shorter training (half the capture ticks
    and Adam steps, one DAgger round) and 4 plain warm-up ticks, then the
    plain and learned solves over the same 3 ticks, each solve to the same
    residual bound."""
    from stroemung_amd import learned, workloads
    log = log or (lambda *x: None)
    train = workloads.fixed_channels(64, seed=1001, eps=1e-7)
    params, info = learned.train_net(train, coarsen=8, capture_ticks=20, points=4,
                                     capture_sweeps=32, steps=1500, batch=32, dagger_rounds=1,
                                     dagger_ticks=12, dagger_steps=800, cycles=64, sweeps=32,
                                     step="momentum", log=log)
    test = workloads.fixed_channels(64, seed=1, eps=1e-7)
    res = learned.compare(test, params, coarsen=8, cycles=64, sweeps=32, warm=4, ticks=3,
                          step="momentum", log=log)
    pl, lr = res["plain"], res["learned"]
    return {"workload": "64 channels 512x256, 1-4 random obstacles, Re 100-1000, eps 1e-7 "
                        "(as --workload learned_channels, shorter training): trained on seed "
                        "1001, measured on seed 1 over 3 ticks after 4 plain ticks",
            "plain_cell_updates_per_s": pl["cell_updates_per_s"],
            "learned_cell_updates_per_s": lr["cell_updates_per_s"],
            "speedup": res["speedup"],
            "plain_sweeps_per_solve": pl["sweeps_per_solve"],
            "learned_sweeps_per_solve": lr["sweeps_per_solve"],
            "residual_bound_met": pl["residual_bound_met"] and lr["residual_bound_met"],
            "training_wall_s": info.get("wall_s")}


def learned_uneven(ticks, log=None):
    """The 256 uneven channels of the batched workload (64x32 to 512x256, seed
    0, eps 1e-5): network trained on the 256 channels of seed 1001, one frame
    of 64 x 32 at coarsening 8 for every size; both runs whole-tick tuned."""
    from stroemung_amd import learned, workloads
    log = log or (lambda *x: None)
    train = workloads.random_channels(count=256, seed=1001, eps=1e-5)
    params, info = learned.train_net(train, coarsen=8, capture_ticks=24, points=4,
                                     capture_sweeps=32, steps=3000, batch=32, dagger_rounds=2,
                                     dagger_ticks=12, dagger_steps=1500, cycles=64, sweeps=32,
                                     step="momentum", log=log)
    test = workloads.random_channels(count=256, seed=0, eps=1e-5)
    res = learned.compare(test, params, coarsen=8, cycles=64, sweeps=32, warm=1, ticks=ticks,
                          step="momentum", tune="ticks", log=log)
    res["training"] = info
    res["workload"] = ("256 channels, imax 64-512 x jmax 32-256 log-uniform, 1-4 random "
                       "obstacles, Re 50-1000, inflow u=1, red-black SOR omega=1.7 to relative "
                       "residual 1e-5, tau=0.5, fp64, seed 0; trained on the 256 channels of "
                       "seed 1001; measured after 1 plain tick and 8 tuning ticks")
    return res


def learned_channels_500(log=None):
    """The learned + batched trajectory: the 64 channels 512 x 256 of seed 1 advanced 500
    learned ticks from rest, with plain ticks timed on the same state at
    checkpoints; network trained as for learned_channels."""
    from stroemung_amd import learned, workloads
    log = log or (lambda *x: None)
    train = workloads.fixed_channels(64, seed=1001, eps=1e-7)
    params, info = learned.train_net(train, coarsen=8, capture_ticks=40, points=4,
                                     capture_sweeps=32, steps=3000, batch=32, dagger_rounds=2,
                                     dagger_ticks=24, dagger_steps=1500, cycles=64, sweeps=32,
                                     step="momentum", log=log)
    test = workloads.fixed_channels(64, seed=1, eps=1e-7)
    res = learned.trajectory(test, params, coarsen=8, cycles=64, sweeps=32, step="momentum",
                             ticks=500, log=log)
    res["training"] = info
    return res


def learned_500_line(a):
    log = lambda *x: print(*x, file=sys.stderr, flush=True)  # noqa: E731
    res = learned_channels_500(log)
    lr = res["learned"]
    out = {
        "metric": METRIC,
        "value": lr["cell_updates_per_s"],
        "unit": "cell-updates/s",
        "n_gpus": 1, "steps": res["ticks"], "warmup": 1,
        "ms_per_step": lr["ms_per_tick"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (solve), bf16 (network)", "data": "synthetic",
        "config": {"workload": "64 channels 512x256 with 1-4 random obstacles, Re 100-1000, "
                               "inflow u=1, red-black SOR omega=1.7 to relative residual 1e-7, "
                               "tau=0.5, fp64, seed 1: 500 learned steps from rest, plain "
                               "steps timed on the same state at checkpoints; trained on "
                               "seed 1001",
                   "learned": lr, "checkpoints": res["checkpoints"],
                   "training": res["training"]},
        "roofline": None,
    }
    print(json.dumps(out), flush=True)


def net_line(a):
    """The U-Net's forward and training step alone, on the frames the learned
    solves use: TFLOP/s against the BF16 dense MFMA peak."""
    from stroemung_amd.learned import PressureNet
    from stroemung_amd.simulation import Context
    ctx = Context(0)
    rows = []
    for (h, w, nb) in ((64, 64, 32), (128, 64, 64), (256, 256, 8)):
        net = PressureNet(ctx, h, w, nb, seed=1)
        fwd, trn = net.benchmark(nb, max(a.steps, 3), train=True)
        flops = net.stats()["flops_forward_per_sample"] * nb
        rows.append({"frame": [h, w], "batch": nb, "forward_ms": fwd, "train_step_ms": trn,
                     "forward_tflops": flops / fwd / 1e9,
                     "train_tflops": 3 * flops / trn / 1e9,
                     "forward_frac": flops / fwd / 1e9 / MFMA_BF16_PEAK_TFLOPS})
        net.close()
    ctx.close()
    best = max(rows, key=lambda r: r["forward_tflops"])
    out = {
        "metric": "U-Net forward TFLOP/s (bf16 MFMA, fp32 accumulate)",
        "value": rows[0]["forward_tflops"], "unit": "TFLOP/s", "n_gpus": 1,
        "steps": max(a.steps, 3), "warmup": 1, "ms_per_step": rows[0]["forward_ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
        "data": "synthetic",
        "config": {"workload": "4-level U-Net, 32 base channels, frames of the learned solves",
                   "frames": rows},
        "roofline": {"kernel": "U-Net forward", "bound": "mfma",
                     "achieved": best["forward_tflops"], "peak": MFMA_BF16_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": best["forward_frac"], "traffic": None},
    }
    print(json.dumps(out), flush=True)


def learned_line(a):
    log = lambda *x: print(*x, file=sys.stderr, flush=True)  # noqa: E731
    if a.workload == "learned_cavities":
        res = learned_cavities(a.steps, log, graphs=not a.no_graphs)
    elif a.workload == "learned_uneven":
        res = learned_uneven(a.steps, log)
    else:
        res = learned_channels(a.steps, log)
    lr = res["learned"]
    out = {
        "metric": METRIC,
        "value": lr["cell_updates_per_s"],
        "unit": "cell-updates/s",
        "n_gpus": 1, "steps": a.steps, "warmup": res["warm"],
        "ms_per_step": lr["ms_per_tick"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (solve), bf16 (network)", "data": "synthetic",
        "config": {"workload": res.pop("workload"), "plain": res["plain"],
                   "speedup_vs_plain": res["speedup"], "sweeps_saved": res["sweeps_saved"],
                   "training": res["training"], "learned": lr},
        "roofline": {
            "kernel": "U-Net forward (k_conv3x3 and companions)", "bound": "mfma",
            "achieved": lr["forward_tflops"], "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": (lr["forward_tflops"] / MFMA_BF16_PEAK_TFLOPS) if lr["forward_tflops"] else None,
            "traffic": None,
        },
    }
    print(json.dumps(out), flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(--gpus %d requested)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    if a.profile:
        profile_run(a)
        return
    if a.workload == "channels":
        channels_run(a)
        return
    if a.workload == "sweeps":
        sweeps_run(a)
        return
    if a.workload == "sync":
        sync_run(a)
        return
    if a.workload == "c1":
        c1_run(a)
        return
    if a.workload == "learned_channels_500":
        learned_500_line(a)
        return
    if a.workload.startswith("learned"):
        learned_line(a)
        return
    if a.workload == "net":
        net_line(a)
        return
    n = a.n
    ctx, ch = make_context(a)
    copy_gbs = ctx.stream_copy(1 << 30, 10)
    for _ in range(a.warmup):
        ctx.tick()
    # one grid: the launch-level tuner (no real ticks, so ticks 2-4 are timed)
    tuned = ctx.autotune(8)
    variant, ksweeps, tile_x, tile_y = ctx.solver_variant()
    ctx.synchronize()
    ctx.reset_stats()
    ctx.profile(a.sample)
    sweeps = 0
    t0 = time.perf_counter()
    for _ in range(a.steps):
        its, _ = ctx.tick()
        sweeps += its[ch]
    ctx.synchronize()
    t1 = time.perf_counter()
    ctx.profile(0)
    st = ctx.state(ch)
    elapsed = t1 - t0
    cells = n * n
    value = cells * a.steps / elapsed

    kname = "rb_tb"
    kernel_id = "k_rb_strip<%d, %s>" % (ksweeps, VARIANT_SHAPES.get(variant, "?"))
    traffic = None
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as fh:
            traffic = json.load(fh).get(kernel_id, {}).get("traffic_bytes_per_launch")
    except OSError:
        pass
    # batches of 8 to 64 launches queued back to back between host polls,
    # each timed by one pair of HIP events on the library's stream; the
    # per-launch event pairs ("rb_tb", every --sample-th launch) also time the
    # event packets around each launch and read a few us high
    ms, timed, launches = ctx.kernel_stats("rb_tb_batch")
    sms, stimed, _ = ctx.kernel_stats(kname)
    if not timed:
        ms, timed = sms, stimed
    _, _, launches = ctx.kernel_stats(kname)
    avg_s = ms / 1e3 / max(timed, 1)
    # one launch advances every cell by ksweeps red-black sweeps on chip: it
    # must move each cell's p, rhs and flags in and p out once (25 B per cell);
    # its tiles' overlapping halos load more (loaded / unique = halo factor)
    unique, loaded = ctx.solver_bytes()
    achieved = unique / avg_s / 1e9 if timed else None
    effective = SWEEP_BYTES_PER_CELL * ksweeps * cells / avg_s / 1e9 if timed else None
    working_set = (3 * 8 + 1) * (n + 2) * (n + 2)
    kernels = {}
    for k in ("rb_tb", "pbuf_fix", "solve_begin", "rhs_stats", "fg_rhs", "fg", "rhs",
              "adap_ranges", "restore", "bc_velocity", "delt"):
        kms, kt, kl = ctx.kernel_stats(k)
        if kl:
            kernels[k] = {"launches": kl, "timed": kt,
                          "avg_us": round(kms * 1e3 / kt, 3) if kt else None}
            if k in STAGE_BYTES_PER_CELL and kt:
                # single-launch event pairs: they also time the event packets
                kernels[k]["gbs_event_timed"] = round(
                    STAGE_BYTES_PER_CELL[k] * (n + 2) ** 2 / (kms * 1e-3 / kt) / 1e9, 1)
    sweeps_per_step = sweeps / a.steps
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": elapsed * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "lid-driven cavity %dx%d, Re=1000, tau=0.5, gamma=0.9, red-black SOR "
                        "omega=1.8 to relative residual %g, fp64, zero initial fields" % (n, n, a.eps),
            "solver": "rb",
            "variant": {"id": variant, "sweeps_per_launch": ksweeps, "tile": [tile_x, tile_y],
                        "autotune_us_per_sweep": tuned[1]},
            "sweeps_per_step": sweeps_per_step,
            "us_per_sweep": elapsed * 1e6 / max(sweeps, 1),
            "final_delt": st["delt"],
            "final_time": st["time"],
            "final_residual_ratio": (st["residual_squared"] / st["tolerance_squared"]
                                     if st["tolerance_squared"] else None),
        },
        "roofline": {
            "kernel": kernel_id,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS if achieved else None,
            "traffic": traffic,
            "traffic_source": ("FETCH_SIZE and WRITE_SIZE per working dispatch from two "
                               "rocprofv3 --pmc runs of `bench.py --profile --variant 9`, "
                               "scaled by a 256 MiB copy (profiles/traffic.json); not "
                               "measured in this run" if traffic else None),
            "bytes_per_launch": unique,
            "model": "25 B per cell per launch: p read, rhs read, flags read, p written",
            "loaded_bytes_per_launch": loaded,
            "halo_factor": loaded / unique,
            "effective_sweep_gbs": effective,
            "sweeps_per_launch": ksweeps,
            "avg_launch_us": avg_s * 1e6,
            "timed_launches": timed,
            "timing": "HIP events around queued batches of 8-64 launches",
            "single_launch_event_us": sms * 1e3 / stimed if stimed else None,
            "measured_copy_gbs": copy_gbs,
            # p, the second p buffer, rhs and flags: under the 256 MB Infinity Cache
            # on this grid, so the kernel's data may stay in the MALL between launches
            "working_set_bytes": working_set,
            "cache_resident": working_set < MALL_BYTES,
        },
        "kernels": kernels,
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, sweeps_per_step, a.cpu_sweeps)
        # beside the reference-order port: the red-black sweeps on all host threads
        out["cpu_baseline"]["parallel_red_black"] = cpu_parallel(n, sweeps_per_step,
                                                                 4 * a.cpu_sweeps)
    ctx.close()
    if not a.no_channels:
        try:
            out["channels"] = channels_summary(a)
        except Exception as e:   # the plain measurement stands on its own
            out["channels"] = {"error": "%s: %s" % (type(e).__name__, e)}
    if not a.no_learned:
        log = lambda *x: print("learned:", *x, file=sys.stderr, flush=True)  # noqa: E731
        try:
            trained = train_cavity_net(log)
            out["learned_same_grid"] = learned_same_grid(a, trained[0], value, sweeps_per_step,
                                                         log)
            out["learned"] = learned_cavities(10, log, trained)
        except Exception as e:   # the plain measurement stands on its own
            out["learned"] = {"error": "%s: %s" % (type(e).__name__, e)}
        try:
            out["learned_channels"] = learned_channels_summary(log)
        except Exception as e:
            out["learned_channels"] = {"error": "%s: %s" % (type(e).__name__, e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
