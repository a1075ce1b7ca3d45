"""The pressure network and the learned pressure solve (host side).

The network is a 4-level U-Net (32 base channels, 3x3 convolutions, ReLU)
whose forward, backward and Adam kernels are the library's own (BF16 on the
matrix cores, FP32 accumulation and master weights).  Its input is the
normalised residual of the pressure equation and the interior-fluid mask of a
grid; its output is a correction of p.  Inside a solve the correction is
scaled by the step that minimises the residual norm, and the FP64 red-black
solver always finishes to the same eps, so the network changes how fast a
solve converges, never where it ends.

Training pairs come from the solver itself: `capture` makes every solve of a
context store the residual at its start (and after a few sweeps) with the
normalised change of p the solve then made.
"""
import ctypes as C
import time

import numpy as np

from ._lib import NetStats, StmError, check, lib


class PressureNet:
    """stm_net_*: a network on the device and stream of a Context."""

    def __init__(self, ctx, height, width, batch, seed=0):
        self.ctx = ctx
        self.height, self.width, self.batch = int(height), int(width), int(batch)
        self._net = C.c_void_p()
        rc = lib().stm_net_create(ctx._ctx, self.height, self.width, self.batch, seed,
                                  C.byref(self._net))
        if rc:   # no net exists: the message is the calling thread's last one
            raise StmError(rc, lib().stm_net_last_error(None).decode(errors="replace"))
        n = C.c_uint64()
        self._check(lib().stm_net_param_count(self._net, C.byref(n)))
        self.nparam = int(n.value)

    def _check(self, rc):
        check(rc, net=self._net)

    def close(self):
        if self._net:
            lib().stm_net_destroy(self._net)
            self._net = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- parameters
    def get_params(self):
        out = np.empty(self.nparam, np.float32)
        self._check(lib().stm_net_get_params(self._net, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def set_params(self, params):
        a = np.ascontiguousarray(params, np.float32)
        if a.size != self.nparam:
            raise ValueError("expected %d parameters, got %d" % (self.nparam, a.size))
        self._check(lib().stm_net_set_params(self._net, a.ctypes.data_as(C.POINTER(C.c_float))))

    # -- compute on host batches
    def forward(self, x):
        """x: (nb, H, W, 2) float32 -> (nb, H, W) float32."""
        x = np.ascontiguousarray(x, np.float32)
        nb = x.shape[0]
        assert x.shape == (nb, self.height, self.width, 2)
        y = np.empty((nb, self.height, self.width), np.float32)
        self._check(lib().stm_net_forward(self._net, nb, x.ctypes.data_as(C.POINTER(C.c_float)),
                                          y.ctypes.data_as(C.POINTER(C.c_float))))
        return y

    def gradients(self, x, t):
        """Loss and parameter gradients of the masked relative squared error."""
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        nb = x.shape[0]
        assert x.shape == (nb, self.height, self.width, 2)
        assert t.shape == (nb, self.height, self.width)
        loss = C.c_float()
        g = np.empty(self.nparam, np.float32)
        fp = C.POINTER(C.c_float)
        self._check(lib().stm_net_gradients(self._net, nb, x.ctypes.data_as(fp),
                                            t.ctypes.data_as(fp), C.byref(loss),
                                            g.ctypes.data_as(fp)))
        return float(loss.value), g

    # -- training set and training
    def reserve(self, capacity):
        self._check(lib().stm_net_dataset_reserve(self._net, capacity))

    def dataset_size(self):
        n = C.c_uint32()
        self._check(lib().stm_net_dataset_size(self._net, C.byref(n)))
        return int(n.value)

    def sample(self, index):
        """(x (H, W, 2), t (H, W)) of training sample `index`."""
        x = np.empty((self.height, self.width, 2), np.float32)
        t = np.empty((self.height, self.width), np.float32)
        fp = C.POINTER(C.c_float)
        self._check(lib().stm_net_dataset_get(self._net, index, x.ctypes.data_as(fp),
                                              t.ctypes.data_as(fp)))
        return x, t

    def train(self, steps, batch, lr=1e-3, seed=0):
        loss = C.c_float()
        self._check(lib().stm_net_train(self._net, steps, batch, lr, seed, C.byref(loss)))
        return float(loss.value)

    def stats(self):
        s = NetStats()
        self._check(lib().stm_net_stats(self._net, C.byref(s)))
        return {k: getattr(s, k) for k, _ in NetStats._fields_ if k != "pad"}

    def benchmark(self, nb, reps=5, train=True):
        f, t = C.c_double(), C.c_double()
        self._check(lib().stm_net_benchmark(self._net, nb, reps, C.byref(f),
                                            C.byref(t) if train else None))
        return f.value, (t.value if train else None)

    # -- parameter layout (the header's description, for tests and tools)
    def layout(self):
        """[(name, offset, shape)] of every parameter block."""
        plan = [(32, 32), (32, 32), (32, 64), (64, 64), (64, 128), (128, 128), (128, 256),
                (256, 256), (384, 128), (128, 128), (192, 64), (64, 64), (96, 32), (32, 32)]
        out, off = [], 0
        for k, (cin, cout) in enumerate(plan):
            out.append(("conv%d.w" % k, off, (cout, 9, cin)))
            off += cout * 9 * cin
            out.append(("conv%d.b" % k, off, (cout,)))
            off += cout
            off = (off + 7) // 8 * 8
        out.append(("head.w", off, (32,)))
        off += 32
        out.append(("head.b", off, (1,)))
        off += 8
        assert off == self.nparam, (off, self.nparam)
        return out


def round8(n):
    return (int(n) + 7) // 8 * 8


def frame_for(grids, coarsen):
    """Smallest network frame (multiples of 8) holding every grid's interior
    coarsened by `coarsen`."""
    h = max(-(-(ct.shape[0] - 2) // coarsen) for ct, _, _ in grids)
    w = max(-(-(ct.shape[1] - 2) // coarsen) for ct, _, _ in grids)
    return round8(h), round8(w)


def make_context(grids, device=0):
    from .simulation import Context
    ctx = Context(device)
    for ct, vel, prm in grids:
        ctx.add_grid(prm, ct, vel)
    ctx.construct()
    return ctx


def _quiet(*_):
    pass


def train_net(grids, coarsen=4, capture_ticks=96, points=4, capture_sweeps=32, steps=3000,
              batch=32, lr=1e-3, dagger_rounds=2, dagger_ticks=32, dagger_steps=1500,
              cycles=16, sweeps=32, step="energy", seed=1, log=_quiet):
    """Train a pressure network on pairs the solver produces on `grids`:
    `capture_ticks` plain ticks storing `points` pairs per channel, `steps`
    Adam steps, then `dagger_rounds` rounds of learned ticks that store pairs
    from the learned solver's own states, each followed by `dagger_steps`
    more steps.  Returns (params, info)."""
    H, W = frame_for(grids, coarsen)
    n = len(grids)
    ctx = make_context(grids)
    net = PressureNet(ctx, H, W, max(n, batch), seed=seed)
    t0 = time.perf_counter()
    try:
        net.reserve(points * n * (capture_ticks + dagger_rounds * dagger_ticks))
        ctx.tick()
        ctx.autotune(8)
        ctx.set_capture(net, points=points, sweeps=capture_sweeps, coarsen=coarsen)
        plain_sweeps = []
        for i in range(capture_ticks):
            its, _ = ctx.tick()
            plain_sweeps.append(float(np.mean(its)))
        ctx.set_capture(None)
        log("captured %d pairs from %d plain ticks" % (net.dataset_size(), capture_ticks))

        def fit(nsteps, rate):
            done, loss = 0, None
            while done < nsteps:
                k = min(500, nsteps - done)
                loss = net.train(k, batch, lr=rate if done < 0.7 * nsteps else rate * 0.3,
                                 seed=seed * 1000003 + done + 7919 * nsteps)
                done += k
            return loss

        loss = fit(steps, lr)
        log("trained %d steps, loss %.4f" % (steps, loss))
        rounds = []
        for r in range(dagger_rounds):
            ctx.set_learned(net, cycles=cycles, sweeps=sweeps, step=step, coarsen=coarsen)
            ctx.set_capture(net, points=points, sweeps=capture_sweeps, coarsen=coarsen)
            sw = []
            for _ in range(dagger_ticks):
                its, _ = ctx.tick()
                sw.append(float(np.mean(its)))
            ctx.set_capture(None)
            ctx.set_learned(None)
            loss = fit(dagger_steps, lr * 0.5)
            rounds.append({"learned_sweeps_per_solve": float(np.mean(sw)), "loss": loss})
            log("round %d: learned sweeps/solve %.1f, loss %.4f" % (r, np.mean(sw), loss))
        st = net.stats()
        info = {
            "frame": [H, W], "coarsen": coarsen, "params": net.nparam,
            "pairs": net.dataset_size(), "capture_ticks": capture_ticks, "points": points,
            "plain_sweeps_per_solve_during_capture": float(np.mean(plain_sweeps)),
            "train_steps": int(st["train_steps"]), "batch": batch, "final_loss": loss,
            "dagger_rounds": rounds,
            "train_ms": st["train_ms"],
            "train_tflops": (3.0 * st["flops_forward_per_sample"] * st["train_samples"]
                             / (st["train_ms"] * 1e-3) / 1e12) if st["train_ms"] else None,
            "wall_s": time.perf_counter() - t0,
        }
        return net.get_params(), info
    finally:
        ctx.set_learned(None)
        ctx.set_capture(None)
        net.close()
        ctx.close()


def _run_ticks(ctx, ticks, cells):
    ctx.synchronize()
    sweeps = 0.0
    t = time.perf_counter()
    for _ in range(ticks):
        its, _ = ctx.tick()
        sweeps += float(np.sum(its))
    ctx.synchronize()
    el = time.perf_counter() - t
    bound = True
    for c in range(len(cells)):
        s = ctx.state(c)
        bound = bound and s["residual_squared"] <= s["tolerance_squared"] and not s["exhausted"]
    return el, sweeps, bound


def compare(grids, params, coarsen=4, cycles=16, sweeps=32, warm=20, ticks=10, graphs=True,
            step="energy", tune="launches", log=_quiet):
    """Plain red-black solve vs the learned solve on `grids`, each from the
    same state after `warm` plain ticks and the same tuning (two contexts,
    same deterministic warm-up; tune = "ticks" uses the whole-tick tuner,
    whose ticks are real, in both).  Every channel of both runs ends each
    solve at its residual bound.  Returns a dict of throughputs."""
    cells = [(ct.shape[0] - 2) * (ct.shape[1] - 2) for ct, _, _ in grids]
    H, W = frame_for(grids, coarsen)
    out = {"channels": len(grids), "interior_cells": int(sum(cells)), "ticks": ticks,
           "warm": warm}
    def tuned(ctx):
        if tune == "ticks":
            ctx.autotune_ticks()
        else:
            ctx.autotune(8)

    ctx = make_context(grids)
    for _ in range(warm):
        ctx.tick()
    tuned(ctx)
    el, sw, ok = _run_ticks(ctx, ticks, cells)
    variant = ctx.solver_variant()
    ctx.close()
    out["plain"] = {"cell_updates_per_s": sum(cells) * ticks / el, "ms_per_tick": el * 1e3 / ticks,
                    "sweeps_per_solve": sw / ticks / len(grids), "residual_bound_met": ok,
                    "variant": {"id": variant[0], "sweeps_per_launch": variant[1]}}
    log("plain: %.2f ms/tick, %.1f sweeps/solve" % (el * 1e3 / ticks, sw / ticks / len(grids)))
    ctx = make_context(grids)
    net = PressureNet(ctx, H, W, len(grids), seed=1)
    try:
        net.set_params(params)
        for _ in range(warm):
            ctx.tick()
        tuned(ctx)
        ctx.set_learned(net, cycles=cycles, sweeps=sweeps, step=step, coarsen=coarsen)
        ctx.set_graphs(graphs)
        el, sw, ok = _run_ticks(ctx, ticks, cells)
        st = net.stats()
        ctx.set_learned(None)
    finally:
        net.close()
        ctx.close()
    fwd_flops = st["flops_forward_per_sample"] * st["forward_samples"]
    out["learned"] = {
        "cell_updates_per_s": sum(cells) * ticks / el, "ms_per_tick": el * 1e3 / ticks,
        "sweeps_per_solve": sw / ticks / len(grids), "residual_bound_met": ok,
        "cycles": cycles, "sweeps_per_cycle": sweeps, "step": step, "coarsen": coarsen,
        "frame": [H, W],
        "graphs": graphs,
        # forward passes timed with events (with graphs: the first cycle of each solve)
        "forward_ms_per_call": st["forward_ms"] / st["forward_calls"] if st["forward_calls"] else None,
        "forward_tflops": fwd_flops / (st["forward_ms"] * 1e-3) / 1e12 if st["forward_ms"] else None,
    }
    out["speedup"] = out["learned"]["cell_updates_per_s"] / out["plain"]["cell_updates_per_s"]
    out["sweeps_saved"] = 1.0 - out["learned"]["sweeps_per_solve"] / out["plain"]["sweeps_per_solve"]
    log("learned: %.2f ms/tick, %.1f sweeps/solve, %.2fx" % (el * 1e3 / ticks,
                                                             sw / ticks / len(grids),
                                                             out["speedup"]))
    return out


def trajectory(grids, params, coarsen=8, cycles=64, sweeps=32, step="momentum", ticks=500,
               checkpoints=(10, 50, 100, 200, 300, 400, 500), log=_quiet):
    """A long learned run: `ticks` learned ticks from rest (after one plain
    tick and the autotuner), each timed; after the learned tick numbered in
    `checkpoints`, one plain tick on the same state is timed as well (it
    advances the simulation like any tick).  Every learned tick's channels
    end at their residual bound.  Returns the learned totals and, per
    checkpoint, learned (mean of the last 10 ticks) vs plain."""
    cells = [(ct.shape[0] - 2) * (ct.shape[1] - 2) for ct, _, _ in grids]
    H, W = frame_for(grids, coarsen)
    ctx = make_context(grids)
    net = PressureNet(ctx, H, W, len(grids), seed=1)
    marks = []
    try:
        net.set_params(params)
        ctx.tick()
        ctx.autotune(8)
        ctx.set_graphs(True)
        ctx.set_learned(net, cycles=cycles, sweeps=sweeps, step=step, coarsen=coarsen)
        ctx.synchronize()
        ms, sw, bound = [], [], True
        for k in range(1, ticks + 1):
            t = time.perf_counter()
            its, _ = ctx.tick()
            ms.append((time.perf_counter() - t) * 1e3)
            sw.append(float(np.mean(its)))
            for c in range(len(grids)):
                st = ctx.state(c)
                bound = bound and st["residual_squared"] <= st["tolerance_squared"] \
                    and not st["exhausted"]
            if k in checkpoints:
                ctx.set_learned(None)
                t = time.perf_counter()
                pits, _ = ctx.tick()
                pm = (time.perf_counter() - t) * 1e3
                ctx.set_learned(net, cycles=cycles, sweeps=sweeps, step=step, coarsen=coarsen)
                lm = float(np.mean(ms[-10:]))
                marks.append({"tick": k, "time": ctx.state(0)["time"],
                              "learned_ms_per_tick": lm,
                              "learned_sweeps_per_solve": float(np.mean(sw[-10:])),
                              "plain_ms_per_tick": pm,
                              "plain_sweeps_per_solve": float(np.mean(pits)),
                              "plain_over_learned": pm / lm})
                log("tick %d: learned %.1f ms (%.0f sweeps), plain %.0f ms (%.0f sweeps), %.1fx"
                    % (k, lm, marks[-1]["learned_sweeps_per_solve"], pm,
                       marks[-1]["plain_sweeps_per_solve"], pm / lm))
        ctx.set_learned(None)
    finally:
        net.close()
        ctx.close()
    wall = sum(ms) * 1e-3
    return {"ticks": ticks, "channels": len(grids), "interior_cells": int(sum(cells)),
            "learned": {"wall_s": wall, "ms_per_tick": wall * 1e3 / ticks,
                        "cell_updates_per_s": sum(cells) * ticks / wall,
                        "sweeps_per_solve": float(np.mean(sw)), "residual_bound_met": bound,
                        "cycles": cycles, "sweeps_per_cycle": sweeps, "step": step,
                        "coarsen": coarsen, "frame": [H, W], "graphs": True},
            "checkpoints": marks}

