"""stm_render_static, the entry point that allocates nothing: the image
staged in the device's static buffer equals stm_render's bit for bit, on two
calls in a row (the second on a smaller grid, so the buffer holds the first
call's larger image underneath), and a channel larger than the buffer is
refused."""
import sys

import numpy as np
import pytest

from stroemung_amd import presets, workloads
from stroemung_amd._lib import StmError
from stroemung_amd.simulation import Context, make_params

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _running(ct, vel, prm, ticks):
    ctx = Context(0)
    ch = ctx.add_grid(prm, ct, vel)
    ctx.construct()
    for _ in range(ticks):
        ctx.tick()
    return ctx, ch


def test_render_static_equals_render_on_two_calls():
    say("test: stm_render_static vs stm_render")
    ct1, vel1 = presets.square_obstacle(300, 200)
    p1 = workloads.square_channel_params(solver="rb", eps=1e-6, itermax=20000)
    ct2, vel2, p2 = workloads.cavity(62, eps=1e-6)
    a, ca = _running(ct1, vel1, p1, 3)
    b, cb = _running(ct2, vel2, p2, 3)
    for ctx, ch, shape in ((a, ca, ct1.shape), (b, cb, ct2.shape)):   # larger, then smaller
        for mode in ("pressure", "speed"):
            got = ctx.render(ch, mode, static=True)
            want = ctx.render(ch, mode)
            assert got.shape == (shape[1], shape[0], 4)
            np.testing.assert_array_equal(got, want)
            if mode == "speed":                     # a real image (the pressure range of a
                assert len(np.unique(got.reshape(-1, 4), axis=0)) > 8   # solve that never hit
                                                    # itermax keeps its construction value)
    # a second round in the other order: the smaller image first
    for ctx, ch in ((b, cb), (a, ca)):
        np.testing.assert_array_equal(ctx.render(ch, "pressure", static=True),
                                      ctx.render(ch, "pressure"))
    a.close()
    b.close()


def test_render_static_refuses_a_channel_larger_than_its_buffer():
    say("test: stm_render_static size limit")
    n = 2048                       # 2050 x 2050 cells > 4 * 2^20 pixels
    ct, vel = presets.lid_cavity(n, n)
    prm = make_params(1.0 / n, 1.0 / n, 1e-4, 0.9, 100.0, 1e-6, 10, 1.7, solver="rb")
    ctx = Context(0)
    ch = ctx.add_grid(prm, ct, vel)
    ctx.upload(ch, p=np.zeros(ct.shape))
    with pytest.raises(StmError, match="static image buffer"):
        ctx.render(ch, "pressure", static=True)
    ctx.close()
    # not yet holding fields: refused, not allocated
    ctx = Context(0)
    ch = ctx.add_grid(prm, ct, vel)
    with pytest.raises(StmError, match="before rendering"):
        ctx.render(ch, "pressure", static=True)
    ctx.close()
