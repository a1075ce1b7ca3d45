"""The launch and wait measurement behind the blocked solver's relaunch design
(stm_sync_benchmark): every figure is a finite, positive time per step, every
bounded neighbour wait ended on its flag, and bad arguments are refused."""
import ctypes as C
import math
import sys

import pytest

from stroemung_amd import _lib
from stroemung_amd.simulation import Context

pytestmark = pytest.mark.gpu


def test_sync_benchmark_reports_every_step_form():
    print("stm_sync_benchmark, 50 steps", flush=True)
    sys.stdout.flush()
    ctx = Context(0)
    try:
        us = ctx.sync_benchmark(50)
        assert len(us) == 8
        for k, v in us.items():
            assert math.isfinite(v) and v > 0.0, (k, v)
        # a kernel that holds every workgroup 10 us cannot take less per launch
        assert us["launch_10us"] >= 10.0 and us["graph_10us"] >= 10.0
        out = (C.c_double * 8)()
        assert _lib.lib().stm_sync_benchmark(ctx._ctx, 5, out) == -6   # STM_ERR_INVALID_ARG
    finally:
        ctx.close()
