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
