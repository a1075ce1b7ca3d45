"""Python host mirroring the reference's `Simulation` / `SimulationGrid` API
(src/simulation.rs:49-333, src/grid/mod.rs:112-268) over the C ABI.

`Context` is the library's unit: a device stream with one or more grids
(channels) that advance together.  `Simulation` is one channel presented
with the reference's field names; its arrays live on the GPU and are read
back when accessed (the reference's fields are public, so a drop-in must
keep them readable between ticks).  Arrays written by the caller go back
with `upload()`.
"""
import ctypes as C

import numpy as np

from . import io as sio
from ._lib import SOLVER_LEX, SOLVER_RB, Params, State, check, lib

_D = C.POINTER(C.c_double)


def _dp(a):
    return None if a is None else a.ctypes.data_as(_D)


def _f64(a, shape):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.shape != tuple(shape):
        raise ValueError("expected shape %s, got %s" % (shape, a.shape))
    return a


SOLVERS = {"lex": SOLVER_LEX, "rb": SOLVER_RB, SOLVER_LEX: SOLVER_LEX, SOLVER_RB: SOLVER_RB}


def make_params(delx, dely, delt, gamma, reynolds, sor_absolute_epsilon, max_iterations, omega,
                initial_norm_squared=None, iterations=0, time=0.0, tau=0.0, solver="lex",
                check_every=1):
    p = Params()
    p.delx, p.dely, p.delt = delx, dely, delt
    p.gamma, p.reynolds = gamma, reynolds
    p.sor_absolute_epsilon = sor_absolute_epsilon
    p.max_iterations = int(max_iterations)
    p.omega = omega
    p.has_initial_norm_squared = int(initial_norm_squared is not None)
    p.initial_norm_squared = 0.0 if initial_norm_squared is None else initial_norm_squared
    p.iterations = int(iterations)
    p.time = time
    p.tau = tau
    p.solver = SOLVERS[solver]
    p.check_every = int(check_every)
    return p


def rebuild_boundary_list(cell_type):
    """SimulationGrid::rebuild_boundary_list (src/grid/mod.rs:202-235) on the
    host side of the library.  Returns (edge [NX,NY] int8 with -1 = fluid,
    sorted boundary list of (x, y), fluid cell count)."""
    ct = np.ascontiguousarray(cell_type, dtype=np.uint8)
    nx, ny = ct.shape
    edge = np.zeros((nx, ny), np.int8)
    lst = np.zeros(nx * ny, np.uint32)
    n = C.c_uint32(0)
    fluid = C.c_double(0)
    err = C.create_string_buffer(256)
    rc = lib().stm_rebuild_boundary_list(nx, ny, ct.ctypes.data_as(C.POINTER(C.c_uint8)),
                                         edge.ctypes.data_as(C.POINTER(C.c_int8)),
                                         lst.ctypes.data_as(C.POINTER(C.c_uint32)),
                                         C.byref(n), C.byref(fluid), err, 256)
    check(rc)
    return edge, [divmod(int(i), ny) for i in lst[:n.value]], fluid.value


class Context:
    """One device stream holding one or more grids."""

    def __init__(self, device=0):
        self._ctx = C.c_void_p()
        check(lib().stm_create(device, C.byref(self._ctx)))
        self.shapes = []

    def close(self):
        if self._ctx:
            lib().stm_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        check(rc, self._ctx)

    def add_grid(self, params, cell_type, cell_velocity=None):
        ct = np.ascontiguousarray(cell_type, dtype=np.uint8)
        nx, ny = ct.shape
        vel = None if cell_velocity is None else _f64(cell_velocity, (nx, ny, 2))
        ch = C.c_uint32(0)
        self._check(lib().stm_add_grid(self._ctx, C.byref(params), nx, ny,
                                       ct.ctypes.data_as(C.POINTER(C.c_uint8)), _dp(vel),
                                       C.byref(ch)))
        self.shapes.append((nx, ny))
        return ch.value

    def set_cells(self, ch, cell_type, cell_velocity=None):
        ct = np.ascontiguousarray(cell_type, dtype=np.uint8)
        vel = None if cell_velocity is None else _f64(cell_velocity, ct.shape + (2,))
        self._check(lib().stm_set_cells(self._ctx, ch, ct.ctypes.data_as(C.POINTER(C.c_uint8)),
                                        _dp(vel)))

    def set_params(self, ch, params):
        self._check(lib().stm_set_params(self._ctx, ch, C.byref(params)))

    def upload(self, ch, u=None, v=None, p=None, f=None, g=None, rhs=None):
        sh = self.shapes[ch]
        arrs = [None if a is None else _f64(a, sh) for a in (u, v, p, f, g, rhs)]
        self._check(lib().stm_upload(self._ctx, ch, *[_dp(a) for a in arrs]))

    def download(self, ch, fields=("u", "v", "p", "f", "g", "rhs")):
        sh = self.shapes[ch]
        out = {k: np.empty(sh) for k in fields}
        args = [_dp(out[k]) if k in out else None for k in ("u", "v", "p", "f", "g", "rhs")]
        self._check(lib().stm_download(self._ctx, ch, *args))
        return out

    def state(self, ch):
        s = State()
        self._check(lib().stm_get_state(self._ctx, ch, C.byref(s)))
        return {name: (list(getattr(s, name)) if isinstance(getattr(s, name), C.Array)
                       else getattr(s, name)) for name, _ in State._fields_}

    def construct(self):
        self._check(lib().stm_construct(self._ctx))

    def tick(self):
        n = len(self.shapes)
        its = (C.c_uint32 * n)()
        nrm = (C.c_double * n)()
        self._check(lib().stm_run_simulation_tick(self._ctx, its, nrm))
        return list(its), list(nrm)

    def run_ticks(self, n):
        self._check(lib().stm_run_ticks(self._ctx, n))

    def stage(self, name, *args):
        self._check(getattr(lib(), "stm_" + name)(self._ctx, *args))

    def synchronize(self):
        self._check(lib().stm_synchronize(self._ctx))

    def solver_variant(self, variant=-1):
        """Select (variant >= 0) or query the blocked red-black variant:
        returns (variant, sweeps per launch, tile_x, tile_y)."""
        out = [C.c_int32() for _ in range(4)]
        self._check(lib().stm_solver_variant(self._ctx, variant, *[C.byref(o) for o in out]))
        return tuple(o.value for o in out)

    def solver_bytes(self):
        """(unique, loaded) bytes of one blocked launch of the current
        variant: every cell once (25 B) and what the overlapping tiles move."""
        u, l = C.c_double(), C.c_double()
        self._check(lib().stm_solver_bytes(self._ctx, C.byref(u), C.byref(l)))
        return u.value, l.value

    def stream_copy(self, nbytes, reps=20):
        """Measured device copy bandwidth, GB/s (read + write bytes)."""
        g = C.c_double()
        self._check(lib().stm_stream_copy(self._ctx, nbytes, reps, C.byref(g)))
        return g.value

    def sync_benchmark(self, steps=1000):
        """Microseconds per step of the blocked kernel's launch shape: separate
        launches, graph replay, a resident kernel with a grid-wide barrier or
        with neighbour flags, each with and without its tile write, and a
        10-us kernel launched and replayed (keys below)."""
        us = (C.c_double * 8)()
        self._check(lib().stm_sync_benchmark(self._ctx, steps, us))
        keys = ("launch", "graph", "grid_barrier", "neighbour_flags", "grid_barrier_no_write",
                "neighbour_flags_no_write", "launch_10us", "graph_10us")
        return dict(zip(keys, list(us)))

    def autotune(self, launches=8):
        """Times `launches` blocked launches of each variant on the current
        fields (restored afterwards); keeps the fastest per sweep.  Returns
        (variant, us per sweep)."""
        chosen = C.c_int32()
        us = C.c_double()
        self._check(lib().stm_autotune(self._ctx, launches, C.byref(chosen), C.byref(us)))
        return chosen.value, us.value

    def autotune_ticks(self):
        """One real tick per variant (the simulation advances), scored by
        interior cell-sweeps per second of the whole tick; keeps the best.
        Returns (variant, cell-sweeps/s)."""
        chosen = C.c_int32()
        rate = C.c_double()
        self._check(lib().stm_autotune_ticks(self._ctx, C.byref(chosen), C.byref(rate)))
        return chosen.value, rate.value

    def profile(self, sample):
        self._check(lib().stm_profile(self._ctx, sample))

    def kernel_stats(self, name):
        ms = C.c_double(0)
        timed = C.c_uint64(0)
        launches = C.c_uint64(0)
        self._check(lib().stm_kernel_stats(self._ctx, name.encode(), C.byref(ms),
                                           C.byref(timed), C.byref(launches)))
        return ms.value, timed.value, launches.value

    def reset_stats(self):
        self._check(lib().stm_reset_stats(self._ctx))

    def set_learned(self, net, cycles=1, sweeps=32, step="network", coarsen=1):
        """Two-grid cycles with the network before every red-black solve (net
        None: off).  step: "network" takes the network's magnitude, "energy"
        a line search, "momentum" the energy minimum over the correction and
        the previous cycle's change of p, "momentum2" also over the change of
        the cycle before; coarsen: 1, 2, 4, 8 or 16."""
        rule = {"energy": 0, "network": 1, "momentum": 2, "momentum2": 3}[step]
        self._check(lib().stm_set_learned(self._ctx, net._net if net else None, cycles, sweeps,
                                          rule, coarsen))

    def render(self, ch, mode="pressure", static=False):
        """The reference's GUI image of channel ch: (ny, nx, 4) uint8 RGBA.
        static=True stages it in the device's static buffer (no allocation)."""
        nx, ny = self.shapes[ch]
        out = np.empty((ny, nx, 4), np.uint8)
        fn = lib().stm_render_static if static else lib().stm_render
        self._check(fn(self._ctx, ch, {"pressure": 0, "speed": 1}[mode],
                       out.ctypes.data_as(C.POINTER(C.c_uint8))))
        return out

    def set_graphs(self, on=True):
        """Replay learned cycles from a captured hipGraph."""
        self._check(lib().stm_set_graphs(self._ctx, 1 if on else 0))

    def set_capture(self, net, points=2, sweeps=32, coarsen=1):
        """Store `points` training pairs per channel from every solve in net's
        training set (None: off)."""
        self._check(lib().stm_set_capture(self._ctx, net._net if net else None, points, sweeps,
                                          coarsen))


class Simulation:
    """The reference's `Simulation` on one GPU channel.

    Construction mirrors `Simulation::try_from` (src/simulation.rs:71-99):
    the boundary is classified, f = g = rhs = 0, F/G, RHS and, unless given,
    initial_norm_squared are computed.  `run_simulation_tick()` returns
    (sor_iterations, norm_squared) like src/simulation.rs:324.
    """

    def __init__(self, size, cell_size, delt, gamma, reynolds, sor_absolute_epsilon,
                 max_iterations, omega, cell_type, cell_velocity=None, u=None, v=None,
                 pressure=None, initial_norm_squared=None, iterations=0, time=0.0, tau=0.0,
                 solver="lex", check_every=1, device=0, context=None):
        self.size = tuple(size)
        self.cell_size = tuple(cell_size)
        self.gamma, self.reynolds, self.omega = gamma, reynolds, omega
        self.sor_absolute_epsilon = sor_absolute_epsilon
        self.max_iterations = max_iterations
        self.tau, self.solver, self.check_every = tau, solver, check_every
        self.cell_type = np.ascontiguousarray(cell_type, dtype=np.uint8)
        self.cell_velocity = (np.zeros(self.size + (2,)) if cell_velocity is None
                              else _f64(cell_velocity, self.size + (2,)))
        self._params = make_params(cell_size[0], cell_size[1], delt, gamma, reynolds,
                                   sor_absolute_epsilon, max_iterations, omega,
                                   initial_norm_squared, iterations, time, tau, solver,
                                   check_every)
        self.ctx = context or Context(device)
        self.channel = self.ctx.add_grid(self._params, self.cell_type, self.cell_velocity)
        z = np.zeros(self.size)
        self.ctx.upload(self.channel, u=z if u is None else u, v=z if v is None else v,
                        p=z if pressure is None else pressure)
        self.ctx.construct()
        self._cache = None
        self._state = None

    # --------------------------------------------------------------- I/O
    @classmethod
    def from_reader(cls, fh, exact=False, **ext):
        """Simulation::from_reader (src/simulation.rs:117-120): serde JSON.

        Floats are parsed as the reference's serde_json parses them, which
        can move a 17-digit decimal by one ulp; exact=True parses them
        correctly rounded instead, so a state written by to_json() resumes
        bit for bit (f, g and rhs are not in the JSON: construction
        recomputes them, and a tick recomputes every value of them it
        reads)."""
        d = sio.load_reference_json(fh, exact=exact)
        g = d["grid"]
        ct, vel = sio.decode_cells(g["cell_type"])
        return cls(size=d["size"], cell_size=d["cell_size"], delt=d["delt"], gamma=d["gamma"],
                   reynolds=d["reynolds"], sor_absolute_epsilon=d["sor_absolute_epsilon"],
                   max_iterations=d["max_iterations"], omega=d["omega"], cell_type=ct,
                   cell_velocity=vel, u=sio.decode_array(g["u"]), v=sio.decode_array(g["v"]),
                   pressure=sio.decode_array(g["pressure"]),
                   initial_norm_squared=d.get("initial_norm_squared"),
                   iterations=d["iterations"], time=d["time"], **ext)

    def to_json(self):
        """The reference's serialised form (f, g, rhs are #[serde(skip)])."""
        st = self.state
        return {
            "size": list(self.size), "cell_size": list(self.cell_size), "delt": st["delt"],
            "gamma": self.gamma, "reynolds": self.reynolds,
            "initial_norm_squared": (st["initial_norm_squared"]
                                     if st["has_initial_norm_squared"] else None),
            "sor_absolute_epsilon": self.sor_absolute_epsilon,
            "max_iterations": self.max_iterations, "iterations": st["iterations"],
            "time": st["time"], "omega": self.omega,
            "grid": {"size": list(self.size), "pressure": sio.encode_array(self.pressure),
                     "u": sio.encode_array(self.u), "v": sio.encode_array(self.v),
                     "cell_type": sio.encode_cells(self.cell_type, self.cell_velocity)},
        }

    # ------------------------------------------------------------ stepping
    def run_simulation_tick(self):
        its, nrm = self.ctx.tick()
        self._cache = None
        self._state = None
        return its[self.channel], nrm[self.channel]

    def _fields(self):
        if self._cache is None:
            self._cache = self.ctx.download(self.channel)
        return self._cache

    def upload(self, **fields):
        """Send caller-modified arrays (u, v, p, f, g, rhs) to the device."""
        self.ctx.upload(self.channel, **fields)
        self._cache = None

    def set_cell_type(self, cell_type, cell_velocity=None):
        """Edit the grid and rebuild the boundary list, as draw_cells does
        (src/lib.rs:38-78): u, v and p of every cell whose type changes
        become 0; a too-thin boundary raises InvalidGridError and leaves the
        grid and the fields unchanged."""
        self.ctx.set_cells(self.channel, cell_type, cell_velocity)
        self._cache = None
        self.cell_type = np.ascontiguousarray(cell_type, dtype=np.uint8)
        if cell_velocity is not None:
            self.cell_velocity = _f64(cell_velocity, self.size + (2,))

    @property
    def state(self):
        if self._state is None:
            self._state = self.ctx.state(self.channel)
        return self._state

    u = property(lambda self: self._fields()["u"])
    v = property(lambda self: self._fields()["v"])
    pressure = property(lambda self: self._fields()["p"])
    f = property(lambda self: self._fields()["f"])
    g = property(lambda self: self._fields()["g"])
    rhs = property(lambda self: self._fields()["rhs"])
    delt = property(lambda self: self.state["delt"])
    time = property(lambda self: self.state["time"])
    iterations = property(lambda self: self.state["iterations"])
    initial_norm_squared = property(lambda self: self.state["initial_norm_squared"])
    pressure_range = property(lambda self: self.state["pressure_range"])
    speed_range = property(lambda self: self.state["speed_range"])
