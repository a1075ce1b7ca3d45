"""ctypes wrapper of the CPU oracle (oracle.c) — TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg,
as the checker.  The product package never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")

FLUID, NOSLIP, INFLOW, OUTFLOW, MOVING_WALL = 0, 1, 2, 3, 4
SOLVER_LEX, SOLVER_RB = 0, 1


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


class Restore(C.Structure):
    _fields_ = [("idx", C.c_int32), ("has_u", C.c_int32), ("has_v", C.c_int32),
                ("u", C.c_double), ("v", C.c_double)]


class OrcSim(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32),
        ("delx", C.c_double), ("dely", C.c_double), ("delt", C.c_double),
        ("gamma", C.c_double), ("reynolds", C.c_double),
        ("sor_absolute_epsilon", C.c_double),
        ("max_iterations", C.c_int32),
        ("omega", C.c_double),
        ("has_initial_norm_squared", C.c_int32),
        ("initial_norm_squared", C.c_double),
        ("iterations", C.c_uint32),
        ("time", C.c_double),
        ("tau", C.c_double),
        ("solver", C.c_int32),
        ("check_every", C.c_int32),
        ("masked_stop", C.c_int32),
        ("fix_mean", C.c_int32),
        ("rhs_mean", C.c_double),
        ("u", C.POINTER(C.c_double)), ("v", C.POINTER(C.c_double)),
        ("p", C.POINTER(C.c_double)), ("f", C.POINTER(C.c_double)),
        ("g", C.POINTER(C.c_double)), ("rhs", C.POINTER(C.c_double)),
        ("cell_type", C.POINTER(C.c_uint8)),
        ("cell_vel", C.POINTER(C.c_double)),
        ("edge", C.POINTER(C.c_int8)),
        ("blist", C.POINTER(C.c_int32)),
        ("nb", C.c_int32),
        ("fluid_cells", C.c_double),
        ("restore", C.POINTER(Restore)),
        ("nrestore", C.c_int32),
        ("pressure_range", C.c_double * 2),
        ("speed_range", C.c_double * 2),
        ("last_res2", C.c_double),
        ("last_tol2", C.c_double),
    ]


_D = C.POINTER(C.c_double)


def _declare(L):
    for name in ("orc_du2dx", "orc_dv2dy"):
        getattr(L, name).argtypes = [_D, C.c_double, C.c_double]
        getattr(L, name).restype = C.c_double
    for name in ("orc_duvdx", "orc_duvdy"):
        getattr(L, name).argtypes = [_D, _D, C.c_double, C.c_double]
        getattr(L, name).restype = C.c_double
    L.orc_laplacian.argtypes = [_D, C.c_double, C.c_double]
    L.orc_laplacian.restype = C.c_double
    L.orc_residual.argtypes = [_D, C.c_double, C.c_double, C.c_double]
    L.orc_residual.restype = C.c_double
    for name in ("orc_calculate_f", "orc_calculate_g"):
        getattr(L, name).argtypes = [_D, _D] + [C.c_double] * 5
        getattr(L, name).restype = C.c_double
    L.orc_classify.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_uint8), C.POINTER(C.c_int8),
                               C.POINTER(C.c_int32), C.POINTER(C.c_int32), _D]
    L.orc_classify.restype = C.c_int
    P = C.POINTER(OrcSim)
    for name in ("orc_rebuild_boundary_list", "orc_copy_pressure_to_boundaries",
                 "orc_set_boundary_u_and_v", "orc_construct"):
        getattr(L, name).argtypes = [P]
        getattr(L, name).restype = C.c_int
    for name in ("orc_calculate_f_and_g", "orc_calculate_rhs", "orc_set_u_and_v",
                 "orc_calculate_pressure_range", "orc_calculate_speed_range"):
        getattr(L, name).argtypes = [P]
        getattr(L, name).restype = None
    L.orc_calculate_norm_squared.argtypes = [P]
    L.orc_calculate_norm_squared.restype = C.c_double
    L.orc_get_initial_norm_squared.argtypes = [P]
    L.orc_get_initial_norm_squared.restype = C.c_double
    for name in ("orc_solve_sor", "orc_solve_rb", "orc_run_simulation_tick"):
        getattr(L, name).argtypes = [P, C.POINTER(C.c_uint32), _D]
        getattr(L, name).restype = C.c_int
    L.orc_max_uv.argtypes = [P, _D, _D]
    L.orc_max_uv.restype = None
    L.orc_comp_delt.argtypes = [P, C.c_double, C.c_double]
    L.orc_comp_delt.restype = C.c_double
    L.orc_ghost_refresh.argtypes = [P, C.c_int]
    L.orc_rb_half_sweep.argtypes = [P, C.c_int]
    L.orc_rb_sweeps.argtypes = [P, C.c_int]
    L.orc_rb_sweeps_mt.argtypes = [P, C.c_int, C.c_int]
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_set_threads.restype = None
    L.orc_masked_residual_sum.argtypes = [P, _D]
    L.orc_masked_residual_sum.restype = C.c_double
    L.orc_rhs_mean.argtypes = [P]
    L.orc_rhs_mean.restype = C.c_double
    L.orc_render.argtypes = [P, C.c_int, C.POINTER(C.c_uint8)]
    L.orc_render.restype = None


def _dp(a):
    return a.ctypes.data_as(_D)


def set_threads(n=None):
    """Threads for the red-black solves (n None: every core); the results
    are the same bits for any count."""
    lib().orc_set_threads(int(n or os.cpu_count() or 1))


def view3(a):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(9))
    return a


def classify(cell_type):
    """-> (edge int8 [NX,NY] with -1 for fluid, sorted boundary flat indices, fluid count)."""
    ct = np.ascontiguousarray(cell_type, dtype=np.uint8)
    nx, ny = ct.shape
    edge = np.zeros((nx, ny), dtype=np.int8)
    blist = np.zeros(nx * ny, dtype=np.int32)
    nb = C.c_int32(0)
    fc = C.c_double(0)
    rc = lib().orc_classify(nx, ny, ct.ctypes.data_as(C.POINTER(C.c_uint8)),
                            edge.ctypes.data_as(C.POINTER(C.c_int8)),
                            blist.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(nb), C.byref(fc))
    if rc != 0:
        idx = -rc - 1
        raise ValueError("BoundaryTooThinError at (%d, %d)" % divmod(idx, ny))
    return edge, blist[:nb.value].copy(), fc.value


class Sim:
    """One oracle simulation.  Fields are numpy arrays [NX, NY] owned here."""

    def __init__(self, cell_type, cell_vel=None, *, delx, dely, delt, gamma, reynolds,
                 sor_absolute_epsilon, max_iterations, omega, initial_norm_squared=None,
                 iterations=0, time=0.0, u=None, v=None, p=None, tau=0.0,
                 solver=SOLVER_LEX, check_every=1, masked_stop=0):
        self.cell_type = np.ascontiguousarray(cell_type, dtype=np.uint8)
        nx, ny = self.cell_type.shape
        if cell_vel is None:
            cell_vel = np.zeros((nx, ny, 2))
        self.cell_vel = np.ascontiguousarray(cell_vel, dtype=np.float64)
        z = lambda a: np.zeros((nx, ny)) if a is None else np.array(a, dtype=np.float64)  # noqa
        self.u, self.v, self.p = z(u), z(v), z(p)
        self.f, self.g, self.rhs = z(None), z(None), z(None)
        self.edge = np.zeros((nx, ny), dtype=np.int8)
        self.blist = np.zeros(nx * ny, dtype=np.int32)
        self.restore = (Restore * (2 * nx * ny))()
        s = OrcSim()
        s.nx, s.ny = nx, ny
        s.delx, s.dely, s.delt, s.gamma, s.reynolds = delx, dely, delt, gamma, reynolds
        s.sor_absolute_epsilon = sor_absolute_epsilon
        s.max_iterations = max_iterations
        s.omega = omega
        s.has_initial_norm_squared = int(initial_norm_squared is not None)
        s.initial_norm_squared = initial_norm_squared or 0.0
        s.iterations = iterations
        s.time = time
        s.tau = tau
        s.solver = solver
        s.check_every = check_every
        s.masked_stop = masked_stop
        s.u, s.v, s.p = _dp(self.u), _dp(self.v), _dp(self.p)
        s.f, s.g, s.rhs = _dp(self.f), _dp(self.g), _dp(self.rhs)
        s.cell_type = self.cell_type.ctypes.data_as(C.POINTER(C.c_uint8))
        s.cell_vel = _dp(self.cell_vel)
        s.edge = self.edge.ctypes.data_as(C.POINTER(C.c_int8))
        s.blist = self.blist.ctypes.data_as(C.POINTER(C.c_int32))
        s.restore = C.cast(self.restore, C.POINTER(Restore))
        self.s = s

    def __getattr__(self, name):
        if name in ("s",):
            raise AttributeError(name)
        return getattr(self.s, name)

    def set(self, **kw):
        for k, v in kw.items():
            setattr(self.s, k, v)

    def construct(self):
        rc = lib().orc_construct(C.byref(self.s))
        if rc != 0:
            raise ValueError("construct failed (%d)" % rc)
        return self

    def tick(self):
        it = C.c_uint32(0)
        n2 = C.c_double(0)
        rc = lib().orc_run_simulation_tick(C.byref(self.s), C.byref(it), C.byref(n2))
        if rc != 0:
            raise RuntimeError("tick failed (%d)" % rc)
        return it.value, n2.value

    def call(self, name, *args):
        return getattr(lib(), "orc_" + name)(C.byref(self.s), *args)

    def solve(self):
        it = C.c_uint32(0)
        n2 = C.c_double(0)
        fn = lib().orc_solve_rb if self.s.solver == SOLVER_RB else lib().orc_solve_sor
        fn(C.byref(self.s), C.byref(it), C.byref(n2))
        return it.value, n2.value

    def max_uv(self):
        a, b = C.c_double(0), C.c_double(0)
        lib().orc_max_uv(C.byref(self.s), C.byref(a), C.byref(b))
        return a.value, b.value

    def render(self, mode=0):
        """src/visualization.rs: (ny, nx, 4) RGBA u8; mode 0 pressure, 1 speed."""
        out = np.zeros((self.s.ny, self.s.nx, 4), np.uint8)
        lib().orc_render(C.byref(self.s), mode, out.ctypes.data_as(C.POINTER(C.c_uint8)))
        return out

    def restore_list(self):
        return [(r.idx, r.has_u, r.u, r.has_v, r.v)
                for r in self.restore[:self.s.nrestore]]
