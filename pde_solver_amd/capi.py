"""ctypes mirror of include/pdehip.h and a small builder for pdehip_problem.

The structures here are byte-for-byte the C structures of the ABI; the same definitions are
used to drive the HIP library (libpdehip.so), the CPU oracle and the reference build, so a
parity test hands one problem description to all of them.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

# ----------------------------------------------------------------------------- enums
OK, ERR_INVALID, ERR_UNSUPPORTED, ERR_NOT_CONVERGED, ERR_BREAKDOWN, ERR_DEGENERATE = 0, 1, 2, 3, 4, 5
ERR_DEVICE, ERR_OOM, ERR_CANCELLED, ERR_ADMISSION, PENDING, ERR_TIMEOUT = 6, 7, 8, 9, 10, 11
STATUS_NAMES = {0: "ok", 1: "invalid", 2: "unsupported", 3: "not_converged", 4: "breakdown",
                5: "degenerate", 6: "device", 7: "oom", 8: "cancelled", 9: "admission",
                10: "pending", 11: "timeout"}

DIRICHLET, NEUMANN, ROBIN = 0, 1, 2
LEFT, RIGHT, BOTTOM, TOP, FRONT, BACK = range(6)
BCCONV_AXIS, BCCONV_OUTWARD = 0, 1
OP_CONST, OP_POINTWISE, OP_FLUX = 0, 1, 2
CONV_CENTERED, CONV_UPWIND = 0, 1
NL_POWER, NL_SIN, NL_COS, NL_EXP, NL_ABS = range(5)
COEF_A, COEF_B, COEF_AZ, COEF_C, COEF_D, COEF_DZ, COEF_E = range(7)

JACOBI, GAUSS_SEIDEL, SOR, CG, BICGSTAB, GMRES, MG, MGPCG = range(8)
FP64, MIXED = 0, 1
PC_NONE, PC_JACOBI, PC_MG = 0, 1, 2
TOL_ABS, TOL_REL = 0, 1
ROWS_REFERENCE, ROWS_FOLDED = 0, 1
TI_FE, TI_RK2, TI_RK4, TI_SSPRK2, TI_SSPRK3 = range(5)

MAX_NONLINEAR = 4

DP = C.POINTER(C.c_double)


class Expr(C.Structure):
    _fields_ = [("c", C.c_double), ("cx", C.c_double), ("cy", C.c_double), ("cz", C.c_double),
                ("samples", DP)]


class FaceBC(C.Structure):
    _fields_ = [("kind", C.c_int), ("value", Expr), ("alpha", Expr), ("beta", Expr), ("gamma", Expr)]


class Nonlinear(C.Structure):
    _fields_ = [("kind", C.c_int), ("coeff", C.c_double), ("power", C.c_int)]


class Problem(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("xmin", C.c_double), ("xmax", C.c_double), ("ymin", C.c_double),
        ("ymax", C.c_double), ("zmin", C.c_double), ("zmax", C.c_double),
        ("op_kind", C.c_int),
        ("a", C.c_double), ("b", C.c_double), ("az", C.c_double), ("c", C.c_double),
        ("d", C.c_double), ("dz", C.c_double), ("e", C.c_double),
        ("coef", DP * 7),
        ("kappa", DP),
        ("convection", C.c_int),
        ("f", C.c_double),
        ("f_field", DP),
        ("bc", FaceBC * 6),
        ("bc_convention", C.c_int),
        ("n_nonlinear", C.c_int),
        ("nonlinear", Nonlinear * MAX_NONLINEAR),
        ("ut", C.c_double), ("utt", C.c_double),
        ("initial", DP), ("initial_velocity", DP),
    ]


class Options(C.Structure):
    _fields_ = [
        ("method", C.c_int), ("max_iter", C.c_int), ("tol", C.c_double), ("tol_mode", C.c_int),
        ("sor_omega", C.c_double), ("gmres_restart", C.c_int),
        ("mg_pre", C.c_int), ("mg_post", C.c_int), ("mg_coarse_iters", C.c_int),
        ("mg_max_levels", C.c_int), ("mg_single_cycle", C.c_int),
        ("precond", C.c_int), ("precision", C.c_int), ("bc_rows", C.c_int),
        ("inner_tol", C.c_double), ("check_every", C.c_int),
        ("cancel", C.POINTER(C.c_int)),
    ]


class Report(C.Structure):
    _fields_ = [
        ("status", C.c_int), ("iterations", C.c_int), ("inner_iterations", C.c_int),
        ("refinements", C.c_int), ("converged", C.c_int),
        ("final_metric", C.c_double), ("rel_residual", C.c_double),
        ("residual_l2", C.c_double), ("residual_linf", C.c_double), ("solve_ms", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class TimeConfig(C.Structure):
    _fields_ = [("integrator", C.c_int), ("t_start", C.c_double), ("dt", C.c_double),
                ("frames", C.c_int), ("precision", C.c_int), ("buffer_frames", C.c_int)]


class BatchConfig(C.Structure):
    _fields_ = [("precision", C.c_int), ("use_graph", C.c_int), ("autotune", C.c_int),
                ("tile_x", C.c_int), ("tile_y", C.c_int), ("waves_per_cu", C.c_int)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_size_t), ("max_inflight", C.c_int), ("streams", C.c_int),
                ("cache_operators", C.c_int), ("timeout_ms", C.c_double),
                ("batch_lanes", C.c_int), ("batch_max", C.c_int)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("rejected", C.c_uint64),
                ("cache_hits", C.c_uint64), ("bytes_in_use", C.c_size_t), ("peak_bytes", C.c_size_t),
                ("batched", C.c_uint64), ("batches", C.c_uint64), ("timed_out", C.c_uint64)]


FRAME_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_double, DP, DP)


def dptr(a: Optional[np.ndarray]):
    if a is None:
        return DP()
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "arrays must be C-contiguous float64"
    return a.ctypes.data_as(DP)


# ----------------------------------------------------------------------------- builder
@dataclass
class BC:
    """One face's boundary condition; values are constant + linear (c, cx, cy, cz) or samples."""
    kind: int = DIRICHLET
    value: Sequence[float] = (0.0, 0.0, 0.0, 0.0)
    alpha: Sequence[float] = (0.0, 0.0, 0.0, 0.0)
    beta: Sequence[float] = (0.0, 0.0, 0.0, 0.0)
    gamma: Sequence[float] = (0.0, 0.0, 0.0, 0.0)
    value_samples: Optional[np.ndarray] = None

    @staticmethod
    def dirichlet(c=0.0, cx=0.0, cy=0.0, cz=0.0):
        return BC(DIRICHLET, value=(c, cx, cy, cz))

    @staticmethod
    def neumann(g=0.0):
        return BC(NEUMANN, value=(g, 0.0, 0.0, 0.0))

    @staticmethod
    def robin(alpha, beta, gamma):
        return BC(ROBIN, alpha=(alpha, 0, 0, 0), beta=(beta, 0, 0, 0), gamma=(gamma, 0, 0, 0))


@dataclass
class Spec:
    """A pdehip_problem with numpy-owned arrays (kept alive by this object)."""
    nx: int
    ny: int
    nz: int = 1
    bounds: Sequence[float] = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    op_kind: int = OP_CONST
    a: float = 0.0
    b: float = 0.0
    az: float = 0.0
    c: float = 0.0
    d: float = 0.0
    dz: float = 0.0
    e: float = 0.0
    coef: Dict[int, np.ndarray] = field(default_factory=dict)
    kappa: Optional[np.ndarray] = None
    convection: int = CONV_CENTERED
    f: float = 0.0
    f_field: Optional[np.ndarray] = None
    bcs: List[BC] = field(default_factory=lambda: [BC() for _ in range(6)])
    bc_convention: int = BCCONV_AXIS
    nonlinear: List[tuple] = field(default_factory=list)   # (kind, coeff, power)
    ut: float = 0.0
    utt: float = 0.0
    initial: Optional[np.ndarray] = None
    initial_velocity: Optional[np.ndarray] = None

    @property
    def shape(self):
        return (self.nz, self.ny, self.nx) if self.nz > 1 else (self.ny, self.nx)

    @property
    def n(self) -> int:
        return self.nx * self.ny * max(1, self.nz)

    def spacing(self):
        x0, x1, y0, y1, z0, z1 = self.bounds
        dz = (z1 - z0) / (self.nz - 1) if self.nz > 1 else 1.0
        return (x1 - x0) / (self.nx - 1), (y1 - y0) / (self.ny - 1), dz

    def coords(self):
        x0, x1, y0, y1, z0, z1 = self.bounds
        dx, dy, dz = self.spacing()
        x = x0 + np.arange(self.nx) * dx
        y = y0 + np.arange(self.ny) * dy
        if self.nz > 1:
            z = z0 + np.arange(self.nz) * dz
            Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
            return X, Y, Z
        Y, X = np.meshgrid(y, x, indexing="ij")
        return X, Y, np.zeros_like(X)

    def _arr(self, a):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        return a

    def to_c(self) -> Problem:
        p = Problem()
        self._keep = []
        p.nx, p.ny, p.nz = self.nx, self.ny, self.nz
        (p.xmin, p.xmax, p.ymin, p.ymax, p.zmin, p.zmax) = [float(v) for v in self.bounds]
        p.op_kind = self.op_kind
        p.a, p.b, p.az, p.c, p.d, p.dz, p.e = self.a, self.b, self.az, self.c, self.d, self.dz, self.e
        for m in range(7):
            arr = self._arr(self.coef.get(m))
            self._keep.append(arr)
            p.coef[m] = dptr(arr)
        kap = self._arr(self.kappa)
        self._keep.append(kap)
        p.kappa = dptr(kap)
        p.convection = self.convection
        p.f = self.f
        ff = self._arr(self.f_field)
        self._keep.append(ff)
        p.f_field = dptr(ff)
        for fidx, bc in enumerate(self.bcs[:6]):
            fb = p.bc[fidx]
            fb.kind = bc.kind
            for name in ("value", "alpha", "beta", "gamma"):
                vals = list(getattr(bc, name)) + [0.0] * 4
                ex = getattr(fb, name)
                ex.c, ex.cx, ex.cy, ex.cz = [float(v) for v in vals[:4]]
                ex.samples = DP()
            if bc.value_samples is not None:
                vs = self._arr(bc.value_samples)
                self._keep.append(vs)
                fb.value.samples = dptr(vs)
        p.bc_convention = self.bc_convention
        p.n_nonlinear = len(self.nonlinear)
        for t, (kind, coeff, power) in enumerate(self.nonlinear[:MAX_NONLINEAR]):
            p.nonlinear[t].kind, p.nonlinear[t].coeff, p.nonlinear[t].power = kind, coeff, power
        p.ut, p.utt = self.ut, self.utt
        ini = self._arr(self.initial)
        vel = self._arr(self.initial_velocity)
        self._keep += [ini, vel]
        p.initial = dptr(ini)
        p.initial_velocity = dptr(vel)
        return p


def options(method: int = JACOBI, **kw) -> Options:
    """pdehip_options with SolverConfig defaults (include/pde_types.h:282-307)."""
    o = Options()
    o.method = method
    o.max_iter = 10000
    o.tol = 1e-6
    o.tol_mode = TOL_ABS
    o.sor_omega = 1.5
    o.gmres_restart = 30
    o.mg_pre, o.mg_post, o.mg_coarse_iters, o.mg_max_levels = 2, 2, 50, 10
    o.mg_single_cycle = 0
    o.precond = PC_NONE
    o.precision = FP64
    o.bc_rows = ROWS_REFERENCE
    o.inner_tol = 0.0
    o.check_every = 0
    o.cancel = C.POINTER(C.c_int)()
    for k, v in kw.items():
        if not hasattr(o, k):
            raise AttributeError(k)
        setattr(o, k, v)
    return o


def time_config(integrator=TI_FE, t_start=0.0, dt=0.01, frames=1, precision=FP64, buffer_frames=0):
    t = TimeConfig()
    t.integrator, t.t_start, t.dt, t.frames = integrator, t_start, dt, frames
    t.precision, t.buffer_frames = precision, buffer_frames
    return t
