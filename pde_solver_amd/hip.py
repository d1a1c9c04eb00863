"""Python access to libpdehip.so (the HIP backend's C ABI).  Plumbing for tests and benchmarks:
every call goes straight to the native library, and importing this module fails loudly when
the library is missing — there is no CPU fallback."""
from __future__ import annotations

import ctypes as C
import weakref
import os
from typing import List, Optional, Sequence

import numpy as np

from . import capi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpdehip.so")

_lib = None


class HipError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"[{capi.STATUS_NAMES.get(status, status)}] {msg}")
        self.status = status


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m pde_solver_amd.build`")
        L = C.CDLL(LIB_PATH)
        P, O, R = C.POINTER(capi.Problem), C.POINTER(capi.Options), C.POINTER(capi.Report)
        DP, VP = capi.DP, C.c_void_p
        L.pdehip_abi_version.restype = C.c_int
        L.pdehip_is_available.argtypes = [C.c_char_p, C.c_int]
        L.pdehip_create.argtypes = [C.c_int, C.c_size_t, C.POINTER(VP)]
        L.pdehip_destroy.argtypes = [VP]
        L.pdehip_last_error.argtypes = [VP]
        L.pdehip_last_error.restype = C.c_char_p
        L.pdehip_solve.argtypes = [VP, P, O, DP, R]
        L.pdehip_apply.argtypes = [VP, P, DP, DP]
        L.pdehip_residual_norms.argtypes = [VP, P, DP, DP, DP]
        L.pdehip_residual_norms_noalloc.argtypes = [VP, P, VP, DP, DP]
        L.pdehip_plan_create.argtypes = [VP, P, O, C.POINTER(VP)]
        L.pdehip_plan_set_source.argtypes = [VP, DP]
        L.pdehip_plan_solve.argtypes = [VP, R]
        L.pdehip_plan_download.argtypes = [VP, DP]
        L.pdehip_plan_bytes.argtypes = [VP]
        L.pdehip_plan_bytes.restype = C.c_size_t
        L.pdehip_plan_time_kernel.argtypes = [VP, C.c_int, C.c_int, C.c_int, DP, DP]
        L.pdehip_plan_destroy.argtypes = [VP]
        L.pdehip_time_series.argtypes = [VP, P, C.POINTER(capi.TimeConfig), capi.FRAME_FN, VP, DP, R]
        L.pdehip_solve_batch.argtypes = [VP, C.c_int, P, O, C.POINTER(capi.BatchConfig),
                                         C.POINTER(DP), R]
        L.pdehip_server_create.argtypes = [VP, C.POINTER(capi.ServerConfig), C.POINTER(VP)]
        L.pdehip_server_submit.argtypes = [VP, P, O, DP, C.POINTER(C.c_uint64)]
        L.pdehip_server_poll.argtypes = [VP, C.c_uint64, R]
        L.pdehip_server_wait.argtypes = [VP, C.c_uint64, R]
        L.pdehip_server_stats_get.argtypes = [VP, C.POINTER(capi.ServerStats)]
        L.pdehip_server_destroy.argtypes = [VP]
        L.pdehip_batch_create.argtypes = [VP, C.c_int, P, O, C.POINTER(capi.BatchConfig), C.POINTER(VP)]
        L.pdehip_batch_solve.argtypes = [VP, R]
        L.pdehip_batch_download.argtypes = [VP, C.c_int, DP]
        L.pdehip_batch_bytes.argtypes = [VP]
        L.pdehip_batch_bytes.restype = C.c_size_t
        L.pdehip_batch_time_kernel.argtypes = [VP, C.c_int, C.c_int, C.POINTER(C.c_double),
                                               C.POINTER(C.c_double)]
        L.pdehip_batch_destroy.argtypes = [VP]
        L.pdehip_tune_set.argtypes = [C.c_char_p, C.c_int]
        L.pdehip_tune_get.argtypes = [C.c_char_p, C.POINTER(C.c_int)]
        _lib = L
    return _lib


def tune_set(key: str, value: int) -> None:
    """Set a process-wide launch-shape knob (pdehip_tune_set)."""
    st = lib().pdehip_tune_set(key.encode(), int(value))
    if st != 0:
        raise HipError(st, f"invalid tuning {key}={value}")


def tune_get(key: str) -> int:
    v = C.c_int()
    st = lib().pdehip_tune_get(key.encode(), C.byref(v))
    if st != 0:
        raise HipError(st, f"unknown tuning key {key}")
    return v.value


def is_available():
    buf = C.create_string_buffer(256)
    ok = lib().pdehip_is_available(buf, 256)
    return bool(ok), buf.value.decode()


class Context:
    """One device context (stream + HBM budget)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self.h = C.c_void_p()
        self._children = weakref.WeakSet()   # plans, batches and servers made from this context
        st = lib().pdehip_create(device, hbm_budget, C.byref(self.h))
        if st != 0:
            raise HipError(st, lib().pdehip_last_error(None).decode())

    def close(self):
        """Destroy the context; plans, batches and servers still open on it are closed first
        (the C API requires them to go before their context)."""
        if self.h:
            for child in list(self._children):
                child.close()
            lib().pdehip_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def error(self) -> str:
        return lib().pdehip_last_error(self.h).decode()

    def _check(self, st, allow=(0,)):
        if st not in allow:
            raise HipError(st, self.error())
        return st

    # ------------------------------------------------------------------ calls
    def solve(self, spec: capi.Spec, opt: capi.Options, check: bool = True):
        p = spec.to_c()
        u = np.zeros(spec.n)
        rep = capi.Report()
        st = lib().pdehip_solve(self.h, C.byref(p), C.byref(opt), capi.dptr(u), C.byref(rep))
        if check:
            self._check(st)
        return st, u.reshape(spec.shape), rep

    def apply(self, spec: capi.Spec, x: np.ndarray) -> np.ndarray:
        p = spec.to_c()
        xx = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        y = np.zeros(spec.n)
        self._check(lib().pdehip_apply(self.h, C.byref(p), capi.dptr(xx), capi.dptr(y)))
        return y.reshape(spec.shape)

    def residual_norms(self, spec: capi.Spec, u: np.ndarray):
        p = spec.to_c()
        uu = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        l2, linf = C.c_double(), C.c_double()
        self._check(lib().pdehip_residual_norms(self.h, C.byref(p), capi.dptr(uu), C.byref(l2), C.byref(linf)))
        return l2.value, linf.value

    def residual_norms_noalloc(self, prob: capi.Problem, u_dev_ptr: int):
        """pdehip_residual_norms_noalloc: prob's arrays and u are device addresses (ints)."""
        l2, linf = C.c_double(), C.c_double()
        st = lib().pdehip_residual_norms_noalloc(self.h, C.byref(prob), C.c_void_p(u_dev_ptr),
                                                 C.byref(l2), C.byref(linf))
        return st, l2.value, linf.value

    def time_series(self, spec: capi.Spec, tc: capi.TimeConfig, keep_frames: bool = True,
                    on_frame=None):
        """Frames (if keep_frames or on_frame), final state, report.  With keep_frames False
        and no on_frame, no frame callback is installed: only the final state is copied."""
        p = spec.to_c()
        frames: List[np.ndarray] = []

        def cb(user, frame, t, grid, vel):
            arr = np.ctypeslib.as_array(grid, shape=(spec.n,)).copy().reshape(spec.shape)
            if keep_frames:
                frames.append(arr)
            if on_frame is not None:
                return 1 if on_frame(frame, t, arr) is not False else 0
            return 1

        fn = capi.FRAME_FN(cb) if (keep_frames or on_frame is not None) else C.cast(None, capi.FRAME_FN)
        u = np.zeros(spec.n)
        rep = capi.Report()
        self._check(lib().pdehip_time_series(self.h, C.byref(p), C.byref(tc), fn, None,
                                             capi.dptr(u), C.byref(rep)))
        return (np.stack(frames) if frames else None), u.reshape(spec.shape), rep

    def plan(self, spec: capi.Spec, opt: capi.Options) -> "Plan":
        return Plan(self, spec, opt)

    def solve_batch(self, specs: Sequence[capi.Spec], opt: capi.Options,
                    cfg: Optional[capi.BatchConfig] = None, check: bool = True):
        n = len(specs)
        probs = (capi.Problem * n)(*[s.to_c() for s in specs])
        outs = [np.zeros(s.n) for s in specs]
        ptrs = (capi.DP * n)(*[capi.dptr(o) for o in outs])
        reps = (capi.Report * n)()
        st = lib().pdehip_solve_batch(self.h, n, probs, C.byref(opt),
                                      C.byref(cfg) if cfg is not None else None, ptrs, reps)
        if check:
            self._check(st)
        return st, [o.reshape(s.shape) for o, s in zip(outs, specs)], list(reps)

    def server(self, cfg: Optional[capi.ServerConfig] = None) -> "Server":
        return Server(self, cfg)

    def batch(self, specs: Sequence[capi.Spec], opt: capi.Options,
              cfg: Optional[capi.BatchConfig] = None) -> "Batch":
        return Batch(self, specs, opt, cfg)


class Batch:
    """A resident batch: many problems solved together in lock step (pdehip_batch_*)."""

    def __init__(self, ctx: Context, specs: Sequence[capi.Spec], opt: capi.Options,
                 cfg: Optional[capi.BatchConfig] = None):
        self.ctx = ctx
        ctx._children.add(self)
        self.specs = list(specs)
        self._probs = (capi.Problem * len(specs))(*[s.to_c() for s in specs])
        self.h = C.c_void_p()
        ctx._check(lib().pdehip_batch_create(ctx.h, len(specs), self._probs, C.byref(opt),
                                             C.byref(cfg) if cfg is not None else None,
                                             C.byref(self.h)))

    def solve(self, check: bool = True):
        reps = (capi.Report * len(self.specs))()
        st = lib().pdehip_batch_solve(self.h, reps)
        if check:
            self.ctx._check(st)
        return st, list(reps)

    def download(self, i: int) -> np.ndarray:
        s = self.specs[i]
        u = np.zeros(s.n)
        self.ctx._check(lib().pdehip_batch_download(self.h, i, capi.dptr(u)))
        return u.reshape(s.shape)

    def bytes(self) -> int:
        return int(lib().pdehip_batch_bytes(self.h))

    def time_kernel(self, kernel: int, reps: int = 10):
        """(avg ms, algorithmic bytes) of one batched level-0 launch (pdehip_batch_time_kernel)."""
        ms, nb = C.c_double(), C.c_double()
        self.ctx._check(lib().pdehip_batch_time_kernel(self.h, kernel, reps, C.byref(ms), C.byref(nb)))
        return ms.value, nb.value

    def close(self):
        if self.h:
            lib().pdehip_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Plan:
    """A resident solve: operator, hierarchy and work vectors stay on the device."""

    def __init__(self, ctx: Context, spec: capi.Spec, opt: capi.Options):
        self.ctx = ctx
        ctx._children.add(self)
        self.spec = spec
        self.h = C.c_void_p()
        p = spec.to_c()
        ctx._check(lib().pdehip_plan_create(ctx.h, C.byref(p), C.byref(opt), C.byref(self.h)))

    def set_source(self, f: Optional[np.ndarray]):
        ff = None if f is None else np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
        self.ctx._check(lib().pdehip_plan_set_source(self.h, capi.dptr(ff)))

    def solve(self, check: bool = True):
        rep = capi.Report()
        st = lib().pdehip_plan_solve(self.h, C.byref(rep))
        if check:
            self.ctx._check(st)
        return st, rep

    def download(self) -> np.ndarray:
        u = np.zeros(self.spec.n)
        self.ctx._check(lib().pdehip_plan_download(self.h, capi.dptr(u)))
        return u.reshape(self.spec.shape)

    def bytes(self) -> int:
        return int(lib().pdehip_plan_bytes(self.h))

    def time_kernel(self, kernel: int, precision: int = capi.FP64, reps: int = 20):
        ms, b = C.c_double(), C.c_double()
        self.ctx._check(lib().pdehip_plan_time_kernel(self.h, kernel, precision, reps, C.byref(ms), C.byref(b)))
        return ms.value, b.value

    def close(self):
        if self.h:
            lib().pdehip_plan_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request loop."""

    def __init__(self, ctx: Context, cfg: Optional[capi.ServerConfig] = None):
        self.ctx = ctx
        ctx._children.add(self)
        self.h = C.c_void_p()
        self._keep = {}
        ctx._check(lib().pdehip_server_create(ctx.h, C.byref(cfg) if cfg is not None else None,
                                              C.byref(self.h)))

    def submit(self, spec: capi.Spec, opt: capi.Options) -> int:
        p = spec.to_c()
        out = np.zeros(spec.n)
        t = C.c_uint64()
        st = lib().pdehip_server_submit(self.h, C.byref(p), C.byref(opt), capi.dptr(out), C.byref(t))
        if st != 0:
            raise HipError(st, lib().pdehip_last_error(None).decode())
        self._keep[t.value] = (spec, out, p)
        return t.value

    def poll(self, ticket: int):
        rep = capi.Report()
        st = lib().pdehip_server_poll(self.h, ticket, C.byref(rep))
        if st == capi.PENDING:
            return st, None, None
        spec, out, _ = self._keep.pop(ticket)
        return st, out.reshape(spec.shape), rep

    def wait(self, ticket: int):
        rep = capi.Report()
        st = lib().pdehip_server_wait(self.h, ticket, C.byref(rep))
        spec, out, _ = self._keep.pop(ticket)
        return st, out.reshape(spec.shape), rep

    def stats(self) -> dict:
        s = capi.ServerStats()
        lib().pdehip_server_stats_get(self.h, C.byref(s))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def close(self):
        if self.h:
            lib().pdehip_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
