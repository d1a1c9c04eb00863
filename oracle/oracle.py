"""ctypes access to the CPU oracle (liboracle.so) and the reference build (_ref/libpderef.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, always as the checker or the timed CPU baseline — never by the product.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

from pde_solver_amd import capi

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libpderef.so")

_oracle = None
_ref = None


def use_library(path: str) -> None:
    """Load the oracle from another build of oracle.cpp (for example the -march=native one the
    CPU baseline builds on the measuring host).  Must be called before the first oracle call."""
    global ORACLE_SO
    if _oracle is not None:
        raise RuntimeError("the oracle library is already loaded")
    ORACLE_SO = path


def _load_oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} missing: run `make -C oracle oracle`")
        lib = C.CDLL(ORACLE_SO)
        P, O, R = C.POINTER(capi.Problem), C.POINTER(capi.Options), C.POINTER(capi.Report)
        DP = capi.DP
        lib.oracle_solve.argtypes = [P, O, DP, R]
        lib.oracle_apply.argtypes = [P, DP, DP]
        lib.oracle_apply_folded.argtypes = [P, DP, DP]
        lib.oracle_build_rhs.argtypes = [P, DP]
        lib.oracle_apply_bc.argtypes = [P, DP, C.c_int]
        lib.oracle_residual_norms.argtypes = [P, DP, DP, DP]
        lib.oracle_time_series.argtypes = [P, C.POINTER(capi.TimeConfig), DP, DP]
        lib.oracle_mg_levels.argtypes = [P, C.c_int]
        lib.oracle_mg_precond.argtypes = [P, O, DP, DP]
        lib.oracle_last_error.restype = C.c_char_p
        lib.oracle_set_threads.argtypes = [C.c_int]
        _oracle = lib
    return _oracle


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def _load_ref():
    global _ref
    if _ref is None:
        lib = C.CDLL(REF_SO)
        P, O = C.POINTER(capi.Problem), C.POINTER(capi.Options)
        DP = capi.DP
        CP = C.POINTER(C.c_char_p)
        lib.ref_solve.argtypes = [P, O, C.c_char_p, CP, DP, DP, DP]
        lib.ref_apply.argtypes = [P, CP, DP, DP]
        lib.ref_residual_norms.argtypes = [P, C.c_char_p, CP, DP, DP, DP]
        lib.ref_time_series.argtypes = [P, C.c_char_p, CP, C.POINTER(capi.TimeConfig), DP]
        lib.ref_last_error.restype = C.c_char_p
        _ref = lib
    return _ref


def set_threads(n: int) -> None:
    _load_oracle().oracle_set_threads(int(n))


def max_threads() -> int:
    return int(_load_oracle().oracle_max_threads())


# ----------------------------------------------------------------------------- oracle
def solve(spec: capi.Spec, opt: capi.Options):
    lib = _load_oracle()
    p = spec.to_c()
    u = np.zeros(spec.n)
    rep = capi.Report()
    st = lib.oracle_solve(C.byref(p), C.byref(opt), capi.dptr(u), C.byref(rep))
    return st, u.reshape(spec.shape), rep


def apply(spec: capi.Spec, x: np.ndarray, folded: bool = False) -> np.ndarray:
    lib = _load_oracle()
    p = spec.to_c()
    xx = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    y = np.zeros(spec.n)
    fn = lib.oracle_apply_folded if folded else lib.oracle_apply
    fn(C.byref(p), capi.dptr(xx), capi.dptr(y))
    return y.reshape(spec.shape)


def build_rhs(spec: capi.Spec) -> np.ndarray:
    lib = _load_oracle()
    p = spec.to_c()
    b = np.zeros(spec.n)
    lib.oracle_build_rhs(C.byref(p), capi.dptr(b))
    return b.reshape(spec.shape)


def apply_bc(spec: capi.Spec, u: np.ndarray, which: int = 3) -> np.ndarray:
    lib = _load_oracle()
    p = spec.to_c()
    v = np.ascontiguousarray(u, dtype=np.float64).reshape(-1).copy()
    lib.oracle_apply_bc(C.byref(p), capi.dptr(v), int(which))
    return v.reshape(spec.shape)


def residual_norms(spec: capi.Spec, u: np.ndarray):
    lib = _load_oracle()
    p = spec.to_c()
    uu = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
    l2, linf = C.c_double(), C.c_double()
    st = lib.oracle_residual_norms(C.byref(p), capi.dptr(uu), C.byref(l2), C.byref(linf))
    return st, l2.value, linf.value


def time_series(spec: capi.Spec, tc: capi.TimeConfig) -> np.ndarray:
    lib = _load_oracle()
    p = spec.to_c()
    frames = np.zeros((max(1, tc.frames), spec.n))
    st = lib.oracle_time_series(C.byref(p), C.byref(tc), capi.dptr(frames), capi.DP())
    if st != 0:
        raise RuntimeError(lib.oracle_last_error().decode())
    return frames.reshape((max(1, tc.frames),) + spec.shape)


def mg_levels(spec: capi.Spec, max_levels: int) -> int:
    return int(_load_oracle().oracle_mg_levels(C.byref(spec.to_c()), int(max_levels)))


def mg_precond(spec: capi.Spec, opt: capi.Options, b: np.ndarray) -> np.ndarray:
    lib = _load_oracle()
    p = spec.to_c()
    bb = np.ascontiguousarray(b, dtype=np.float64).reshape(-1)
    z = np.zeros(spec.n)
    lib.oracle_mg_precond(C.byref(p), C.byref(opt), capi.dptr(bb), capi.dptr(z))
    return z.reshape(spec.shape)


# ----------------------------------------------------------------------------- reference
def _coef_arr(coef_latex: Optional[Sequence[Optional[str]]]):
    if not coef_latex:
        return None
    arr = (C.c_char_p * 7)(*[(s.encode() if s else None) for s in list(coef_latex) + [None] * 7][:7])
    return C.cast(arr, C.POINTER(C.c_char_p)), arr


def ref_solve(spec: capi.Spec, opt: capi.Options, rhs_latex: Optional[str] = None,
              coef_latex: Optional[Sequence[Optional[str]]] = None):
    lib = _load_ref()
    p = spec.to_c()
    u = np.zeros(spec.n)
    l2, linf = C.c_double(), C.c_double()
    ca = _coef_arr(coef_latex)
    st = lib.ref_solve(C.byref(p), C.byref(opt), rhs_latex.encode() if rhs_latex else None,
                       ca[0] if ca else None, capi.dptr(u), C.byref(l2), C.byref(linf))
    err = lib.ref_last_error().decode()
    return st, err, u.reshape(spec.shape), l2.value, linf.value


def ref_apply(spec: capi.Spec, x: np.ndarray, coef_latex=None) -> np.ndarray:
    lib = _load_ref()
    p = spec.to_c()
    xx = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    y = np.zeros(spec.n)
    ca = _coef_arr(coef_latex)
    if lib.ref_apply(C.byref(p), ca[0] if ca else None, capi.dptr(xx), capi.dptr(y)) != 0:
        raise RuntimeError(lib.ref_last_error().decode())
    return y.reshape(spec.shape)


def ref_residual_norms(spec: capi.Spec, u: np.ndarray, rhs_latex: Optional[str] = None,
                       coef_latex=None):
    lib = _load_ref()
    p = spec.to_c()
    uu = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
    l2, linf = C.c_double(), C.c_double()
    ca = _coef_arr(coef_latex)
    st = lib.ref_residual_norms(C.byref(p), rhs_latex.encode() if rhs_latex else None,
                                ca[0] if ca else None, capi.dptr(uu), C.byref(l2), C.byref(linf))
    return st, l2.value, linf.value


def ref_time_series(spec: capi.Spec, tc: capi.TimeConfig, rhs_latex: Optional[str] = None,
                    coef_latex=None) -> np.ndarray:
    lib = _load_ref()
    p = spec.to_c()
    frames = np.zeros((max(1, tc.frames), spec.n))
    ca = _coef_arr(coef_latex)
    st = lib.ref_time_series(C.byref(p), rhs_latex.encode() if rhs_latex else None,
                             ca[0] if ca else None, C.byref(tc), capi.dptr(frames))
    if st != 0:
        raise RuntimeError(lib.ref_last_error().decode())
    return frames.reshape((max(1, tc.frames),) + spec.shape)
