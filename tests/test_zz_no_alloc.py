"""pdehip_residual_norms_noalloc (device-resident inputs, static scratch, no allocation) against
pdehip_residual_norms (host inputs) on the same data: 2D CONST with Dirichlet / Neumann / Robin
faces and a nonlinear term, 2D FLUX with per-node face samples, 3D FLUX at the 257^3 workload
size.  Two calls in a row, the second with a smaller grid, so stale scratch contents would
show.  Size and operator limits are rejected before any launch."""
import ctypes as C

import numpy as np
import pytest

from pde_solver_amd import capi
from pde_solver_amd.capi import BC

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from pde_solver_amd import hip
    c = hip.Context()
    yield c
    c.close()


_hiprt = None


def hiprt():
    """The HIP runtime the library itself links (the test allocates the caller's buffers)."""
    global _hiprt
    if _hiprt is None:
        from pde_solver_amd import hip
        hip.lib()
        _hiprt = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
        _hiprt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hiprt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hiprt.hipFree.argtypes = [C.c_void_p]
    return _hiprt


class DevArray:
    """A caller-owned device copy of a float64 array."""

    def __init__(self, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
        self.ptr = C.c_void_p()
        assert hiprt().hipMalloc(C.byref(self.ptr), max(8, a.nbytes)) == 0
        assert hiprt().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, 1) == 0   # host to device

    def data_ptr(self):
        return self.ptr.value

    def __del__(self):
        if self.ptr:
            hiprt().hipFree(self.ptr)


def dev(arr, keep):
    t = DevArray(arr)
    keep.append(t)
    return t


def device_problem(spec, keep, face_samples=None):
    """spec.to_c() with every array member replaced by a device copy."""
    p = spec.to_c()
    if spec.f_field is not None:
        p.f_field = C.cast(C.c_void_p(dev(spec.f_field, keep).data_ptr()), capi.DP)
    if spec.kappa is not None:
        p.kappa = C.cast(C.c_void_p(dev(spec.kappa, keep).data_ptr()), capi.DP)
    for (face, which), arr in (face_samples or {}).items():
        ex = getattr(p.bc[face], which)
        ex.samples = C.cast(C.c_void_p(dev(arr, keep).data_ptr()), capi.DP)
    return p


def case_2d_const(n, seed):
    rng = np.random.default_rng(seed)
    s = capi.Spec(n, n + 6, a=-1.0, b=-0.7, c=0.3, e=0.4)
    s.f_field = rng.standard_normal(s.shape)
    s.bcs = [BC.robin(1.0, 0.5, 0.2), BC.neumann(0.3), BC.dirichlet(0.1, 0.5, -0.2),
             BC.robin(2.0, 1.0, -0.1), BC(), BC()]
    s.bc_convention = capi.BCCONV_OUTWARD
    s.nonlinear = [(capi.NL_POWER, 0.5, 3)]
    u = rng.standard_normal(s.shape)
    return s, u, None


def case_2d_flux(n, seed):
    rng = np.random.default_rng(seed)
    s = capi.Spec(n, n, a=-1.0, op_kind=capi.OP_FLUX)
    s.kappa = np.exp(0.5 * rng.standard_normal(s.shape))
    s.f_field = rng.standard_normal(s.shape)
    s.bcs = [BC.dirichlet(0.0), BC.neumann(0.0), BC.robin(1.0, 1.0, 0.0), BC.dirichlet(0.2), BC(), BC()]
    s.bc_convention = capi.BCCONV_AXIS
    samples = rng.standard_normal(n)       # Dirichlet data of the left face, per node
    u = rng.standard_normal(s.shape)
    return s, u, {(capi.LEFT, "value"): samples}


def case_3d_flux(n, seed):
    rng = np.random.default_rng(seed)
    s = capi.Spec(n, n, n, a=-1.0, op_kind=capi.OP_FLUX)
    s.kappa = np.exp(0.5 * rng.standard_normal(s.shape))
    s.f_field = rng.standard_normal(s.shape)
    s.bcs = [BC.dirichlet(0.0)] * 4 + [BC.robin(1.0, 1.0, 0.0)] * 2
    s.bc_convention = capi.BCCONV_OUTWARD
    u = rng.standard_normal(s.shape)
    return s, u, None


def host_norms(ctx, spec, u, samples):
    if samples:   # the host entry takes host samples
        p = spec.to_c()
        keep = []
        for (face, which), arr in samples.items():
            a = np.ascontiguousarray(arr, dtype=np.float64)
            keep.append(a)
            getattr(p.bc[face], which).samples = capi.dptr(a)
        uu = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        from pde_solver_amd import hip
        l2, linf = C.c_double(), C.c_double()
        st = hip.lib().pdehip_residual_norms(ctx.h, C.byref(p), capi.dptr(uu), C.byref(l2), C.byref(linf))
        assert st == 0, ctx.error()
        return l2.value, linf.value
    return ctx.residual_norms(spec, u)


@pytest.mark.parametrize("make,n_big,n_small", [(case_2d_const, 129, 33), (case_2d_flux, 257, 17),
                                                (case_3d_flux, 257, 33)])
def test_noalloc_residual_norms_equal_allocating_entry(ctx, make, n_big, n_small):
    for n in (n_big, n_small):
        print(f"no-alloc residual norms {make.__name__} n={n}", flush=True)
        spec, u, samples = make(n, seed=n)
        keep = []
        p = device_problem(spec, keep, samples)
        ud = dev(u, keep)
        st, l2, linf = ctx.residual_norms_noalloc(p, ud.data_ptr())
        assert st == 0, ctx.error()
        l2_ref, linf_ref = host_norms(ctx, spec, u, samples)
        print(f"  noalloc l2 {l2:.17g} linf {linf:.17g}  host l2 {l2_ref:.17g} linf {linf_ref:.17g}",
              flush=True)
        assert l2 > 0 and linf > 0
        assert l2 == l2_ref and linf == linf_ref


def test_noalloc_limits_rejected_before_launch(ctx):
    keep = []
    s = capi.Spec(16386, 3, a=-1.0, b=-1.0)        # 2D side above the scratch's face capacity
    p = s.to_c()
    u = dev(np.zeros(8), keep)                      # never read: rejected first
    st, _, _ = ctx.residual_norms_noalloc(p, u.data_ptr())
    assert st == capi.ERR_INVALID and "16385" in ctx.error()
    s = capi.Spec(9, 9, a=-1.0, b=-1.0, op_kind=capi.OP_POINTWISE)
    s.coef = {capi.COEF_A: np.ones(s.shape)}
    st, _, _ = ctx.residual_norms_noalloc(s.to_c(), u.data_ptr())
    assert st == capi.ERR_UNSUPPORTED
