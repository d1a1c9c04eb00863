"""Mixed-precision BiCGStab and GMRES (FP32 Krylov vectors, FP64 solution and residual
refinement) against the FP64 GPU path and the CPU oracle: upwind convection-diffusion in 2D
(the nonsymmetric batch's operator), a FLUX problem with mixed boundaries, and a 3D case."""
import numpy as np
import pytest

from oracle import oracle as O
from pde_solver_amd import capi
from pde_solver_amd.capi import BC

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def ctx():
    from pde_solver_amd import hip
    c = hip.Context()
    yield c
    c.close()


def opts(method, precision, tol):
    return capi.options(method, tol=tol, tol_mode=capi.TOL_REL, precond=capi.PC_JACOBI,
                        bc_rows=capi.ROWS_FOLDED, gmres_restart=30, max_iter=20000,
                        precision=precision)


def convdiff(n, eps, theta, seed):
    s = capi.Spec(n, n, a=-eps, b=-eps, c=float(np.cos(theta)), d=float(np.sin(theta)),
                  convection=capi.CONV_UPWIND)
    s.f_field = np.random.default_rng(seed).standard_normal(s.shape)
    return s


@pytest.mark.parametrize("method", [capi.BICGSTAB, capi.GMRES])
@pytest.mark.parametrize("n,eps", [(65, 1e-2), (129, 1e-1), (129, 1e-3)])
def test_mixed_krylov_convdiff(ctx, method, n, eps):
    s = convdiff(n, eps, 2.1, seed=n)
    print(f"mixed krylov method={method} n={n} eps={eps}", flush=True)
    st64, u64, r64 = ctx.solve(s, opts(method, capi.FP64, 1e-10))
    st, u, rep = ctx.solve(s, opts(method, capi.MIXED, 1e-10))
    assert st == 0 and st64 == 0, ctx.error()
    assert rep.rel_residual < 1.5e-10
    # both stop at rel-resid 1e-10; their difference is bounded by the conditioning
    assert rel(u, u64) < 1e-7
    if n <= 65:
        _, uo, _ = O.solve(s, opts(method, capi.FP64, 1e-12))
        assert rel(u, uo) < 1e-7
    print(f"  fp64 its {r64.iterations}, mixed refinements {rep.refinements} inner {rep.inner_iterations}", flush=True)


@pytest.mark.parametrize("method", [capi.BICGSTAB, capi.GMRES])
def test_mixed_krylov_flux_robin(ctx, method):
    n = 65
    rng = np.random.default_rng(4)
    s = capi.Spec(n, n, a=-1.0, b=-1.0, e=3.0, op_kind=capi.OP_FLUX)
    s.kappa = np.exp(0.5 * rng.standard_normal((n, n)))
    s.f_field = rng.standard_normal((n, n))
    s.bcs = [BC.dirichlet(0.0), BC.robin(1.0, 1.0, 0.0), BC.neumann(0.0), BC.dirichlet(0.0), BC(), BC()]
    s.bc_convention = capi.BCCONV_OUTWARD
    st, u, rep = ctx.solve(s, opts(method, capi.MIXED, 1e-10))
    assert st == 0, ctx.error()
    _, uo, _ = O.solve(s, opts(method, capi.FP64, 1e-12))
    assert rel(u, uo) < 1e-7


@pytest.mark.parametrize("method", [capi.BICGSTAB, capi.GMRES])
def test_mixed_krylov_3d(ctx, method):
    n = 17
    s = capi.Spec(n, n, n, a=-1.0, b=-1.0, az=-1.0, c=0.5, d=-0.3, dz=0.2, convection=capi.CONV_UPWIND)
    s.f_field = np.random.default_rng(5).standard_normal(s.shape)
    st, u, rep = ctx.solve(s, opts(method, capi.MIXED, 1e-10))
    assert st == 0, ctx.error()
    _, uo, _ = O.solve(s, opts(method, capi.FP64, 1e-12))
    assert rel(u, uo) < 1e-7


def test_mixed_bicgstab_diffusion_dominated_completes(ctx):
    """eps = 0.1 at 513^2: the FP32 BiCGStab iteration diverges there; the solve must still
    return the FP64 answer (completed by the FP64 iteration, refinements = -1) or converge."""
    s = convdiff(513, 0.1, 1.0, seed=1)
    st64, u64, _ = ctx.solve(s, opts(capi.BICGSTAB, capi.FP64, 1e-8))
    st, u, rep = ctx.solve(s, opts(capi.BICGSTAB, capi.MIXED, 1e-8))
    assert st == 0 and st64 == 0, ctx.error()
    assert rep.rel_residual < 1.5e-8
    assert rel(u, u64) < 1e-5


@pytest.mark.parametrize("case", ["pointwise2d", "flux3d"])
@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
def test_jacobi_bicgstab_operator_kinds(ctx, case, precision):
    """Jacobi-preconditioned BiCGStab (the Jacobi scaling carried by the p and s updates) on a
    point-wise variable-coefficient 2D operator and a 3D flux operator with Robin faces,
    against the oracle."""
    rng = np.random.default_rng(17)
    if case == "pointwise2d":
        n = 49
        s = capi.Spec(n, n, op_kind=capi.OP_POINTWISE, c=0.4, d=-0.2)
        s.coef = {capi.COEF_A: -1.0 - rng.random(s.shape), capi.COEF_B: -0.5 - rng.random(s.shape),
                  capi.COEF_E: rng.random(s.shape)}
    else:
        n = 17
        s = capi.Spec(n, n, n, a=-1.0, e=1.0, op_kind=capi.OP_FLUX)
        s.kappa = np.exp(0.4 * rng.standard_normal(s.shape))
        s.bcs = [BC.dirichlet(0.0)] * 4 + [BC.robin(1.0, 1.0, 0.0)] * 2
        s.bc_convention = capi.BCCONV_OUTWARD
    s.f_field = rng.standard_normal(s.shape)
    print(f"jacobi bicgstab {case} precision={precision}", flush=True)
    st, u, rep = ctx.solve(s, opts(capi.BICGSTAB, precision, 1e-10))
    assert st == 0, ctx.error()
    _, uo, ro = O.solve(s, opts(capi.BICGSTAB, capi.FP64, 1e-12))
    print(f"  its {rep.iterations} (oracle {ro.iterations}) rel {rel(u, uo):.2e}", flush=True)
    assert rel(u, uo) < (1e-8 if precision == capi.FP64 else 1e-7)
