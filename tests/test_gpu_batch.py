"""The batched lock-step MG-PCG engine (pdehip_batch_*, and pdehip_solve_batch for eligible
sets) against one-at-a-time resident-plan solves and the CPU oracle: uneven 2D sizes, Poisson
and Helmholtz, Dirichlet / Neumann / Robin faces drawn per face, FP64 and mixed precision."""
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


def problems(sizes, seed):
    rng = np.random.default_rng(seed)
    kinds = [BC.dirichlet(0.0), BC.neumann(0.0), BC.robin(1.0, 1.0, 0.0)]
    out = []
    for t, n in enumerate(sizes):
        c = 0.0 if t % 2 == 0 else float(rng.uniform(0, 50))
        s = capi.Spec(n, n, a=-1.0, b=-1.0, e=c)
        x, y = np.meshgrid(np.linspace(0, 1, n), np.linspace(0, 1, n))
        f = np.zeros((n, n))
        for _ in range(8):
            p, q = rng.integers(1, 9, size=2)
            f += rng.standard_normal() * np.sin(np.pi * p * x + rng.uniform(0, 2 * np.pi)) * \
                np.sin(np.pi * q * y + rng.uniform(0, 2 * np.pi))
        s.f_field = -f
        faces = [kinds[k] for k in rng.integers(0, 3, size=4)]
        if c == 0.0 and all(fc.kind == capi.NEUMANN for fc in faces):
            faces[0] = BC.dirichlet(0.0)
        s.bcs = faces + [BC(), BC()]
        s.bc_convention = capi.BCCONV_OUTWARD
        out.append(s)
    return out


def opts(precision, tol):
    return capi.options(capi.MGPCG, tol=tol, tol_mode=capi.TOL_REL, precond=capi.PC_MG,
                        bc_rows=capi.ROWS_FOLDED, precision=precision, max_iter=200)


SIZES = [33, 65, 129, 257, 33, 65, 17, 129, 513, 33]


@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
def test_batch_engine_matches_plans_and_oracle(ctx, precision):
    specs = problems(SIZES, seed=3)
    print(f"batch engine precision={precision} sizes={SIZES}", flush=True)
    opt = opts(precision, 1e-10)
    b = ctx.batch(specs, opt)
    st, reps = b.solve()
    assert st == 0
    for i, s in enumerate(specs):
        u = b.download(i)
        plan = ctx.plan(s, opt)
        _, rp = plan.solve()
        up = plan.download()
        plan.close()
        tol_plan = 1e-8 if precision == capi.FP64 else 5e-7
        d = rel(u, up)
        print(f"  {i} n={s.nx} its batch {reps[i].iterations}/{reps[i].inner_iterations} "
              f"plan {rp.iterations}/{rp.inner_iterations} rel {d:.2e} "
              f"relres {reps[i].rel_residual:.1e}", flush=True)
        assert reps[i].status == 0 and reps[i].converged
        assert reps[i].rel_residual < 2e-10
        assert d < tol_plan
        if s.nx <= 129:
            _, uo, _ = O.solve(s, opts(capi.FP64, 1e-13))
            assert rel(u, uo) < (1e-8 if precision == capi.FP64 else 5e-7)
    b.close()


def test_batch_repeat_solves_identical(ctx):
    specs = problems([65, 33, 129], seed=8)
    b = ctx.batch(specs, opts(capi.MIXED, 1e-8))
    b.solve()
    first = [b.download(i) for i in range(len(specs))]
    b.solve()
    for i in range(len(specs)):
        assert np.array_equal(first[i], b.download(i))
    b.close()
