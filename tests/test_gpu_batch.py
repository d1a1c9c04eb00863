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
        # the batched finish (closures, reference residual norms) against the oracle's norms
        # of the same downloaded solution
        _, l2, linf = O.residual_norms(s, u)
        noise = 64 * 2.2e-16 * (8.0 * (s.nx - 1) ** 2 + s.e) * np.abs(u).max()   # per-node rounding
        assert abs(reps[i].residual_l2 - l2) <= 1e-3 * l2 + noise * np.sqrt(s.n)
        assert abs(reps[i].residual_linf - linf) <= 1e-3 * linf + noise
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


@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
def test_batch_autotuned_graph_matches_default(ctx, precision):
    """Autotuned work-item shapes and the captured iteration graph change only the order of
    the partial sums: same answers to the solve tolerance, and graph replays are repeatable."""
    specs = problems([129, 33, 513, 65, 257, 17, 1025, 33], seed=5)
    print(f"batch autotune+graph precision={precision}", flush=True)
    opt = opts(precision, 1e-10)
    b0 = ctx.batch(specs, opt)
    _, r0 = b0.solve()
    u0 = [b0.download(i) for i in range(len(specs))]
    b0.close()
    b1 = ctx.batch(specs, opt, capi.BatchConfig(autotune=1, use_graph=1))
    _, r1 = b1.solve()
    u1 = [b1.download(i) for i in range(len(specs))]
    _, r2 = b1.solve()
    for i, s in enumerate(specs):
        assert r1[i].status == 0 and r1[i].rel_residual < 2e-10
        assert abs(r1[i].inner_iterations - r0[i].inner_iterations) <= 1
        assert rel(u1[i], u0[i]) < 2e-8
        # every size against the oracle, 513^2 and 1025^2 included
        _, uo, _ = O.solve(s, opts(capi.FP64, 1e-13))
        print(f"  n={s.nx} rel vs oracle {rel(u1[i], uo):.2e}", flush=True)
        assert rel(u1[i], uo) < (1e-8 if precision == capi.FP64 else 5e-7)
        assert np.array_equal(u1[i], b1.download(i))
        assert r2[i].iterations == r1[i].iterations
    b1.close()


def test_batch_mixed_fp32_coarse_levels(ctx):
    """mixed.fp16 = 0: coarse levels in FP32; same answers against the oracle."""
    from pde_solver_amd import hip
    specs = problems([65, 129, 33], seed=6)
    hip.tune_set("mixed.fp16", 0)
    try:
        b = ctx.batch(specs, opts(capi.MIXED, 1e-10))
        _, reps = b.solve()
        for i, s in enumerate(specs):
            _, uo, _ = O.solve(s, opts(capi.FP64, 1e-13))
            assert reps[i].status == 0 and reps[i].rel_residual < 2e-10
            assert rel(b.download(i), uo) < 5e-7
        b.close()
    finally:
        hip.tune_set("mixed.fp16", 1)


def test_context_close_with_open_batch_and_plan():
    """Closing a context closes the plans and batches still open on it first (regression: the
    batch's later destructor used the freed context)."""
    import gc
    from pde_solver_amd import hip
    c = hip.Context()
    specs = problems([33, 65], seed=9)
    b = c.batch(specs, opts(capi.FP64, 1e-8))
    p = c.plan(specs[0], opts(capi.FP64, 1e-8))
    b.solve()
    c.close()
    assert not b.h and not p.h
    del b, p
    gc.collect()


@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
def test_nonfinite_source_is_an_error_not_convergence(ctx, precision):
    """A NaN in the data must surface as an error: reductions propagate NaN (a NaN-dropping max
    or sqrt(max(0, nan)) would read it as a zero residual and report convergence)."""
    specs = problems([33, 65], seed=10)
    specs[1].f_field[10, 10] = np.nan
    opt = opts(precision, 1e-8)
    b = ctx.batch(specs, opt)
    st, reps = b.solve(check=False)
    assert st != 0
    assert reps[0].status == 0
    assert reps[1].status != 0
    b.close()
    st, u, rep = ctx.solve(specs[1], opt, check=False)
    assert st != 0


def test_batch_time_kernel_then_solve(ctx):
    """pdehip_batch_time_kernel times the level-0 launches over every problem; a solve after it
    starts afresh and gives the same answer as before."""
    specs = problems([65, 33, 129], seed=11)
    b = ctx.batch(specs, opts(capi.MIXED, 1e-10))
    b.solve()
    before = [b.download(i) for i in range(len(specs))]
    n0 = sum(s.n for s in specs)
    for k, per_node in ((0, 12.0), (1, 12.0), (2, 8.0), (3, 24.0)):
        ms, nb = b.time_kernel(k, 3)
        assert ms > 0.0
        assert nb == per_node * n0
    with pytest.raises(Exception):
        b.time_kernel(7, 1)
    b.solve()
    for i in range(len(specs)):
        assert np.array_equal(before[i], b.download(i))
    b.close()
