"""The batched lock-step BiCGStab (pdehip_batch_* / pdehip_solve_batch with method BiCGStab)
against one-at-a-time resident-plan BiCGStab solves and the CPU oracle: uneven 2D sizes,
upwind convection-diffusion with random directions, Dirichlet and folded Neumann/Robin faces,
Jacobi and no preconditioner, FP64 and mixed precision."""
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


def problems(sizes, seed, mixed_faces=False):
    rng = np.random.default_rng(seed)
    out = []
    for n in sizes:
        eps = float(10.0 ** rng.uniform(-2.0, -1.0))
        th = float(rng.uniform(0.0, 2 * np.pi))
        s = capi.Spec(n, n, a=-eps, b=-eps, c=float(np.cos(th)), d=float(np.sin(th)),
                      convection=capi.CONV_UPWIND)
        x, y = np.meshgrid(np.linspace(0, 1, n), np.linspace(0, 1, n))
        f = np.zeros((n, n))
        for _ in range(4):
            p, q = rng.integers(1, 6, size=2)
            f += rng.standard_normal() * np.sin(np.pi * p * x) * np.sin(np.pi * q * y)
        s.f_field = -f
        if mixed_faces:
            s.bcs = [BC.dirichlet(0.0), BC.neumann(0.0), BC.robin(1.0, 1.0, 0.0), BC.dirichlet(0.2),
                     BC(), BC()]
            s.bc_convention = capi.BCCONV_OUTWARD
        out.append(s)
    return out


def kopts(precision=capi.FP64, precond=capi.PC_JACOBI, tol=1e-10):
    return capi.options(capi.BICGSTAB, tol=tol, tol_mode=capi.TOL_REL, precond=precond,
                        bc_rows=capi.ROWS_FOLDED, max_iter=5000, precision=precision)


SIZES = [33, 65, 129, 33, 257, 65, 129, 33]


@pytest.mark.parametrize("precond", [capi.PC_JACOBI, capi.PC_NONE])
@pytest.mark.parametrize("mixed_faces", [False, True])
def test_batch_bicgstab_fp64_matches_plans(ctx, precond, mixed_faces):
    """FP64 batch = the single-problem BiCGStab of every problem: both reach the true residual
    target (rel-resid 1e-10) and the solutions agree to the conditioning of the problem (1e-7; up
    to 9e-9 measured on the near-singular Neumann/Robin sets).  On these convection-dominated
    problems BiCGStab's iteration count is rounding-sensitive (the batch sums its dot products in
    another order; changing the batch's item shapes alone moves one problem between 507 and 985
    iterations where the single path takes 350), so counts are only bounded."""
    print(f"batch bicgstab fp64 precond={precond} mixed_faces={mixed_faces}", flush=True)
    specs = problems(SIZES, seed=11 + precond, mixed_faces=mixed_faces)
    st, us, reps = ctx.solve_batch(specs, kopts(precond=precond), check=False)
    for i, s in enumerate(specs):
        st1, u1, r1 = ctx.solve(s, kopts(precond=precond), check=False)
        print(f"  {i} n={s.nx} its batch {reps[i].iterations} plan {r1.iterations} "
              f"rel {rel(us[i], u1):.2e} status {reps[i].status}/{st1}", flush=True)
        assert reps[i].status == st1 == capi.OK
        assert reps[i].iterations <= 3 * r1.iterations + 10
        assert rel(us[i], u1) < 1e-7
        assert reps[i].rel_residual < 1e-10 * 1.01
    assert st == capi.OK


def test_batch_bicgstab_matches_oracle(ctx):
    """Small problems against the CPU oracle's BiCGStab on the same folded system."""
    specs = problems([17, 33, 33, 17], seed=5, mixed_faces=True)
    st, us, reps = ctx.solve_batch(specs, kopts(tol=1e-11))
    for i, s in enumerate(specs):
        st_o, u_o, _ = O.solve(s, kopts(tol=1e-11))
        print(f"  {i} n={s.nx} its {reps[i].iterations} rel vs oracle {rel(us[i], u_o):.2e}", flush=True)
        assert st_o == 0
        assert rel(us[i], u_o) < 1e-8


def test_batch_bicgstab_mixed(ctx):
    """Mixed precision (FP32 Krylov vectors and applies, FP64 solution and residual): every
    problem reaches the FP64 residual target, and the solution matches the FP64 batch."""
    specs = problems(SIZES, seed=21)
    st64, u64, r64 = ctx.solve_batch(specs, kopts(tol=1e-8))
    stm, um, rm = ctx.solve_batch(specs, kopts(precision=capi.MIXED, tol=1e-8), check=False)
    for i in range(len(specs)):
        print(f"  {i} n={specs[i].nx} its fp64 {r64[i].iterations} mixed {rm[i].iterations} "
              f"relres {rm[i].rel_residual:.2e} rel vs fp64 {rel(um[i], u64[i]):.2e} status {rm[i].status}",
              flush=True)
        assert rm[i].status == capi.OK
        assert rm[i].rel_residual < 1e-8 * 1.01
        assert rel(um[i], u64[i]) < 1e-5
    assert stm == capi.OK


def test_batch_bicgstab_mixed_completion_pass(ctx):
    """Problems the FP32 pass leaves unconverged start again from x0 in FP64 (refinements = -1,
    as the single-problem mixed path).  With an iteration cap far below what any problem needs,
    neither pass converges, so the completion pass runs on every problem: finite solutions
    reduced from the zero guess, status NOT_CONVERGED."""
    specs = problems([65, 129, 33], seed=41)
    opt = kopts(precision=capi.MIXED, tol=1e-10)
    opt.max_iter = 4
    st, us, reps = ctx.solve_batch(specs, opt, check=False)
    for r, u in zip(reps, us):
        print(f"  status {r.status} refinements {r.refinements} its {r.iterations} relres {r.rel_residual:.2e}",
              flush=True)
        assert r.status == capi.ERR_NOT_CONVERGED
        assert r.refinements == -1
        assert np.all(np.isfinite(u))
        assert 0.0 < r.rel_residual < 1.0
    assert st == capi.ERR_NOT_CONVERGED


def test_batch_bicgstab_graph_and_resident(ctx):
    """The resident batch with graph replay gives the same bits as without, on two solves in a
    row; kernel timing runs on a BiCGStab batch and the next solve starts afresh."""
    specs = problems([65, 33, 129, 65], seed=31)
    b0 = ctx.batch(specs, kopts(), capi.BatchConfig(use_graph=0))
    b1 = ctx.batch(specs, kopts(), capi.BatchConfig(use_graph=1))
    try:
        _, r0 = b0.solve()
        u0 = [b0.download(i) for i in range(len(specs))]
        for rep_round in range(2):
            _, r1 = b1.solve()
            u1 = [b1.download(i) for i in range(len(specs))]
            for i in range(len(specs)):
                assert r1[i].iterations == r0[i].iterations
                assert np.array_equal(u1[i], u0[i])
        ms, nbytes = b1.time_kernel(2, reps=3)
        print(f"  apply over the batch: {ms * 1e3:.1f} us, {nbytes / (ms * 1e-3) / 1e9:.0f} GB/s", flush=True)
        assert ms > 0 and nbytes > 0
        _, r2 = b1.solve()
        for i in range(len(specs)):
            assert r2[i].iterations == r0[i].iterations
    finally:
        b0.close()
        b1.close()
