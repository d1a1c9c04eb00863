"""The batched lock-step GMRES(m) (pdehip_batch_* / pdehip_solve_batch with method GMRES)
against one-at-a-time resident-plan GMRES solves and the CPU oracle's GMRESSolve (reference
krylov.cpp:187-333): uneven 2D sizes, upwind convection-diffusion with random directions,
Dirichlet and folded Neumann/Robin faces, Jacobi and no preconditioner, restart lengths below
and above 64, a final cycle shortened by max_iter, graph replay and resident reuse."""
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


def gopts(precond=capi.PC_JACOBI, tol=1e-10, restart=30, max_iter=5000, precision=capi.FP64):
    return capi.options(capi.GMRES, tol=tol, tol_mode=capi.TOL_REL, precond=precond,
                        bc_rows=capi.ROWS_FOLDED, gmres_restart=restart, max_iter=max_iter,
                        precision=precision)


SIZES = [33, 65, 129, 33, 257, 65, 17]


@pytest.mark.parametrize("precond", [capi.PC_JACOBI, capi.PC_NONE])
@pytest.mark.parametrize("mixed_faces", [False, True])
def test_batch_gmres_matches_plans(ctx, precond, mixed_faces):
    """The batch runs each problem's single-path GMRES: the same algorithm, the Gram-Schmidt
    sums in another order, so iteration counts agree to a few and solutions to the
    conditioning (both at rel-resid 1e-10).  A problem on which restarted GMRES stagnates
    (one of the near-singular Neumann/Robin sets) fails the same way in both: NOT_CONVERGED at
    the iteration cap."""
    print(f"batch gmres precond={precond} mixed_faces={mixed_faces}", flush=True)
    specs = problems(SIZES, seed=13 + precond, mixed_faces=mixed_faces)
    st, us, reps = ctx.solve_batch(specs, gopts(precond=precond), check=False)
    for i, s in enumerate(specs):
        st1, u1, r1 = ctx.solve(s, gopts(precond=precond), check=False)
        print(f"  {i} n={s.nx} its batch {reps[i].iterations} plan {r1.iterations} "
              f"rel {rel(us[i], u1):.2e} relres {reps[i].rel_residual:.2e} status {reps[i].status}/{st1}",
              flush=True)
        assert reps[i].status == st1
        if st1 != capi.OK:
            assert st1 == capi.ERR_NOT_CONVERGED
            assert reps[i].iterations == r1.iterations == 5000
            assert np.all(np.isfinite(us[i]))
            continue
        # over dozens of restart cycles the rounding differences move the count a little
        # (2366 against 2226 on the slowest unpreconditioned case)
        assert abs(reps[i].iterations - r1.iterations) <= max(3, r1.iterations // 8)
        assert rel(us[i], u1) < 1e-7
        assert reps[i].rel_residual < 1e-10 * 1.01
    assert (st == capi.OK) == all(r.status == capi.OK for r in reps)


def test_batch_gmres_matches_oracle(ctx):
    """Small problems against the CPU oracle's GMRES on the same folded system."""
    specs = problems([17, 33, 33, 17, 65], seed=5, mixed_faces=True)
    st, us, reps = ctx.solve_batch(specs, gopts(tol=1e-11))
    for i, s in enumerate(specs):
        st_o, u_o, r_o = O.solve(s, gopts(tol=1e-11))
        print(f"  {i} n={s.nx} its {reps[i].iterations} oracle {r_o.iterations} "
              f"rel vs oracle {rel(us[i], u_o):.2e}", flush=True)
        assert st_o == 0
        assert rel(us[i], u_o) < 1e-8
        assert abs(reps[i].iterations - r_o.iterations) <= 2


def test_batch_gmres_long_restart_and_short_last_cycle(ctx):
    """Restart 80 (Hessenberg and partials sized by the restart) converges like the plans; a cap
    of 45 iterations with restart 30 ends in a 15-step cycle: every problem then reports
    NOT_CONVERGED after exactly 45 iterations, with a finite, reduced residual."""
    specs = problems([65, 129, 33], seed=7)
    st, us, reps = ctx.solve_batch(specs, gopts(restart=80, max_iter=3000))
    for i, s in enumerate(specs):
        _, u1, r1 = ctx.solve(s, gopts(restart=80, max_iter=3000))
        assert reps[i].status == capi.OK
        assert rel(us[i], u1) < 1e-7
    hard = problems([129, 257], seed=9)
    st, us, reps = ctx.solve_batch(hard, gopts(precond=capi.PC_NONE, tol=1e-12, max_iter=45),
                                   check=False)
    for r, u in zip(reps, us):
        print(f"  capped: status {r.status} its {r.iterations} relres {r.rel_residual:.2e}", flush=True)
        assert r.status == capi.ERR_NOT_CONVERGED
        assert r.iterations == 45
        assert np.all(np.isfinite(u))
        assert 0.0 < r.rel_residual < 1.0
    assert st == capi.ERR_NOT_CONVERGED


def test_batch_gmres_graph_and_resident(ctx):
    """Graph replay gives the same bits as eager launches, on two solves in a row of a resident
    batch."""
    specs = problems([65, 33, 129, 65], seed=31)
    b0 = ctx.batch(specs, gopts(), capi.BatchConfig(use_graph=0))
    b1 = ctx.batch(specs, gopts(), capi.BatchConfig(use_graph=1))
    try:
        _, r0 = b0.solve()
        u0 = [b0.download(i) for i in range(len(specs))]
        for _ in range(2):
            _, r1 = b1.solve()
            u1 = [b1.download(i) for i in range(len(specs))]
            for i in range(len(specs)):
                assert r1[i].iterations == r0[i].iterations
                assert np.array_equal(u1[i], u0[i])
    finally:
        b0.close()
        b1.close()


def test_batch_gmres_mixed_runs_one_plan_after_the_other(ctx):
    """Mixed-precision GMRES is not taken by the batched engine: creating such a batch is
    refused, and pdehip_solve_batch solves the set one plan after the other instead."""
    from pde_solver_amd import hip
    specs = problems([33, 65], seed=3)
    opt = gopts(precision=capi.MIXED, tol=1e-8)
    with pytest.raises(hip.HipError):
        ctx.batch(specs, opt)
    st, us, reps = ctx.solve_batch(specs, opt)
    for i, s in enumerate(specs):
        _, u1, _ = ctx.solve(s, opt)
        assert reps[i].status == capi.OK
        assert np.array_equal(us[i], u1)
