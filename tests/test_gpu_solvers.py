"""GPU parity of the extension solvers (no reference counterpart) against the CPU oracle, and
the resident-plan / serving / batch entry points against one-shot solves."""
import numpy as np
import pytest

from oracle import oracle as O
from pde_solver_amd import capi
from pde_solver_amd.capi import BC

pytestmark = pytest.mark.gpu


def rel(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


@pytest.fixture(scope="module")
def ctx():
    from pde_solver_amd import hip
    c = hip.Context()
    yield c
    c.close()


def varcoef3d(n, seed=0):
    """-div(k grad u) = f, k = exp(0.5 g) (g smooth Gaussian field), Dirichlet 0 on x/y faces,
    Robin (1, 1, 0) outward on z faces: the structure of the 257^3 workload at small n."""
    s = capi.Spec(n, n, n, a=-1.0, op_kind=capi.OP_FLUX)
    rng = np.random.default_rng(seed)
    g = rng.standard_normal(s.shape)
    for ax in range(3):   # cheap smoothing
        g = (np.roll(g, 1, ax) + 2 * g + np.roll(g, -1, ax)) / 4
    g = (g - g.mean()) / g.std()
    s.kappa = np.exp(0.5 * g)
    f = rng.standard_normal(s.shape)
    s.f_field = -(f - f.mean())
    s.bcs = [BC.dirichlet(0.0)] * 4 + [BC.robin(1.0, 1.0, 0.0)] * 2
    s.bc_convention = capi.BCCONV_OUTWARD
    return s


def mgpcg(tol=1e-12, precision=capi.FP64, levels=10, **kw):
    return capi.options(capi.MGPCG, tol=tol, tol_mode=capi.TOL_REL, precond=capi.PC_MG,
                        bc_rows=capi.ROWS_FOLDED, precision=precision, mg_max_levels=levels,
                        max_iter=200, **kw)


@pytest.mark.parametrize("n", [17, 33])
def test_mgpcg_fp64_varcoef3d_matches_oracle(ctx, n):
    print(f"mgpcg fp64 varcoef3d n={n}", flush=True)
    s = varcoef3d(n)
    st_o, u_o, rep_o = O.solve(s, mgpcg())
    assert st_o == 0
    st, u, rep = ctx.solve(s, mgpcg())
    print(f"  iterations hip {rep.iterations} oracle {rep_o.iterations}  rel {rel(u, u_o):.2e}", flush=True)
    assert rel(u, u_o) < 1e-10
    assert abs(rep.iterations - rep_o.iterations) <= 1


@pytest.mark.parametrize("n", [17, 33])
def test_mgpcg_mixed_varcoef3d_matches_fp64(ctx, n):
    print(f"mgpcg mixed varcoef3d n={n}", flush=True)
    s = varcoef3d(n, seed=1)
    st_o, u_o, _ = O.solve(s, mgpcg(tol=1e-13))
    st, u, rep = ctx.solve(s, mgpcg(tol=1e-10, precision=capi.MIXED))
    _, _, rep64 = ctx.solve(s, mgpcg(tol=1e-10))
    print(f"  replacements {rep.refinements} inner {rep.inner_iterations} (fp64 {rep64.iterations}) "
          f"rel {rel(u, u_o):.2e} relres {rep.rel_residual:.2e}", flush=True)
    assert st == 0
    assert rel(u, u_o) < 1e-8
    assert rep.rel_residual < 1e-10
    # residual replacement keeps one Krylov space: the FP32 recurrence needs about the FP64
    # iteration count (a restarted refinement would need several more)
    assert rep.inner_iterations <= rep64.iterations + 2
    assert 1 <= rep.refinements <= 12


def test_mgpcg_mixed_unconverged_keeps_correction(ctx):
    """A mixed solve stopped by max_iter before any residual replacement still returns the
    FP32 correction folded into the FP64 solution (status NOT_CONVERGED, residual reduced)."""
    s = varcoef3d(33, seed=2)
    opt = mgpcg(tol=1e-12, precision=capi.MIXED, inner_tol=1e-6)
    opt.max_iter = 2
    st, u, rep = ctx.solve(s, opt, check=False)
    print(f"  status {st} inner {rep.inner_iterations} replacements {rep.refinements} "
          f"relres {rep.rel_residual:.2e}", flush=True)
    assert st == capi.ERR_NOT_CONVERGED
    assert rep.refinements == 0
    assert rep.rel_residual < 1e-2


@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
def test_mgpcg_2d_mixed_bcs(ctx, precision):
    s = capi.Spec(129, 129, a=-1.0, b=-1.0, e=5.0)
    s.f_field = np.random.default_rng(4).standard_normal(s.shape)
    s.bcs = [BC.neumann(0.0), BC.robin(1.0, 1.0, 0.2), BC.dirichlet(0.1, 0.5), BC.neumann(0.3), BC(), BC()]
    s.bc_convention = capi.BCCONV_OUTWARD
    st_o, u_o, _ = O.solve(s, mgpcg(tol=1e-13))
    st, u, rep = ctx.solve(s, mgpcg(tol=1e-11, precision=precision))
    assert st == 0
    assert rel(u, u_o) < (1e-10 if precision == capi.FP64 else 1e-9)


def test_mg_iterated_matches_oracle(ctx):
    s = capi.Spec(65, 65, a=-1.0, b=-1.0)
    s.f_field = np.random.default_rng(2).standard_normal(s.shape)
    s.bcs = [BC.dirichlet(0.2)] * 3 + [BC.robin(1, 0.5, 0)] + [BC(), BC()]
    s.bc_convention = capi.BCCONV_OUTWARD
    opt = capi.options(capi.MG, tol=1e-11, tol_mode=capi.TOL_REL, max_iter=100)
    st_o, u_o, rep_o = O.solve(s, opt)
    st, u, rep = ctx.solve(s, opt)
    assert st == 0 and st_o == 0
    assert rel(u, u_o) < 1e-10
    assert rep.iterations == rep_o.iterations


@pytest.mark.parametrize("method", [capi.BICGSTAB, capi.GMRES])
def test_jacobi_pc_krylov_upwind_matches_oracle(ctx, method):
    s = capi.Spec(65, 65, a=-0.01, b=-0.01, c=0.8, d=-0.6, convection=capi.CONV_UPWIND)
    s.f_field = np.random.default_rng(7).standard_normal(s.shape)
    opt = capi.options(method, tol=1e-11, tol_mode=capi.TOL_REL, precond=capi.PC_JACOBI,
                       bc_rows=capi.ROWS_FOLDED, gmres_restart=30, max_iter=5000)
    st_o, u_o, rep_o = O.solve(s, opt)
    st, u, rep = ctx.solve(s, opt)
    assert st == 0 and st_o == 0, ctx.error()
    assert rel(u, u_o) < 1e-9


def test_cg_folded_robin_matches_oracle(ctx):
    s = capi.Spec(65, 65, a=-1.0, b=-1.0)
    s.f_field = np.random.default_rng(3).standard_normal(s.shape)
    s.bcs = [BC.robin(1, 1, 0.1)] * 4 + [BC(), BC()]
    s.bc_convention = capi.BCCONV_OUTWARD
    opt = capi.options(capi.CG, tol=1e-12, tol_mode=capi.TOL_REL, bc_rows=capi.ROWS_FOLDED, max_iter=5000)
    st_o, u_o, _ = O.solve(s, opt)
    st, u, _ = ctx.solve(s, opt)
    assert st == 0 and rel(u, u_o) < 1e-10


def test_plan_reuse_matches_one_shot(ctx):
    s = varcoef3d(17, seed=5)
    opt = mgpcg(tol=1e-10)
    plan = ctx.plan(s, opt)
    rng = np.random.default_rng(9)
    for _ in range(2):
        f = rng.standard_normal(s.shape)
        plan.set_source(f)
        plan.solve()
        u_plan = plan.download()
        s.f_field = f
        _, u_one, _ = ctx.solve(s, opt)
        assert np.array_equal(u_plan, u_one)
    plan.close()


def test_server_results_identical_alone_and_in_mix(ctx):
    from pde_solver_amd import hip
    specs = []
    rng = np.random.default_rng(11)
    for n in (33, 65, 129, 33, 65):
        s = capi.Spec(n, n, a=-1.0, b=-1.0, e=float(rng.uniform(0, 10)))
        s.f_field = rng.standard_normal(s.shape)
        specs.append(s)
    opt = mgpcg(tol=1e-10)
    one_shot = [ctx.solve(s, opt)[1] for s in specs]
    cfg = capi.ServerConfig()
    cfg.max_inflight, cfg.streams, cfg.cache_operators = 8, 4, 1
    srv = ctx.server(cfg)
    alone = [srv.wait(srv.submit(s, opt))[1].copy() for s in specs]
    tickets = [srv.submit(s, opt) for s in specs]
    outs = [srv.wait(t) for t in tickets]
    assert srv.stats()["batched"] == 2 * len(specs)   # 2D MG-PCG requests run in batch lanes
    for (st, u, rep), a, o in zip(outs, alone, one_shot):
        assert st == 0
        assert np.array_equal(u, a)
        assert rel(u, o) < 1e-8
    srv.close()


def test_batch_matches_individual(ctx):
    specs = []
    rng = np.random.default_rng(12)
    for n in (33, 65, 33, 129):
        s = capi.Spec(n, n, a=-1.0, b=-1.0, e=float(rng.uniform(0, 50)))
        s.f_field = rng.standard_normal(s.shape)
        s.bcs = [BC.dirichlet(0.0), BC.neumann(0.0), BC.robin(1, 1, 0), BC.dirichlet(0.0), BC(), BC()]
        s.bc_convention = capi.BCCONV_OUTWARD
        specs.append(s)
    opt = mgpcg(tol=1e-10)
    st, outs, reps = ctx.solve_batch(specs, opt)
    assert st == 0
    for s, u in zip(specs, outs):
        _, u1, _ = O.solve(s, mgpcg(tol=1e-12))
        assert rel(u, u1) < 1e-8


def test_mg_linear_solution_robin_neumann(ctx):
    """u = 1 + x solves -u_xx - u_yy = 0 with Robin (u + du/dn = 0) at x=0, du/dx = 1 at x=1,
    Dirichlet 1 + x on y faces: the iterated V-cycle recovers it to rounding and tracks the
    oracle's iteration count."""
    s = capi.Spec(33, 33, a=-1.0, b=-1.0)
    s.bcs = [BC.robin(1, 1, 0), BC.neumann(1.0), BC.dirichlet(1.0, 1.0), BC.dirichlet(1.0, 1.0),
             BC(), BC()]
    s.bc_convention = capi.BCCONV_OUTWARD
    opt = capi.options(capi.MG, tol=1e-10, max_iter=200, bc_rows=capi.ROWS_FOLDED,
                       precond=capi.PC_MG)
    st_o, u_o, rep_o = O.solve(s, opt)
    st, u, rep = ctx.solve(s, opt)
    print("hip", st, rep.iterations, rep.final_metric, "oracle", rep_o.iterations, rep_o.final_metric)
    assert st == 0 and st_o == 0, ctx.error()
    x = np.linspace(0.0, 1.0, 33)
    assert np.abs(u - (1.0 + x)[None, :]).max() < 1e-11
    assert abs(rep.iterations - rep_o.iterations) <= 1


@pytest.mark.parametrize("shape", [(65, 9, 17), (9, 33, 5), (33, 17, 9)])
def test_mgpcg_fp64_nonsquare3d_matches_oracle(ctx, shape):
    """Flattened boxes: coarse levels whose interior box (all but two nodes per face) is thin or
    empty, so the restriction runs on its face-rim blocks alone."""
    nx, ny, nz = shape
    print(f"mgpcg fp64 nonsquare 3D {shape}", flush=True)
    s = capi.Spec(nx, ny, nz, a=-1.0, op_kind=capi.OP_FLUX)
    rng = np.random.default_rng(nx + 3 * ny + 7 * nz)
    s.kappa = np.exp(0.3 * rng.standard_normal(s.shape))
    s.f_field = rng.standard_normal(s.shape)
    s.bcs = [BC.dirichlet(0.0), BC.robin(1.0, 1.0, 0.0), BC.dirichlet(0.1), BC.neumann(0.0),
             BC.robin(1.0, 2.0, 0.0), BC.dirichlet(0.0)]
    s.bc_convention = capi.BCCONV_OUTWARD
    st_o, u_o, rep_o = O.solve(s, mgpcg())
    assert st_o == 0
    for precision, tol in ((capi.FP64, 1e-10), (capi.MIXED, 1e-8)):
        st, u, rep = ctx.solve(s, mgpcg(tol=1e-12 if precision == capi.FP64 else 1e-10, precision=precision))
        print(f"  precision {precision} its {rep.iterations} (oracle {rep_o.iterations}) rel {rel(u, u_o):.2e}",
              flush=True)
        assert st == 0
        assert rel(u, u_o) < tol


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
def test_flux_two_layer_conductor_harmonic_face(ctx, dim, precision):
    """FLUX faces are harmonic means of their two nodes (FP64 operator, and the FP32 reciprocal
    form of the mixed path): a two-layer conductor is reproduced at every node."""
    from tests.closed_forms import two_layer_conductor
    s, exact = two_layer_conductor(21, 8, 1.0, 10.0, ny=9, nz=9 if dim == 3 else 1)
    st, u, rep = ctx.solve(s, mgpcg(tol=1e-12 if precision == capi.FP64 else 1e-10,
                                    precision=precision))
    err = np.abs(u - exact).max()
    print(f"  dim {dim} precision {precision} its {rep.iterations} max err {err:.2e}", flush=True)
    assert st == 0
    assert err < (1e-10 if precision == capi.FP64 else 1e-8)


@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
def test_gmres_long_restart_matches_oracle(ctx, precision):
    """GMRES with a restart length above 64 (the Hessenberg staging is sized by the restart)."""
    s = capi.Spec(33, 33, a=-0.05, b=-0.05, c=0.7, d=0.4, convection=capi.CONV_UPWIND)
    s.f_field = np.random.default_rng(21).standard_normal(s.shape)
    opt = capi.options(capi.GMRES, tol=1e-10, tol_mode=capi.TOL_REL, precond=capi.PC_JACOBI,
                       bc_rows=capi.ROWS_FOLDED, gmres_restart=80, max_iter=2000,
                       precision=precision)
    st_o, u_o, rep_o = O.solve(s, capi.options(capi.GMRES, tol=1e-12, tol_mode=capi.TOL_REL,
                                               precond=capi.PC_JACOBI, bc_rows=capi.ROWS_FOLDED,
                                               gmres_restart=80, max_iter=2000))
    st, u, rep = ctx.solve(s, opt)
    print(f"  precision {precision} its {rep.iterations} (oracle {rep_o.iterations}) "
          f"rel {rel(u, u_o):.2e}", flush=True)
    assert st == 0 and st_o == 0
    assert rel(u, u_o) < (1e-8 if precision == capi.FP64 else 1e-7)
    if precision == capi.FP64:
        _, _, rep_same = O.solve(s, opt)   # the same tolerance: the same iteration count
        assert abs(rep.iterations - rep_same.iterations) <= 2


def test_plan_set_source_reference_rows_keep_dirichlet_data(ctx):
    """A reference-row BiCGStab plan (x0 = 0) with non-zero Dirichlet data: every new source
    rebuilds b with the boundary datum, and the plan's answers equal one-shot solves bitwise."""
    s = capi.Spec(33, 33, a=-1.0, b=-1.0, c=0.5)
    s.bcs = [BC.dirichlet(1.0, 0.5), BC.dirichlet(0.2), BC.dirichlet(0.0, 0.0, 1.0),
             BC.dirichlet(-0.3)] + [BC(), BC()]
    opt = capi.options(capi.BICGSTAB, tol=1e-11, max_iter=5000)
    rng = np.random.default_rng(31)
    s.f_field = rng.standard_normal(s.shape)
    plan = ctx.plan(s, opt)
    for _ in range(2):
        f = rng.standard_normal(s.shape)
        plan.set_source(f)
        plan.solve()
        u_plan = plan.download()
        s.f_field = f
        _, u_one, _ = ctx.solve(s, opt)
        _, u_o, _ = O.solve(s, opt)
        assert np.array_equal(u_plan, u_one)
        assert rel(u_plan, u_o) < 1e-9
    plan.close()
