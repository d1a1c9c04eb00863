"""The request server (pdehip_server_*) on the serving workload's request kinds at small sizes:
2D and 3D grids; Poisson, flux-form diffusion with a lognormal conductivity and upwind
convection-diffusion; MG-PCG, BiCGStab and GMRES(30); FP64 and mixed precision.

Every result is checked against the CPU oracle (FP64 solve to rel-resid 1e-13) and bit for bit
against the same request served alone.  Admission rejection, the deadline status, cache hits
that differ only in the source and relaxation requests that differ only in their nonlinear
terms are covered too."""
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


def smooth_field(shape, rng, passes=3):
    g = rng.standard_normal(shape)
    for _ in range(passes):
        for ax in range(len(shape)):
            g = (np.roll(g, 1, ax) + 2 * g + np.roll(g, -1, ax)) / 4
    return (g - g.mean()) / g.std()


def request(kind, dim, n, method, precision, seed):
    rng = np.random.default_rng(seed)
    shape3 = (n, n, n) if dim == 3 else (n, n, 1)
    if kind == "poisson":
        s = capi.Spec(*shape3, a=-1.0, b=-1.0, az=-1.0 if dim == 3 else 0.0)
    elif kind == "varcoef":
        s = capi.Spec(*shape3, a=-1.0, op_kind=capi.OP_FLUX)
        s.kappa = np.exp(0.5 * smooth_field(s.shape, rng))
    else:   # upwind convection-diffusion
        eps = 10 ** rng.uniform(-2, -1)
        th = rng.uniform(0, 2 * np.pi)
        s = capi.Spec(*shape3, a=-eps, b=-eps, az=-eps if dim == 3 else 0.0, c=np.cos(th),
                      d=np.sin(th), dz=0.3 if dim == 3 else 0.0, convection=capi.CONV_UPWIND)
    s.f_field = smooth_field(s.shape, rng, passes=1)
    if kind == "convdiff":
        s.bcs = [BC.dirichlet(0.0)] * 6
    else:
        s.bcs = [BC.dirichlet(0.0)] * 4 + [BC.robin(1.0, 1.0, 0.0)] * 2 if dim == 3 else \
            [BC.dirichlet(0.0), BC.neumann(0.0), BC.robin(1.0, 1.0, 0.0), BC.dirichlet(0.1), BC(), BC()]
    s.bc_convention = capi.BCCONV_OUTWARD
    pc = capi.PC_MG if method == capi.MGPCG else capi.PC_JACOBI
    opt = capi.options(method, tol=1e-10, tol_mode=capi.TOL_REL, precond=pc,
                       bc_rows=capi.ROWS_FOLDED, precision=precision, gmres_restart=30,
                       max_iter=200 if method == capi.MGPCG else 20000)
    return s, opt


MIX = [
    ("poisson", 2, 33, capi.MGPCG, capi.FP64),
    ("poisson", 2, 65, capi.MGPCG, capi.MIXED),
    ("poisson", 3, 17, capi.MGPCG, capi.MIXED),
    ("poisson", 3, 33, capi.MGPCG, capi.FP64),
    ("varcoef", 2, 65, capi.MGPCG, capi.FP64),
    ("varcoef", 2, 33, capi.MGPCG, capi.MIXED),
    ("varcoef", 3, 33, capi.MGPCG, capi.MIXED),
    ("varcoef", 3, 17, capi.MGPCG, capi.FP64),
    ("convdiff", 2, 33, capi.BICGSTAB, capi.FP64),
    ("convdiff", 2, 65, capi.BICGSTAB, capi.MIXED),
    ("convdiff", 2, 65, capi.GMRES, capi.FP64),
    ("convdiff", 2, 33, capi.GMRES, capi.MIXED),
    ("convdiff", 3, 17, capi.BICGSTAB, capi.FP64),
    ("convdiff", 3, 17, capi.GMRES, capi.MIXED),
    ("varcoef", 2, 17, capi.BICGSTAB, capi.FP64),
    ("poisson", 2, 17, capi.GMRES, capi.FP64),
]


def server(ctx, **kw):
    cfg = capi.ServerConfig()
    cfg.max_inflight, cfg.streams, cfg.cache_operators = 8, 4, 1
    for k, v in kw.items():
        setattr(cfg, k, v)
    return ctx.server(cfg)


def test_server_mix_matches_oracle_and_alone(ctx):
    reqs = [request(*m, seed=100 + i) for i, m in enumerate(MIX)]
    # alone: each request served by itself (nothing else on the device)
    srv = server(ctx)
    alone = []
    for i, (s, o) in enumerate(reqs):
        print(f"alone {i} {MIX[i]}", flush=True)
        st, u, rep = srv.wait(srv.submit(s, o))
        assert st == 0, (i, MIX[i], st)
        alone.append(u.copy())
    srv.close()
    # in the mix: every request submitted at once to a fresh server
    srv = server(ctx)
    print("mix: submitting all", flush=True)
    tickets = [srv.submit(s, o) for s, o in reqs]
    mixed = [srv.wait(t) for t in tickets]
    stats = srv.stats()
    srv.close()
    assert stats["completed"] == len(reqs) and stats["rejected"] == 0
    # the 2D CONST MG-PCG, FP64 BiCGStab and FP64 GMRES requests ran in batch lanes, the rest on
    # plan workers
    n_lane = sum(1 for m in MIX if m[1] == 2 and m[0] != "varcoef" and
                 (m[3] == capi.MGPCG or (m[3] in (capi.BICGSTAB, capi.GMRES) and m[4] == capi.FP64)))
    print(f"  batched {stats['batched']} of {len(reqs)} in {stats['batches']} lane batches", flush=True)
    assert stats["batched"] == n_lane > 0
    for i, ((s, o), (st, u, rep)) in enumerate(zip(reqs, mixed)):
        o_ref = capi.options(o.method, tol=1e-13, tol_mode=capi.TOL_REL, precond=o.precond,
                             bc_rows=capi.ROWS_FOLDED, gmres_restart=30, max_iter=o.max_iter * 2)
        st_o, u_o, _ = O.solve(s, o_ref)
        err = rel(u, u_o)
        print(f"  {i} {MIX[i]} status {st} its {rep.iterations} rel vs oracle {err:.2e} "
              f"identical alone {np.array_equal(u, alone[i])}", flush=True)
        assert st == 0 and st_o == 0
        assert err <= (1e-8 if o.precision == capi.FP64 else 1e-7)
        assert np.array_equal(u, alone[i])


def test_server_admission_rejects_request_larger_than_budget(ctx):
    from pde_solver_amd import hip
    srv = server(ctx, hbm_budget_bytes=1 << 20)
    s, o = request("poisson", 2, 65, capi.MGPCG, capi.FP64, seed=1)
    with pytest.raises(hip.HipError) as e:
        srv.submit(s, o)
    assert e.value.status == capi.ERR_ADMISSION
    assert srv.stats()["rejected"] == 1
    srv.close()


def test_server_deadline_reports_timeout(ctx):
    srv = server(ctx, timeout_ms=1e-3)
    s, o = request("varcoef", 3, 33, capi.MGPCG, capi.FP64, seed=2)
    st, _, rep = srv.wait(srv.submit(s, o))
    assert st == capi.ERR_TIMEOUT
    assert srv.stats()["timed_out"] == 1
    srv.close()


def slow_poisson(n, method, dim=2):
    """Poisson without a preconditioner to rel-resid 1e-10: hundreds (BiCGStab) to thousands
    (GMRES(30)) of Krylov iterations, so a deadline of a fifth of the solve lands inside it."""
    if dim == 3:
        s = capi.Spec(n, n, n, a=-1.0, b=-1.0, az=-1.0)
        s.bcs = [BC.dirichlet(0.0)] * 6
    else:
        s = capi.Spec(n, n, a=-1.0, b=-1.0)
        s.bcs = [BC.dirichlet(0.0)] * 4 + [BC(), BC()]
    s.f_field = np.random.default_rng(n).standard_normal(s.shape)
    return s, capi.options(method, tol=1e-10, tol_mode=capi.TOL_REL, precond=capi.PC_NONE,
                           gmres_restart=30, max_iter=20000)


@pytest.mark.parametrize("method,dim", [(capi.BICGSTAB, 2), (capi.GMRES, 2), (capi.GMRES, 3)])
def test_server_deadline_stops_running_solve(ctx, method, dim):
    """A deadline that passes while the solve runs ends it on the device: a batch lane marks the
    member broken down between polling chunks (2D BiCGStab and GMRES), a plan worker's watchdog
    raises the solve's cancel flag (3D GMRES).  The request ends with PDEHIP_ERR_TIMEOUT after
    far fewer iterations than the full solve, and the server keeps serving."""
    import time
    s, o = slow_poisson(129 if dim == 2 else 33, method, dim)
    srv = server(ctx)
    t = time.perf_counter()
    st_full, _, rep_full = srv.wait(srv.submit(s, o))
    full_ms = (time.perf_counter() - t) * 1e3
    srv.close()
    srv = server(ctx, timeout_ms=max(5.0, 0.2 * full_ms))
    t = time.perf_counter()
    st, _, rep = srv.wait(srv.submit(s, o))
    ms = (time.perf_counter() - t) * 1e3
    print(f"  method {method}: full {rep_full.iterations} its {full_ms:.0f} ms; with deadline status "
          f"{st} after {rep.iterations} its {ms:.0f} ms", flush=True)
    assert st_full == 0
    assert st == capi.ERR_TIMEOUT
    assert ms < 0.7 * full_ms
    assert rep.iterations < rep_full.iterations
    stats = srv.stats()
    assert stats["timed_out"] == 1
    assert stats["batched"] == (1 if dim == 2 else 0)
    # the server still serves after a cancelled solve
    s2, o2 = request("poisson", 2, 33, capi.MGPCG, capi.FP64, seed=3)
    srv.close()
    srv = server(ctx)
    st2, _, _ = srv.wait(srv.submit(s2, o2))
    assert st2 == 0
    srv.close()


def test_server_lane_batches_compatible_arrivals(ctx):
    """Requests queued together with the same options share one lane batch; a request with other
    options forms its own; each result equals the request served alone."""
    specs = [request("poisson", 2, n, capi.MGPCG, capi.MIXED, seed=40 + i)
             for i, n in enumerate((33, 65, 129, 65, 33, 129))]
    other = request("poisson", 2, 65, capi.MGPCG, capi.FP64, seed=50)
    srv = server(ctx, batch_lanes=1)
    alone = [srv.wait(srv.submit(s, o))[1].copy() for s, o in specs + [other]]
    before = srv.stats()["batches"]
    tickets = [srv.submit(s, o) for s, o in specs + [other]]
    outs = [srv.wait(t) for t in tickets]
    stats = srv.stats()
    srv.close()
    print(f"  lane batches for the mix: {stats['batches'] - before}", flush=True)
    assert stats["batches"] - before <= 4   # far fewer batches than requests
    for (st, u, _), a in zip(outs, alone):
        assert st == 0 and np.array_equal(u, a)


def test_server_cache_hit_with_only_constant_source_changed(ctx):
    """Two requests with the same operator and constant sources f = 1 and f = -3: the second
    hits the resident operator and must use its own f (plan workers only: bitwise against the
    one-shot solve)."""
    srv = server(ctx, streams=1, batch_lanes=-1)
    outs, alone = [], []
    for f in (1.0, -3.0):
        s = capi.Spec(33, 33, a=-1.0, b=-1.0)
        s.f = f
        s.bcs = [BC.dirichlet(0.0)] * 4 + [BC(), BC()]
        o = capi.options(capi.MGPCG, tol=1e-10, tol_mode=capi.TOL_REL, precond=capi.PC_MG,
                         bc_rows=capi.ROWS_FOLDED)
        outs.append(srv.wait(srv.submit(s, o))[1].copy())
        alone.append(ctx.solve(s, o)[1])
    assert srv.stats()["cache_hits"] == 1
    srv.close()
    for u, a in zip(outs, alone):
        assert np.array_equal(u, a)
    assert rel(outs[1], -3.0 * outs[0]) < 1e-9


def test_server_lane_cache_hit_with_only_constant_source_changed(ctx):
    """The same through one batch lane: the second request reuses the lane's resident operator
    with its own constant source (within the solve tolerance of the one-shot solves)."""
    srv = server(ctx, batch_lanes=1)
    outs = []
    for f in (1.0, -3.0):
        s = capi.Spec(65, 65, a=-1.0, b=-1.0)
        s.f = f
        s.bcs = [BC.dirichlet(0.0)] * 4 + [BC(), BC()]
        o = capi.options(capi.MGPCG, tol=1e-10, tol_mode=capi.TOL_REL, precond=capi.PC_MG,
                         bc_rows=capi.ROWS_FOLDED)
        outs.append((srv.wait(srv.submit(s, o))[1].copy(), ctx.solve(s, o)[1]))
    assert srv.stats()["cache_hits"] == 1 and srv.stats()["batched"] == 2
    srv.close()
    for u, a in outs:
        assert rel(u, a) < 1e-8
    assert rel(outs[1][0], -3.0 * outs[0][0]) < 1e-8


def test_server_nonlinear_terms_are_part_of_the_operator(ctx):
    """Jacobi requests that differ only in their nonlinear terms do not share a resident plan."""
    srv = server(ctx, streams=1)
    res = []
    for nl in ([(capi.NL_POWER, 0.5, 3)], [(capi.NL_SIN, 0.8, 0)]):
        s = capi.Spec(17, 17, a=-1.0, b=-1.0)
        X, Y, _ = s.coords()
        s.f_field = -2 * np.pi ** 2 * np.sin(np.pi * X) * np.sin(np.pi * Y)
        s.nonlinear = nl
        o = capi.options(capi.JACOBI, tol=1e-11, max_iter=20000)
        u = srv.wait(srv.submit(s, o))[1].copy()
        _, u_o, _ = O.solve(s, o)
        res.append((u, u_o))
    assert srv.stats()["cache_hits"] == 0
    srv.close()
    for u, u_o in res:
        assert rel(u, u_o) < 1e-10
