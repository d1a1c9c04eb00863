"""The fused out-of-place red-black sweep (one launch per sweep) against the two-launch in-place
sweep: every multigrid solve must come out bitwise identical, whatever the launch shape, on
grids that span several tiles and z chunks in every direction (non-cubic, so a swapped axis
shows), for FP64 and for the mixed FP32/FP16 hierarchy.  The box smoothing of the coarse mixed
levels is held to the same bar against the plane-march sweeps."""
import numpy as np
import pytest

from pde_solver_amd import capi
from pde_solver_amd.capi import BC

pytestmark = pytest.mark.gpu

DEFAULTS = {"smooth.fused": 1, "smooth.rows": 2, "smooth.zchunk": 0, "smooth.ychunk": 64,
            "smooth.pair": 0, "smooth.ring": 1, "smooth.reg": 1, "stencil.ring": 1, "smooth.box": 65,
            "smooth.box2": 33, "smooth.box_z": 4, "smooth.box_rr": 65, "smooth.box_small": 33,
            "smooth.box_sxy": 4, "mixed.fp16": 1,
            "mixed.replace": 1, "mixed.xpay": 1, "restrict.march": 1}


@pytest.fixture(scope="module")
def ctx():
    from pde_solver_amd import hip
    c = hip.Context()
    yield c
    for k, v in DEFAULTS.items():
        hip.tune_set(k, v)
    c.close()


def flux_problem(shape, seed):
    nz, ny, nx = shape if len(shape) == 3 else (1,) + tuple(shape)
    s = capi.Spec(nx, ny, nz, a=-1.0, op_kind=capi.OP_FLUX)
    rng = np.random.default_rng(seed)
    s.kappa = np.exp(0.5 * rng.standard_normal(s.shape))
    s.f_field = rng.standard_normal(s.shape)
    s.bcs = [BC.dirichlet(0.3), BC.robin(1.0, 0.5, 0.1), BC.neumann(0.2), BC.dirichlet(0.0),
             BC.robin(1.0, 1.0, 0.0), BC.robin(1.0, 1.0, 0.0)]
    s.bc_convention = capi.BCCONV_OUTWARD
    return s


def const_problem(shape, seed):
    nz, ny, nx = shape
    s = capi.Spec(nx, ny, nz, a=-1.0, b=-2.0, az=-0.5, e=0.3)
    s.f_field = np.random.default_rng(seed).standard_normal(s.shape)
    s.bcs = [BC.dirichlet(0.1)] * 2 + [BC.neumann(0.0)] * 2 + [BC.robin(1.0, 1.0, 0.0)] * 2
    s.bc_convention = capi.BCCONV_OUTWARD
    return s


def solve(ctx, s, precision, fused, pre=2, post=2, max_iter=100, **tune):
    from pde_solver_amd import hip
    for k, v in DEFAULTS.items():
        hip.tune_set(k, v)
    hip.tune_set("smooth.fused", fused)
    for k, v in tune.items():
        hip.tune_set(k if "." in k else "smooth." + k, v)
    opt = capi.options(capi.MGPCG, tol=1e-10, tol_mode=capi.TOL_REL, precond=capi.PC_MG,
                       bc_rows=capi.ROWS_FOLDED, precision=precision, max_iter=max_iter,
                       mg_pre=pre, mg_post=post)
    st, u, rep = ctx.solve(s, opt, check=False)
    assert st == (0 if max_iter >= 100 else capi.ERR_NOT_CONVERGED), ctx.error()
    return u, rep


CASES_3D = [
    ("flux", (33, 65, 129), {"rows": 2, "zchunk": 32}),
    ("flux", (33, 65, 129), {"rows": 4, "zchunk": 4}),
    ("flux", (65, 33, 17), {"rows": 2, "zchunk": 8}),
    ("const", (17, 129, 65), {"rows": 4, "zchunk": 16}),
    ("flux", (33, 65, 129), {"rows": 2, "zchunk": 0, "pair": 1}),
    ("flux", (65, 33, 257), {"rows": 2, "zchunk": 8, "pair": 1}),
    ("const", (17, 129, 65), {"rows": 2, "zchunk": 0, "pair": 1}),
    # FP32/FP16 levels at rows 2 run the ring sweep (k_rbgs_ring3) with reg = 0: automatic
    # chunks, a chunk whose step count is not a multiple of the 6-step unroll (dead steps), and
    # the plane-march sweep it replaces
    ("flux", (33, 65, 129), {"rows": 2, "zchunk": 0, "ring": 1, "reg": 0}),
    ("flux", (65, 33, 257), {"rows": 2, "zchunk": 9, "ring": 1, "reg": 0}),
    ("const", (17, 129, 65), {"rows": 2, "zchunk": 5, "ring": 1, "reg": 0}),
    ("flux", (33, 65, 129), {"rows": 2, "zchunk": 0, "ring": 0}),
    ("flux", (37, 37, 37), {"rows": 2, "zchunk": 0, "ring": 1, "reg": 0}),
    # the register-row sweep (k_rbgs_reg3, the default): automatic chunks, dead steps, short
    # chunks, sides 4m+1 (1/h^2 not a power of two)
    ("flux", (33, 65, 129), {"rows": 2, "zchunk": 0, "reg": 1}),
    ("flux", (65, 33, 257), {"rows": 2, "zchunk": 9, "reg": 1}),
    ("const", (17, 129, 65), {"rows": 2, "zchunk": 5, "reg": 1}),
    ("flux", (69, 61, 37), {"rows": 2, "zchunk": 0, "reg": 1}),
]


@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
@pytest.mark.parametrize("kind,shape,tune", CASES_3D)
def test_fused_sweep_bitwise_3d(ctx, kind, shape, tune, precision):
    print(f"fused 3D {kind} {shape} {tune} precision={precision}", flush=True)
    s = (flux_problem if kind == "flux" else const_problem)(shape, seed=sum(shape))
    u0, r0 = solve(ctx, s, precision, 0, **tune)
    u1, r1 = solve(ctx, s, precision, 1, **tune)
    print(f"  iterations {r0.iterations}/{r1.iterations}  max diff {np.abs(u0 - u1).max():.3e}", flush=True)
    assert r0.iterations == r1.iterations
    assert np.array_equal(u0, u1)


@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
@pytest.mark.parametrize("kind,shape", [
    ("flux", (33, 65, 129)),    # several march tiles in x and y, z chunks
    ("flux", (65, 33, 17)),     # coarse levels whose box is one or two columns wide
    ("const", (17, 129, 65)),
    ("flux", (129, 65, 65)),    # long z: many chunks
    ("flux", (69, 61, 37)),     # sides 4m+1, odd coarse levels
])
def test_restrict_march_bitwise(ctx, kind, shape, precision):
    """The plane-march restriction (k_restrict3m) against the per-node one (k_restrict3): the
    same coarse right-hand sides, so the same multigrid solve bit for bit, FP64 and mixed, with
    the FP32/FP16 levels on plane marches (box residual-restriction off)."""
    print(f"restrict march {kind} {shape} precision={precision}", flush=True)
    s = (flux_problem if kind == "flux" else const_problem)(shape, seed=sum(shape) + 7)
    u0, r0 = solve(ctx, s, precision, 1, **{"restrict.march": 0, "box_rr": 0})
    u1, r1 = solve(ctx, s, precision, 1, **{"restrict.march": 1, "box_rr": 0})
    print(f"  iterations {r0.iterations}/{r1.iterations}  max diff {np.abs(u0 - u1).max():.3e}", flush=True)
    assert r0.iterations == r1.iterations
    assert np.array_equal(u0, u1)


@pytest.mark.parametrize("kind,shape", [("flux", (33, 65, 129)), ("flux", (65, 33, 17)),
                                        ("const", (17, 129, 65))])
def test_plane_ring_bitwise(ctx, kind, shape):
    """The ring-buffered FP32/FP16 plane march (apply + residual of every mixed 3D level)
    against k_plane3: the same outputs and partials, so the same mixed solve bit for bit."""
    print(f"plane ring {kind} {shape}", flush=True)
    s = (flux_problem if kind == "flux" else const_problem)(shape, seed=sum(shape) + 1)
    u0, r0 = solve(ctx, s, capi.MIXED, 1, **{"stencil.ring": 0})
    u1, r1 = solve(ctx, s, capi.MIXED, 1, **{"stencil.ring": 1})
    print(f"  iterations {r0.iterations}/{r1.iterations}  max diff {np.abs(u0 - u1).max():.3e}", flush=True)
    assert r0.iterations == r1.iterations
    assert np.array_equal(u0, u1)


BOX_CASES = [
    # (kind, (nz, ny, nx), box_z, FP16 coarse levels, pre, post, box2)
    ("flux", (33, 65, 129), 8, 1, 2, 2, 65),
    ("flux", (65, 33, 17), 4, 1, 2, 2, 65),
    ("flux", (129, 65, 33), 8, 0, 2, 2, 65),
    ("const", (17, 129, 65), 8, 1, 2, 2, 65),
    ("const", (65, 65, 65), 4, 0, 2, 2, 65),
    ("flux", (33, 65, 129), 8, 1, 1, 3, 65),
    ("flux", (65, 33, 129), 4, 1, 3, 1, 65),
    # sides 4m+1 whose third level is even and whose 1/h^2 is not a power of two
    ("flux", (69, 61, 37), 8, 1, 2, 2, 65),
    ("flux", (65, 129, 129), 8, 1, 2, 2, 33),
    ("const", (129, 65, 129), 4, 0, 3, 2, 0),
]


@pytest.mark.parametrize("kind,shape,bz,fp16,pre,post,box2", BOX_CASES)
def test_box_smoothing_bitwise(ctx, kind, shape, bz, fp16, pre, post, box2):
    """Coarse levels of the mixed V-cycle on LDS boxes (two sweeps per launch, the
    prolongation folded into the first post-smoothing launch) against the plane-march sweeps
    and the prolongation kernel: the same mixed solve bit for bit, on grids whose coarse levels
    span several boxes in every direction and end in partial boxes, FP16 and FP32 coarse
    storage, odd sweep counts and levels with one sweep per launch (buffers alternating)."""
    print(f"box {kind} {shape} box_z={bz} fp16={fp16} V({pre},{post}) box2={box2}", flush=True)
    s = (flux_problem if kind == "flux" else const_problem)(shape, seed=sum(shape) + 7)
    u0, r0 = solve(ctx, s, capi.MIXED, 1, pre, post, **{"smooth.box": 0, "mixed.fp16": fp16})
    u1, r1 = solve(ctx, s, capi.MIXED, 1, pre, post,
                   **{"smooth.box": 129, "smooth.box2": box2, "smooth.box_z": bz,
                      "mixed.fp16": fp16})
    print(f"  iterations {r0.iterations}/{r1.iterations}  max diff {np.abs(u0 - u1).max():.3e}", flush=True)
    assert r0.iterations == r1.iterations
    assert np.array_equal(u0, u1)


SMALL_BOX_CASES = [
    # (kind, (nz, ny, nx), FP16 coarse levels, y extent of the small boxes)
    ("flux", (33, 65, 129), 1, 8),
    ("flux", (65, 33, 17), 1, 4),
    ("const", (65, 65, 65), 0, 4),
    ("flux", (69, 61, 37), 1, 8),
]


@pytest.mark.parametrize("kind,shape,fp16,sxy", SMALL_BOX_CASES)
def test_small_box_bitwise(ctx, kind, shape, fp16, sxy):
    """Small boxes (8 x sxy x 4 sweep boxes, 4 x 4 x 4 coarse residual-restriction boxes) on
    every coarse level against the plane-march sweeps, residual and restriction: the same
    mixed solve bit for bit."""
    print(f"small box {kind} {shape} fp16={fp16} sxy={sxy}", flush=True)
    s = (flux_problem if kind == "flux" else const_problem)(shape, seed=sum(shape) + 11)
    u0, r0 = solve(ctx, s, capi.MIXED, 1, **{"smooth.box": 0, "smooth.box_rr": 0,
                                             "mixed.fp16": fp16})
    u1, r1 = solve(ctx, s, capi.MIXED, 1, **{"smooth.box": 129, "smooth.box2": 129,
                                             "smooth.box_z": 4, "smooth.box_rr": 129,
                                             "smooth.box_small": 129, "smooth.box_sxy": sxy,
                                             "mixed.fp16": fp16})
    print(f"  iterations {r0.iterations}/{r1.iterations}  max diff {np.abs(u0 - u1).max():.3e}", flush=True)
    assert r0.iterations == r1.iterations
    assert np.array_equal(u0, u1)


RR_CASES = [
    # (kind, (nz, ny, nx), FP16 coarse levels, largest fine side fused)
    ("flux", (33, 65, 129), 1, 129),
    ("flux", (65, 129, 129), 0, 129),
    ("const", (129, 65, 33), 1, 129),
    ("flux", (69, 61, 37), 1, 129),
    ("flux", (65, 65, 257), 1, 257),
    ("const", (17, 33, 65), 0, 257),
]


@pytest.mark.parametrize("kind,shape,fp16,brr", RR_CASES)
def test_box_resrestrict_bitwise(ctx, kind, shape, fp16, brr):
    """The residual and its restriction in one launch (the fine residual formed in LDS)
    against the residual and restriction kernels: the same mixed solve bit for bit, on grids
    whose coarse levels end in partial boxes, with folded faces on every axis, FP16 and FP32
    coarse storage, and the fine level's FP32 -> FP16 transfer (scaled by the V-cycle's input
    normalisation) when the bound admits it."""
    print(f"box rr {kind} {shape} fp16={fp16} bound={brr}", flush=True)
    s = (flux_problem if kind == "flux" else const_problem)(shape, seed=sum(shape) + 11)
    u0, r0 = solve(ctx, s, capi.MIXED, 1, **{"smooth.box_rr": 0, "mixed.fp16": fp16})
    u1, r1 = solve(ctx, s, capi.MIXED, 1, **{"smooth.box_rr": brr, "mixed.fp16": fp16})
    print(f"  iterations {r0.iterations}/{r1.iterations}  max diff {np.abs(u0 - u1).max():.3e}", flush=True)
    assert r0.iterations == r1.iterations
    assert np.array_equal(u0, u1)


@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
@pytest.mark.parametrize("shape,ychunk", [((129, 257), 64), ((257, 65), 16)])
def test_fused_sweep_bitwise_2d(ctx, shape, ychunk, precision):
    print(f"fused 2D {shape} ychunk={ychunk} precision={precision}", flush=True)
    s = flux_problem(shape, seed=shape[0])
    u0, r0 = solve(ctx, s, precision, 0, ychunk=ychunk)
    u1, r1 = solve(ctx, s, precision, 1, ychunk=ychunk)
    print(f"  iterations {r0.iterations}/{r1.iterations}  max diff {np.abs(u0 - u1).max():.3e}", flush=True)
    assert r0.iterations == r1.iterations
    assert np.array_equal(u0, u1)


@pytest.mark.parametrize("kind,shape", [("flux", (33, 65, 129)), ("flux", (65, 33, 17)),
                                        ("const", (17, 129, 65)), ("pointwise", (19, 23, 70))])
def test_plane_apply_3d_matches_oracle(ctx, kind, shape):
    """The plane-march 3D apply (multi-tile, several z chunks) against the CPU oracle's operator."""
    from oracle import oracle as O
    print(f"plane apply {kind} {shape}", flush=True)
    if kind == "flux":
        s = flux_problem(shape, seed=5)
    elif kind == "const":
        s = const_problem(shape, seed=5)
    else:
        nz, ny, nx = shape
        s = capi.Spec(nx, ny, nz, a=-1.0, b=-0.5, az=-2.0, c=0.3, e=0.1, op_kind=capi.OP_POINTWISE)
        rng = np.random.default_rng(6)
        s.coef = {capi.COEF_A: -1.0 - rng.random(s.shape), capi.COEF_E: rng.random(s.shape)}
    x = np.random.default_rng(9).standard_normal(s.shape)
    y = ctx.apply(s, x)
    y_o = O.apply(s, x)
    err = np.abs(y - y_o).max() / np.abs(y_o).max()
    print(f"  max rel diff {err:.2e}", flush=True)
    assert err < 1e-13


@pytest.mark.parametrize("key", ["mixed.replace", "mixed.xpay"])
@pytest.mark.parametrize("kind,shape,max_iter", [
    ("flux", (33, 65, 129), 100), ("flux", (69, 61, 37), 100), ("const", (17, 129, 65), 100),
    # stopped by max_iter: after a replacement (the FP32 correction void) and between two
    ("flux", (33, 65, 129), 3), ("flux", (33, 65, 129), 4), ("flux", (69, 61, 37), 5)])
def test_fused_replacement_bitwise(ctx, kind, shape, max_iter, key):
    """The mixed MG-PCG steps fused into plane marches against the launches they replace: the
    residual replacement in one launch (the FP32 correction folded into the other FP64 solution
    buffer, FP64 residual, scaled FP32 copy; key mixed.replace) and the direction update inside
    the operator apply (key mixed.xpay).  The same solution, iterations, replacements and final
    residual bit for bit, converged or stopped by max_iter."""
    print(f"fused {key} {kind} {shape} max_iter={max_iter}", flush=True)
    s = (flux_problem if kind == "flux" else const_problem)(shape, seed=sum(shape) + 7)
    u0, r0 = solve(ctx, s, capi.MIXED, 1, max_iter=max_iter, **{key: 0})
    u1, r1 = solve(ctx, s, capi.MIXED, 1, max_iter=max_iter, **{key: 1})
    print(f"  iterations {r0.iterations}/{r1.iterations} refinements {r0.refinements}/"
          f"{r1.refinements} max diff {np.abs(u0 - u1).max():.3e}", flush=True)
    assert (r0.iterations, r0.refinements) == (r1.iterations, r1.refinements)
    assert r0.final_metric == r1.final_metric
    assert np.array_equal(u0, u1)
