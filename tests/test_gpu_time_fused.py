"""The one-launch RK4 step (four stages and the boundary closures temporally blocked in LDS)
against the per-stage launches, bit for bit, and against the oracle's RK4 time series
(SolvePDETimeSeries, reference solver.cpp:989-1329).

Grid sides are chosen so that a face column or row sits alone in its 54 x 22 tile (sides of
1 + a multiple of the tile edge: 55, 109, 163 across, 23, 45, 67, 89, 111 down), next to sides
that end mid-tile and a single-tile grid; the boundary mixes put Neumann and Robin closures
next to Dirichlet faces, so corner nodes see every closure order."""
import numpy as np
import pytest

from pde_solver_amd import capi
from pde_solver_amd.capi import BC
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from pde_solver_amd import hip
    c = hip.Context()
    yield c
    c.close()


def case(name, nx, ny, seed=5):
    rng = np.random.default_rng(seed)
    if name == "rd_neumann":
        s = capi.Spec(nx, ny, a=-1.0, b=-1.0, ut=1.0)
        s.nonlinear = [(capi.NL_POWER, -1.0, 1), (capi.NL_POWER, 1.0, 2)]
        s.bcs = [BC.neumann(0.0)] * 4 + [BC(), BC()]
    elif name == "mixed_axis":
        s = capi.Spec(nx, ny, a=-1.0, b=-0.7, c=0.3, d=-0.2, e=-0.1, ut=1.0, f=0.4)
        s.bcs = [BC.dirichlet(0.2, cx=0.0, cy=0.5), BC.robin(1.0, 0.5, 0.3), BC.neumann(0.3),
                 BC.dirichlet(-0.1), BC(), BC()]
        s.nonlinear = [(capi.NL_SIN, 0.2, 1)]
    elif name == "mixed_outward":
        s = capi.Spec(nx, ny, a=-0.8, b=-1.0, ut=2.0)
        s.bcs = [BC.robin(1.0, 1.0, 0.1), BC.neumann(-0.2), BC.dirichlet(0.3), BC.robin(2.0, 0.5, 0.0),
                 BC(), BC()]
        s.bc_convention = capi.BCCONV_OUTWARD
    elif name == "flux":
        s = capi.Spec(nx, ny, a=-1.0, e=-0.5, ut=1.0, op_kind=capi.OP_FLUX)
        s.kappa = np.exp(0.4 * rng.standard_normal((ny, nx)))
        s.f_field = rng.standard_normal((ny, nx))
        s.bcs = [BC.neumann(0.1), BC.dirichlet(0.0), BC.robin(1.0, 0.5, 0.2), BC.neumann(0.0), BC(), BC()]
    else:   # pointwise
        s = capi.Spec(nx, ny, op_kind=capi.OP_POINTWISE, c=0.4, d=-0.2, ut=1.0)
        s.coef = {capi.COEF_A: -1.0 - rng.random((ny, nx)), capi.COEF_B: -0.5 - rng.random((ny, nx)),
                  capi.COEF_E: -rng.random((ny, nx))}
        s.bcs = [BC.dirichlet(0.0), BC.neumann(0.2), BC.dirichlet(1.0), BC.robin(1.0, 1.0, 0.0), BC(), BC()]
    s.initial = rng.uniform(0.0, 1.0, (ny, nx))
    return s


def run(ctx, s, precision, fused, frames=6):
    from pde_solver_amd import hip
    hx, hy, _ = s.spacing()
    tc = capi.time_config(capi.TI_RK4, dt=0.2 * min(hx, hy) ** 2, frames=frames, precision=precision)
    hip.tune_set("time.fused", 1 if fused else 0)
    try:
        fr, u, rep = ctx.time_series(s, tc)
    finally:
        hip.tune_set("time.fused", 1)
    assert rep.iterations == frames - 1
    return fr, u, tc


CASES = [("rd_neumann", 65, 65), ("rd_neumann", 109, 45), ("mixed_axis", 55, 67),
         ("mixed_axis", 50, 37), ("mixed_outward", 163, 23), ("mixed_outward", 33, 70),
         ("flux", 109, 89), ("flux", 70, 41), ("pointwise", 55, 111), ("pointwise", 45, 66),
         ("mixed_axis", 20, 9)]


@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED], ids=["fp64", "fp32"])
@pytest.mark.parametrize("name,nx,ny", CASES, ids=[f"{c[0]}-{c[1]}x{c[2]}" for c in CASES])
def test_fused_step_bitwise_equal_to_stage_launches(ctx, name, nx, ny, precision):
    s = case(name, nx, ny)
    print(f"fused rk4 {name} {nx}x{ny} precision={precision}", flush=True)
    f1, u1, _ = run(ctx, s, precision, True)
    f0, u0, _ = run(ctx, s, precision, False)
    assert np.all(np.isfinite(f1))
    assert np.array_equal(f1, f0), float(np.abs(f1 - f0).max())
    assert np.array_equal(u1, u0)


@pytest.mark.parametrize("name,nx,ny", [("mixed_axis", 109, 67), ("flux", 55, 89),
                                        ("mixed_outward", 163, 45), ("rd_neumann", 109, 111),
                                        ("pointwise", 45, 66)])
def test_fused_step_matches_oracle(ctx, name, nx, ny):
    """rd_neumann takes the cubic (Horner) nonlinear path, mixed_axis the general one."""
    s = case(name, nx, ny)
    print(f"fused rk4 vs oracle {name} {nx}x{ny}", flush=True)
    fr, u, tc = run(ctx, s, capi.FP64, True)
    ref = O.time_series(s, tc)
    for f in range(ref.shape[0]):
        err = float(np.linalg.norm(fr[f] - ref[f]) / np.linalg.norm(ref[f]))
        assert err < 1e-12, (f, err)


def test_fused_knob_roundtrip():
    from pde_solver_amd import hip
    assert hip.tune_get("time.fused") == 1
    hip.tune_set("time.fused", 0)
    assert hip.tune_get("time.fused") == 0
    hip.tune_set("time.fused", 1)
    with pytest.raises(hip.HipError):
        hip.tune_set("time.fused", 2)
