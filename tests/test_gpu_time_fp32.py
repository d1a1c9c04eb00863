"""Time series with an FP32 state (TimeConfig.precision = MIXED; stage arithmetic FP64) against
the FP64 state: reaction-diffusion u_t = D lap u + u(1 - u) with zero-Neumann faces under RK4
(the nonsymmetric workload's time-dependent case), heat with Dirichlet data under SSPRK3 and
forward Euler; frames widened to FP64 on the device."""
import numpy as np
import pytest

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


def rd_spec(n, seed=44):
    from scipy.ndimage import gaussian_filter
    s = capi.Spec(n, n, a=-1.0, b=-1.0, ut=1.0)
    s.nonlinear = [(capi.NL_POWER, -1.0, 1), (capi.NL_POWER, 1.0, 2)]   # N(u) = -u + u^2
    u0 = gaussian_filter(np.random.default_rng(seed).uniform(0, 1, (n, n)), 2.0, mode="reflect")
    s.initial = u0
    s.bcs = [BC.neumann(0.0)] * 4 + [BC(), BC()]
    return s


def test_rd_rk4_fp32_state_matches_fp64(ctx):
    n = 129
    s = rd_spec(n)
    h = 1.0 / (n - 1)
    frames = 41
    tc64 = capi.time_config(capi.TI_RK4, dt=0.2 * h * h, frames=frames, precision=capi.FP64)
    tc32 = capi.time_config(capi.TI_RK4, dt=0.2 * h * h, frames=frames, precision=capi.MIXED)
    f64, u64, _ = ctx.time_series(s, tc64)
    f32, u32, rep = ctx.time_series(s, tc32)
    assert rep.iterations == frames - 1
    assert f32.shape == f64.shape
    assert rel(u32, u64) < 1e-5
    assert rel(f32[frames // 2], f64[frames // 2]) < 1e-5
    _, u_nf, _ = ctx.time_series(s, tc32, keep_frames=False)
    assert np.array_equal(u_nf, u32)


@pytest.mark.parametrize("integ", [capi.TI_SSPRK3, capi.TI_FE])
def test_heat_dirichlet_fp32_state(ctx, integ):
    n = 65
    s = capi.Spec(n, n, a=-1.0, b=-1.0, ut=1.0)
    x = np.linspace(0, 1, n)
    s.initial = np.outer(np.sin(np.pi * x), np.sin(2 * np.pi * x))
    s.bcs = [BC.dirichlet(0.0), BC.dirichlet(0.0), BC.dirichlet(0.0), BC.dirichlet(0.5), BC(), BC()]
    h = 1.0 / (n - 1)
    tc64 = capi.time_config(integ, dt=0.2 * h * h, frames=30, precision=capi.FP64)
    tc32 = capi.time_config(integ, dt=0.2 * h * h, frames=30, precision=capi.MIXED)
    _, u64, _ = ctx.time_series(s, tc64, keep_frames=False)
    _, u32, _ = ctx.time_series(s, tc32, keep_frames=False)
    assert rel(u32, u64) < 1e-5
