"""The reference residual norms of 3D grids (ComputeResidualNorms, residual.cpp:153-583) as an
FP64 plane march (k_plane3 MODE 4, the default) against the row kernel it replaces
(k_resnorm3, with "stencil.plane" = 0) and against the oracle: the same interior rows and
unfolded weights, the neighbour sum in the same order, so the norms agree to the last few
bits (the compilers' contraction choices may differ between the two kernels)."""
import numpy as np
import pytest

from oracle import oracle as O
from pde_solver_amd import capi
from pde_solver_amd.capi import BC

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from pde_solver_amd import hip
    c = hip.Context()
    yield c
    hip.tune_set("stencil.plane", 1)
    c.close()


def problem(kind, shape, seed):
    nz, ny, nx = shape
    rng = np.random.default_rng(seed)
    if kind == "flux":
        s = capi.Spec(nx, ny, nz, a=-1.0, op_kind=capi.OP_FLUX)
        s.kappa = np.exp(0.5 * rng.standard_normal(s.shape))
    elif kind == "const":
        s = capi.Spec(nx, ny, nz, a=-1.0, b=-2.0, az=-0.5, c=0.7, d=-0.3, dz=0.2, e=0.3)
    else:   # per-node coefficients
        s = capi.Spec(nx, ny, nz, a=-1.0, b=-1.0, az=-1.0, e=0.1, op_kind=capi.OP_POINTWISE)
        s.coef[capi.COEF_A] = -1.0 - 0.5 * rng.random(s.shape)
        s.coef[capi.COEF_C] = rng.standard_normal(s.shape)
    s.f_field = rng.standard_normal(s.shape)
    s.bcs = [BC.dirichlet(0.3), BC.robin(1.0, 0.5, 0.1), BC.neumann(0.2), BC.dirichlet(0.0),
             BC.robin(1.0, 1.0, 0.0), BC.neumann(0.0)]
    s.bc_convention = capi.BCCONV_OUTWARD
    return s


CASES = [
    ("flux", (37, 61, 69)),     # several tiles in x and y, partial tiles
    ("flux", (129, 65, 33)),    # several z chunks
    ("const", (33, 65, 129)),
    ("coef", (17, 33, 65)),
    ("flux", (3, 3, 3)),        # one interior node
]


@pytest.mark.parametrize("kind,shape", CASES)
def test_resnorm_plane(ctx, kind, shape):
    from pde_solver_amd import hip
    print(f"resnorm {kind} {shape}", flush=True)
    s = problem(kind, shape, seed=sum(shape))
    u = np.random.default_rng(sum(shape) + 1).standard_normal(s.shape)
    hip.tune_set("stencil.plane", 0)
    l2_row, linf_row = ctx.residual_norms(s, u)
    hip.tune_set("stencil.plane", 1)
    l2, linf = ctx.residual_norms(s, u)
    st, l2_o, linf_o = O.residual_norms(s, u)
    print(f"  l2 {l2:.17g} / {l2_row:.17g} / oracle {l2_o:.17g}", flush=True)
    assert st == 0
    assert l2 == pytest.approx(l2_row, rel=1e-13)
    assert linf == pytest.approx(linf_row, rel=1e-13)
    assert l2 == pytest.approx(l2_o, rel=1e-12)
    assert linf == pytest.approx(linf_o, rel=1e-12)
