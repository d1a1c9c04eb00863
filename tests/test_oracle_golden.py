"""The CPU oracle against the reference's own outputs (golden fixtures written by
tests/golden/make_golden.py from the reference CPU build)."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from pde_solver_amd import capi
from tests.golden import cases as K

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


SOLVE = [c for c in K.CASES if c.kind == "solve"]


@pytest.mark.parametrize("case", SOLVE, ids=lambda c: c.name)
def test_oracle_solve_matches_reference(case):
    g = load(case.name)
    spec, opt, _, _ = case.build()
    st, u, rep = O.solve(spec, opt)
    if case.ref_fails:
        assert int(g["status"]) == 1 and "failed to converge" in str(g["err"])
        if case.name.startswith("cg_"):
            # the oracle restates the reference rows and fails the same way
            assert st != 0
        else:
            # GMRES: the oracle uses the current cycle's step count and converges
            assert st == 0 and rep.final_metric < opt.tol
        return
    assert int(g["status"]) == 0, str(g["err"])
    assert st == 0, (st, O._load_oracle().oracle_last_error())
    # restatement of the same algorithm: agreement to rounding
    assert rel(u, g["u"]) < 1e-11, rel(u, g["u"])
    # ComputeResidualNorms restated on the same grid
    st2, l2, linf = O.residual_norms(spec, g["u"])
    assert st2 == 0
    assert l2 == pytest.approx(float(g["l2"]), rel=1e-9, abs=1e-14)
    assert linf == pytest.approx(float(g["linf"]), rel=1e-9, abs=1e-14)


@pytest.mark.parametrize("case", [c for c in K.CASES if c.kind == "apply"], ids=lambda c: c.name)
def test_oracle_apply_matches_reference(case):
    g = load(case.name)
    spec, _, _, _ = case.build()
    y = O.apply(spec, g["x"])
    assert rel(y, g["y"]) < 1e-14


@pytest.mark.parametrize("case", [c for c in K.CASES if c.kind == "residual"], ids=lambda c: c.name)
def test_oracle_residual_norms_match_reference(case):
    g = load(case.name)
    spec, _, _, _ = case.build()
    st, l2, linf = O.residual_norms(spec, g["u"])
    assert st == 0
    assert l2 == pytest.approx(float(g["l2"]), rel=1e-12)
    assert linf == pytest.approx(float(g["linf"]), rel=1e-12)


@pytest.mark.parametrize("case", [c for c in K.CASES if c.kind == "time"], ids=lambda c: c.name)
def test_oracle_time_series_matches_reference(case):
    g = load(case.name)
    spec, tc, _, _ = case.build()
    frames = O.time_series(spec, tc)
    ref = g["frames"]
    assert frames.shape == ref.shape
    for f in range(ref.shape[0]):
        assert rel(frames[f], ref[f]) < 1e-13, (f, rel(frames[f], ref[f]))
