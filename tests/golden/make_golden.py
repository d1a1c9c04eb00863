"""Write the golden fixtures tests/golden/<case>.npz from the REFERENCE CPU path.

Run in the container that has the reference sources (after `make -C oracle ref`):
    python tests/golden/make_golden.py
Each fixture holds the reference's output only (inputs are rebuilt by tests/golden/cases.py):
  solve    : u (grid), l2, linf (ComputeResidualNorms), err (SolveOutput.error)
  apply    : y = A x for cases.apply_probe(spec)
  residual : l2, linf for cases.residual_probe(spec)
  time     : frames (frames x grid)
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from tests.golden import cases as K  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def main(names=None):
    if not O.ref_available():
        raise SystemExit("reference build missing: run `make -C oracle ref`")
    for c in K.CASES:
        if names and c.name not in names:
            continue
        spec, opt, rhs, coef = c.build()
        data = {}
        if c.kind == "solve":
            st, err, u, l2, linf = O.ref_solve(spec, opt, rhs_latex=rhs, coef_latex=coef)
            data = dict(u=u, l2=l2, linf=linf, status=st, err=np.array(err))
        elif c.kind == "apply":
            x = K.apply_probe(spec)
            data = dict(x=x, y=O.ref_apply(spec, x, coef_latex=coef))
        elif c.kind == "residual":
            u = K.residual_probe(spec)
            st, l2, linf = O.ref_residual_norms(spec, u, rhs_latex=rhs, coef_latex=coef)
            data = dict(u=u, l2=l2, linf=linf, status=st)
        elif c.kind == "time":
            data = dict(frames=O.ref_time_series(spec, opt, rhs_latex=rhs, coef_latex=coef))
        path = os.path.join(OUT, c.name + ".npz")
        np.savez_compressed(path, **data)
        print(f"{c.name:28s} -> {os.path.getsize(path):7d} bytes  status={data.get('status', 0)} "
              f"{str(data.get('err', ''))[:60]}")


if __name__ == "__main__":
    main(sys.argv[1:] or None)
