"""Closed-form discrete solutions shared by the CPU (oracle) and GPU (HIP) tests."""
import numpy as np

from pde_solver_amd import capi
from pde_solver_amd.capi import BC


def two_layer_conductor(nx, m, k1, k2, ny=9, nz=1):
    """-div(kappa grad u) = 0 on [0,1]^d, u = 0 at x = 0, u = 1 at x = 1, zero flux on the other
    faces; kappa = k1 at nodes i <= m, k2 at nodes i > m.

    With the face conductivity taken as the harmonic mean of its two nodes, the one face that
    straddles the jump has resistance h (1/k1 + 1/k2) / 2: exactly that of two layers meeting
    at its midpoint x* = (m + 1/2) h.  The flux q is then the same on every face and the
    discrete solution equals the continuous two-layer solution at every node,
        u(x) = q x / k1 (x < x*),  1 - q (1 - x) / k2 (x > x*),  q = 1 / (x*/k1 + (1 - x*)/k2).
    An arithmetic-mean face ((k1 + k2) / 2) gives a different flux.  Returns (spec, exact)."""
    s = capi.Spec(nx, ny, nz, a=-1.0, op_kind=capi.OP_FLUX)
    X = s.coords()[0]
    idx = np.rint(X * (nx - 1)).astype(int)
    s.kappa = np.where(idx <= m, k1, k2).astype(np.float64)
    s.bcs = [BC.dirichlet(0.0), BC.dirichlet(1.0), BC.neumann(0.0), BC.neumann(0.0),
             BC.neumann(0.0), BC.neumann(0.0)]
    s.bc_convention = capi.BCCONV_OUTWARD
    h = 1.0 / (nx - 1)
    xs = (m + 0.5) * h
    q = 1.0 / (xs / k1 + (1.0 - xs) / k2)
    exact = np.where(idx <= m, q * X / k1, 1.0 - q * (1.0 - X) / k2)
    return s, exact


def arithmetic_mean_solution(nx, m, k1, k2):
    """Nodal solution the same problem would have with arithmetic-mean faces (1D series sum)."""
    kf = np.array([k1] * m + [(k1 + k2) / 2] + [k2] * (nx - 2 - m))
    res = 1.0 / kf
    return np.concatenate([[0.0], np.cumsum(res) / res.sum()])
