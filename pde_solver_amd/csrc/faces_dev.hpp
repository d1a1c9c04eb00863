// faces_dev.hpp — per-node boundary formulas shared by the single-problem and the batched
// kernels: the Neumann/Robin closure value of a face node and the reference residual of a 2D
// node (boundary rows with the residual's own closures).  One definition, so every path rounds
// the same way.
#pragma once

#include "kernels.hpp"

namespace pdehip {

// Closure value of face node q of face f (side 0 low, 1 high) from its inward neighbour `in`
// (reference cpu_utils.cpp:630-720 / 774-901 for AXIS; OUTWARD: alpha*u + beta*du/dn = gamma).
// Returns false when a Robin denominator vanishes (the node keeps its value).
__device__ __forceinline__ bool closure_nr_value(const FaceData& fd, int f, int side, long long q,
                                                 double in, double* out) {
  const double h = fd.h[f];
  if (fd.kind[f] == PDEHIP_BC_NEUMANN) {
    const double g = fd.value[f][q];
    if (fd.convention == PDEHIP_BCCONV_AXIS) *out = side ? in + h * g : in - h * g;
    else *out = in + h * g;
    return true;
  }
  const double al = fd.alpha[f][q], be = fd.beta[f][q], ga = fd.gamma[f][q];
  if (fd.convention == PDEHIP_BCCONV_AXIS && side) {
    const double den = al - be / h;
    if (fabs(den) > 1e-12) { *out = (ga - (be / h) * in) / den; return true; }
  } else {
    const double den = al + be / h;
    if (fabs(den) > 1e-12) { *out = (ga + (be / h) * in) / den; return true; }
  }
  return false;
}

// Reference residual of 2D boundary node id (ComputeResidualNorms, residual.cpp:153-583):
// Dirichlet u - g; Neumann u - (u_in -+ h g) along the axis; Robin u - (gamma + beta/h u_in) /
// (alpha + beta/h) on every face.  Returns false when the node is skipped (vanishing Robin
// denominator).
__device__ __forceinline__ bool resnorm_boundary2(const DevOp& op, const FaceData& fd,
                                                  const double* __restrict__ u, long long id, int i,
                                                  int j, double* r) {
  int face;
  long long inner, q;
  if (i == 0) { face = 0; inner = id + 1; q = j; }
  else if (i == op.nx - 1) { face = 1; inner = id - 1; q = j; }
  else if (j == 0) { face = 2; inner = id + op.nx; q = i; }
  else { face = 3; inner = id - op.nx; q = i; }
  const double h = fd.h[face];
  const int kind = fd.kind[face];
  if (kind == PDEHIP_BC_DIRICHLET) {
    *r = u[id] - fd.value[face][q];
  } else if (kind == PDEHIP_BC_NEUMANN) {
    const double g = fd.value[face][q];
    *r = u[id] - ((face & 1) ? u[inner] + h * g : u[inner] - h * g);
  } else {
    const double al = fd.alpha[face][q], be = fd.beta[face][q], ga = fd.gamma[face][q];
    const double den = al + be / h;
    if (fabs(den) <= 1e-12) return false;
    *r = u[id] - (ga + (be / h) * u[inner]) / den;
  }
  return true;
}

}  // namespace pdehip
