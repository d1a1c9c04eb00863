// transfer_dev.hpp — per-node arithmetic of the 3D multigrid prolongation, shared by the
// prolongation kernel (k_aux.hip) and the box sweeps that prolong on load (k_box.hip).  The
// products are written as explicit FMAs, so both kernels round every fine value the same way
// whatever the compiler's contraction choices.
#pragma once

#include "kernels.hpp"

namespace pdehip {

// In-plane bilinear stencil of fine column (i, j) through the homogeneous closure extension of
// the coarse correction: coarse face nodes are theta times their inward neighbour, so a corner
// on a coarse face maps to that neighbour with weight theta (edges and corners the product).
struct ProlongCol {
  int x0, x1, y0, y1;              // coarse columns / rows read (in range even at weight 0)
  double wx0, wx1, wy0, wy1;       // their weights (0: not read)
  double hx, hy;                   // 1 on a coarse line, 1/2 halfway between two
};

__device__ __forceinline__ ProlongCol prolong_col(const DevOp& c, int i, int j) {
  ProlongCol w;
  const int ic = i >> 1, oi = i & 1, jc = j >> 1, oj = j & 1;
  w.x0 = ic; w.x1 = ic + 1; w.y0 = jc; w.y1 = jc + 1;
  w.wx0 = 1.0; w.wx1 = 1.0; w.wy0 = 1.0; w.wy1 = 1.0;
  if (w.x0 == 0) { w.wx0 = c.theta[0]; w.x0 = 1; }
  if (w.x1 == c.nx - 1) { w.wx1 = c.theta[1]; w.x1 = c.nx - 2; }
  if (w.y0 == 0) { w.wy0 = c.theta[2]; w.y0 = 1; }
  if (w.y1 == c.ny - 1) { w.wy1 = c.theta[3]; w.y1 = c.ny - 2; }
  w.hx = oi ? 0.5 : 1.0;
  w.hy = oj ? 0.5 : 1.0;
  if (!oi) w.wx1 = 0.0;
  if (!oj) w.wy1 = 0.0;
  w.x1 = min(w.x1, c.nx - 2);
  w.y1 = min(w.y1, c.ny - 2);
  return w;
}

// Coarse plane kk through the z closure extension: the plane actually read and its weight.
__device__ __forceinline__ int prolong_zplane(const DevOp& c, int kk, double* wz) {
  *wz = 1.0;
  if (kk == 0) { *wz = c.theta[4]; return 1; }
  if (kk == c.nz - 1) { *wz = c.theta[5]; return c.nz - 2; }
  return kk;
}

// Bilinear value of one coarse plane from its four corner values (l_yx), times the plane's z
// weight.
__device__ __forceinline__ double prolong_plane(const ProlongCol& w, double l00, double l01,
                                                double l10, double l11, double wz) {
  double a = 0.0;
  if (w.wy0 != 0.0) {
    double v = w.wx0 != 0.0 ? w.wx0 * l00 : 0.0;
    if (w.wx1 != 0.0) v = fma(w.wx1, l01, v);
    a = fma(w.wy0, v * w.hx, a);
  }
  if (w.wy1 != 0.0) {
    double v = w.wx0 != 0.0 ? w.wx0 * l10 : 0.0;
    if (w.wx1 != 0.0) v = fma(w.wx1, l11, v);
    a = fma(w.wy1, v * w.hx, a);
  }
  return wz == 0.0 ? 0.0 : wz * (a * w.hy);
}

// Fine value after the correction: fine plane k = 2kc (+1: halfway) takes p_lo (the mean of
// p_lo and p_hi), the bilinear values of coarse planes kc and kc + 1.
__device__ __forceinline__ double prolong_fine(double fa, double s, bool half, double p_lo, double p_hi) {
  const double v = half ? 0.5 * (p_lo + p_hi) : p_lo;
  return fma(s, v, fa);
}

}  // namespace pdehip
