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

// ---------------------------------------------------------------------------- restriction
// Coarse right-hand side of one interior coarse node: full weighting of the fine residual, with
// the transpose of the homogeneous closure extension folded in (coarse nodes next to a folded
// face absorb theta times the weighted value of the face nodes beyond them; tensor product over
// the axes).  `r(i, j, k)` gives the fine residual at fine node (i, j, k).  The explicit FMA
// order makes the restriction kernel and the fused residual-restriction round alike.

// Per-axis fold list of coarse coordinate c on an axis of n nodes: the node itself (weight 1)
// plus the face node it absorbs (weight theta) when it is the first/last interior node.
// Every slot is written at a fixed index (no per-thread array indexing, so the lists stay in
// registers): slot 1 holds the low face when c == 1, else the high one; slot 2 the high face
// (used only when both apply, n == 3).
__device__ __forceinline__ int fold_list(int c, int n, double th_lo, double th_hi, int* idx, double* w) {
  const bool lo = c == 1, hi = c == n - 2;
  idx[0] = c; w[0] = 1.0;
  idx[1] = lo ? 0 : n - 1; w[1] = lo ? th_lo : th_hi;
  idx[2] = n - 1; w[2] = th_hi;
  return 1 + (lo ? 1 : 0) + (hi ? 1 : 0);
}

// 0.5 c + 0.25 w + 0.25 e
__device__ __forceinline__ double fw_row3(double w, double c, double e) {
  const double a = fma(0.25, w, 0.5 * c);
  return fma(0.25, e, a);
}

// 0.25 a0 + 0.5 a1 + 0.25 a2 of three weighted fine rows (south to north)
__device__ __forceinline__ double fw_col3(double a0, double a1, double a2) {
  return fma(0.25, a2, fma(0.5, a1, 0.25 * a0));
}

// 3 x 3 in-plane weighting at (fi, fj) of fine plane fk, all nine nodes inside the grid
template <typename R>
__device__ __forceinline__ double fw_in(const R& r, int fi, int fj, int fk) {
  const double a0 = fw_row3(r(fi - 1, fj - 1, fk), r(fi, fj - 1, fk), r(fi + 1, fj - 1, fk));
  const double a1 = fw_row3(r(fi - 1, fj, fk), r(fi, fj, fk), r(fi + 1, fj, fk));
  const double a2 = fw_row3(r(fi - 1, fj + 1, fk), r(fi, fj + 1, fk), r(fi + 1, fj + 1, fk));
  return fw_col3(a0, a1, a2);
}

// the same, clipped at the plane's edges
template <typename R>
__device__ __forceinline__ double fw_clip(const DevOp& f, const R& r, int fi, int fj, int fk) {
  const double w1[3] = {0.25, 0.5, 0.25};
  double acc = 0.0;
  for (int dj = -1; dj <= 1; ++dj) {
    const int j = fj + dj;
    if (j < 0 || j >= f.ny) continue;
    double a = 0.5 * r(fi, j, fk);
    if (fi > 0) a = fma(0.25, r(fi - 1, j, fk), a);
    if (fi < f.nx - 1) a = fma(0.25, r(fi + 1, j, fk), a);
    acc = fma(w1[dj + 1], a, acc);
  }
  return acc;
}

// Coarse value at interior coarse node (ic, jc, kc) of a 3D level, times s.  RIM: the node may
// sit next to an x/y face, whose folded closure terms take the general path (fold lists,
// clipped weighting); interior nodes (RIM false) never need it.
template <bool RIM, typename R>
__device__ __forceinline__ double restrict_value(const DevOp& f, const DevOp& c, const R& r, int ic,
                                                 int jc, int kc, double s) {
  const int fi = 2 * ic, fj = 2 * jc, fk = 2 * kc;
  double ql = fw_in(r, fi, fj, fk - 1);
  double qm = fw_in(r, fi, fj, fk);
  double qh = fw_in(r, fi, fj, fk + 1);
  bool fold_xy = false;
  int xi[3], yi[3];
  double xw[3], yw[3];
  int mx = 1, my = 1;
  xi[0] = ic; yi[0] = jc; xw[0] = 1.0; yw[0] = 1.0;
  if constexpr (RIM) {
    fold_xy = ic == 1 || jc == 1 || ic == c.nx - 2 || jc == c.ny - 2;
    if (fold_xy) {
      mx = fold_list(ic, c.nx, c.theta[0], c.theta[1], xi, xw);
      my = fold_list(jc, c.ny, c.theta[2], c.theta[3], yi, yw);
    }
  }
  auto Q = [&](int k) -> double {
    double acc = 0.0;
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      if (b >= my) break;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        if (q >= mx) break;
        const double w = yw[b] * xw[q];
        if (w != 0.0) acc = fma(w, fw_clip(f, r, 2 * xi[q], 2 * yi[b], k), acc);
      }
    }
    return acc;
  };
  if (fold_xy) { ql = Q(fk - 1); qm = Q(fk); qh = Q(fk + 1); }
  double v = fma(0.25, qh, fma(0.5, qm, 0.25 * ql));
  if (kc == 1 && c.theta[4] != 0.0)
    v = fma(c.theta[4], fma(0.25, ql, 0.5 * (fold_xy ? Q(0) : fw_in(r, fi, fj, 0))), v);
  if (kc == c.nz - 2 && c.theta[5] != 0.0)
    v = fma(c.theta[5], fma(0.5, fold_xy ? Q(fk + 2) : fw_in(r, fi, fj, fk + 2), 0.25 * qh), v);
  return v * s;
}

}  // namespace pdehip
