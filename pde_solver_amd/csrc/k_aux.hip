// k_aux.hip — boundary closures, multigrid transfers, coarsest-level solve, vector kernels and
// the deterministic two-stage reductions whose second stage updates device-resident scalars.
#include "faces_dev.hpp"
#include "transfer_dev.hpp"

namespace pdehip {
namespace {

template <Prec P> struct PT;
template <> struct PT<P64> { using T = double; using S = double; };
template <> struct PT<P32> { using T = float; using S = float; };
template <> struct PT<P16> { using T = float; using S = __half; };

__device__ __forceinline__ double ldd(const double* p, long long i) { return p[i]; }
__device__ __forceinline__ double ldd(const float* p, long long i) { return p[i]; }
__device__ __forceinline__ double ldd(const __half* p, long long i) { return __half2float(p[i]); }
__device__ __forceinline__ double to_double(double v) { return v; }
__device__ __forceinline__ double to_double(float v) { return v; }
__device__ __forceinline__ double to_double(__half v) { return __half2float(v); }

__device__ __forceinline__ bool skip(const KScalars* sc) { return sc && (sc->done | sc->breakdown); }

long long grid1(long long n, long long cap = 8192) {
  long long b = (n + kBlock - 1) / kBlock;
  if (b < 1) b = 1;
  return b > cap ? cap : b;
}

// Face node q of axis `axis` on the low (side 0) or high (side 1) face: node index and the
// index of its inward neighbour.  Face node order: x faces k*ny+j, y faces k*nx+i, z faces j*nx+i.
__device__ __forceinline__ void face_node(const DevOp& op, int axis, int side, long long q,
                                          long long* node, long long* inner) {
  if (axis == 0) {
    const long long j = q % op.ny, k = q / op.ny;
    const long long i = side ? op.nx - 1 : 0, ii = side ? op.nx - 2 : 1;
    *node = (k * op.ny + j) * op.nx + i;
    *inner = (k * op.ny + j) * op.nx + ii;
  } else if (axis == 1) {
    const long long i = q % op.nx, k = q / op.nx;
    const long long j = side ? op.ny - 1 : 0, jj = side ? op.ny - 2 : 1;
    *node = (k * op.ny + j) * op.nx + i;
    *inner = (k * op.ny + jj) * op.nx + i;
  } else {
    const long long i = q % op.nx, j = q / op.nx;
    const long long k = side ? op.nz - 1 : 0, kk = side ? op.nz - 2 : 1;
    *node = (k * op.ny + j) * op.nx + i;
    *inner = (kk * op.ny + j) * op.nx + i;
  }
}

__host__ __device__ inline long long face_size(const DevOp& op, int axis) {
  return axis == 0 ? (long long)op.ny * op.nz : (axis == 1 ? (long long)op.nx * op.nz : op.sxy);
}

// ----------------------------------------------------------------------------- closures
// Neumann/Robin closure of one axis, low face then high face per thread (reference
// cpu_utils.cpp:630-720 / 774-901 for AXIS; OUTWARD: alpha*u + beta*du/dn = gamma).
template <typename S>
__global__ void k_closure_nr(DevOp op, FaceData fd, int axis, S* u, const KScalars* sc) {
  if (skip(sc)) return;
  const long long nq = face_size(op, axis);
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < nq;
       q += (long long)gridDim.x * blockDim.x) {
    for (int side = 0; side < 2; ++side) {
      const int f = 2 * axis + side;
      const int kind = fd.kind[f];
      if (kind == PDEHIP_BC_DIRICHLET) continue;
      long long node, inner;
      face_node(op, axis, side, q, &node, &inner);
      double v;
      if (closure_nr_value(fd, f, side, q, static_cast<double>(u[inner]), &v)) u[node] = static_cast<S>(v);
    }
  }
}

template <typename S>
__global__ void k_closure_d(DevOp op, FaceData fd, int axis, S* u, const KScalars* sc) {
  if (skip(sc)) return;
  const long long nq = face_size(op, axis);
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < nq;
       q += (long long)gridDim.x * blockDim.x) {
    for (int side = 0; side < 2; ++side) {
      const int f = 2 * axis + side;
      if (fd.kind[f] != PDEHIP_BC_DIRICHLET) continue;
      long long node, inner;
      face_node(op, axis, side, q, &node, &inner);
      u[node] = static_cast<S>(fd.value[f][q]);
    }
  }
}

// ----------------------------------------------------------------------------- transfers
// Both transfers are column marches: a 64 x 4 workgroup owns 64 x 4 (i, j) columns and walks a
// chunk of z planes, carrying the z interpolation (prolongation) or z weighting (restriction)
// in registers, so each in-plane sum is formed once.  Index math is 32-bit.
constexpr int kTX = 64, kTY = 4;

// z planes per workgroup: 16, shortened until the launch has ~4096 workgroups (latency hiding
// on the small levels).
int zchunk_for(int gx, int gy, int nz) {
  int zc = 16;
  while (zc > 1 && (long long)gx * gy * ((nz + zc - 1) / zc) < 4096) zc /= 2;
  return zc;
}

// 3 x 3 in-plane full weighting of fine plane `pl` at (fi, fj), clipped at the plane edges.
template <typename SF>
__device__ __forceinline__ double fw2(const DevOp& f, const SF* pl, int fi, int fj) {
  const double w1[3] = {0.25, 0.5, 0.25};
  double acc = 0.0;
  for (int dj = -1; dj <= 1; ++dj) {
    const int j = fj + dj;
    if (j < 0 || j >= f.ny) continue;
    const SF* row = pl + j * f.nx;
    double a = 0.5 * ldd(row, fi);
    if (fi > 0) a += 0.25 * ldd(row, fi - 1);
    if (fi < f.nx - 1) a += 0.25 * ldd(row, fi + 1);
    acc += w1[dj + 1] * a;
  }
  return acc;
}

// Coarse right-hand side = FoldTranspose(FW(r_fine)) * scale in one pass: interior coarse
// nodes absorb theta times the restricted value of the face nodes next to them (tensor product
// over the axes, which is what folding x, then y, then z faces composes to); coarse boundary
// nodes are zero.
template <Prec PF, Prec PC>
__global__ void __launch_bounds__(kTX * kTY) k_restrict(DevOp f, const void* rfv, DevOp c,
                                                       void* bcv, double scale,
                                                       const double* dscale, int kZC,
                                                       const KScalars* sc) {
  using SF = typename PT<PF>::S;
  using SC = typename PT<PC>::S;
  if (skip(sc)) return;
  const int ic = blockIdx.x * kTX + threadIdx.x;
  const int jc = blockIdx.y * kTY + threadIdx.y;
  if (ic >= c.nx || jc >= c.ny) return;
  const bool d3 = c.nz > 1;
  const int kc0 = d3 ? blockIdx.z * kZC : 0;
  const int kc1 = d3 ? min(kc0 + kZC, c.nz) : 1;
  const SF* rf = static_cast<const SF*>(rfv);
  SC* bc = static_cast<SC*>(bcv);
  const long long csxy = (long long)c.nx * c.ny;
  const long long fsxy = (long long)f.nx * f.ny;
  const long long ccol = (long long)jc * c.nx + ic;
  if (ic == 0 || jc == 0 || ic == c.nx - 1 || jc == c.ny - 1) {
    for (int kc = kc0; kc < kc1; ++kc) bc[kc * csxy + ccol] = store_as<SC>(0.0);
    return;
  }
  const double s = dscale ? scale * (*dscale) : scale;
  int xi[3], yi[3];
  double xw[3], yw[3];
  const int mx = fold_list(ic, c.nx, c.theta[0], c.theta[1], xi, xw);
  const int my = fold_list(jc, c.ny, c.theta[2], c.theta[3], yi, yw);
  // in-plane folded full weighting of fine plane fk at this coarse column
  auto Q = [&](int fk) -> double {
    if (fk < 0 || fk >= f.nz) return 0.0;
    const SF* pl = rf + fk * fsxy;
    double acc = 0.0;
    for (int b = 0; b < my; ++b)
      for (int q = 0; q < mx; ++q) {
        const double w = yw[b] * xw[q];
        if (w != 0.0) acc += w * fw2(f, pl, 2 * xi[q], 2 * yi[b]);
      }
    return acc;
  };
  if (!d3) {
    bc[ccol] = store_as<SC>(Q(0) * s);
    return;
  }
  double q_lo = (kc0 >= 1) ? Q(2 * kc0 - 1) : 0.0;   // Q(2kc - 1) of the current kc
  for (int kc = kc0; kc < kc1; ++kc) {
    if (kc == 0 || kc == c.nz - 1) {
      bc[kc * csxy + ccol] = store_as<SC>(0.0);
      q_lo = (kc == 0) ? Q(1) : q_lo;
      continue;
    }
    const double q_mid = Q(2 * kc), q_hi = Q(2 * kc + 1);
    double v = 0.25 * q_lo + 0.5 * q_mid + 0.25 * q_hi;
    if (kc == 1 && c.theta[4] != 0.0) v += c.theta[4] * (0.5 * Q(0) + 0.25 * q_lo);
    if (kc == c.nz - 2 && c.theta[5] != 0.0) v += c.theta[5] * (0.25 * q_hi + 0.5 * Q(2 * kc + 2));
    bc[kc * csxy + ccol] = store_as<SC>(v * s);
    q_lo = q_hi;
  }
}

// One coarse node of the 3D restriction (restrict_value, transfer_dev.hpp) reading the fine
// residual from memory; coarse boundary nodes are zero.  The 27 loads of an interior node are
// issued unconditionally and wait once.
template <typename SF, typename SC, bool RIM>
__device__ __forceinline__ void restrict3_node(const DevOp& f, const SF* __restrict__ rf, const DevOp& c,
                                               SC* __restrict__ bc, int ic, int jc, int kc, double s) {
  const long long out = ((long long)kc * c.ny + jc) * c.nx + ic;
  if (ic == 0 || jc == 0 || kc == 0 || ic == c.nx - 1 || jc == c.ny - 1 || kc == c.nz - 1) {
    bc[out] = store_as<SC>(0.0);
    return;
  }
  const long long fsxy = (long long)f.nx * f.ny;
  auto r = [&](int i, int j, int k) -> double { return ldd(rf, k * fsxy + (long long)j * f.nx + i); };
  bc[out] = store_as<SC>(restrict_value<RIM>(f, c, r, ic, jc, kc, s));
}

// Rim of a coarse plane: the nodes outside the interior box [2, nx-3] x [2, ny-3], numbered
// row by row (rows 0, 1, ny-2, ny-1 whole, then 4 nodes of every middle row); every node when
// the box is empty.
__device__ __forceinline__ bool rim_node(int nx, int ny, long long r, int* ic, int* jc) {
  const int bx = nx - 4, by = ny - 4;
  if (bx <= 0 || by <= 0) {
    if (r >= (long long)nx * ny) return false;
    *jc = (int)(r / nx); *ic = (int)(r % nx);
    return true;
  }
  if (r < 2LL * nx) { *jc = (int)(r / nx); *ic = (int)(r % nx); return true; }
  if (r < 4LL * nx) { const long long q = r - 2LL * nx; *jc = ny - 2 + (int)(q / nx); *ic = (int)(q % nx); return true; }
  const long long q = r - 4LL * nx;
  if (q >= 4LL * by) return false;
  *jc = 2 + (int)(q / 4);
  const int t = (int)(q % 4);
  *ic = t < 2 ? t : nx - 4 + t;
  return true;
}

// 3D restriction, one coarse node per thread (no z march, so no load waits on a carried
// value): the 27-point full weighting of an interior coarse node reads fine nodes that are all
// in range, so its 27 loads are issued unconditionally and wait once.  Blocks y < gy_in cover
// the interior box; the blocks above it cover the rim, whose nodes next to a face add the
// folded closure terms through the general path (fold lists, clipped weighting).
template <Prec PF, Prec PC>
__global__ void __launch_bounds__(kTX * kTY) k_restrict3(DevOp f, const void* rfv, DevOp c,
                                                        void* bcv, double scale,
                                                        const double* dscale,
                                                        const KScalars* sc, int gy_in) {
  using SF = typename PT<PF>::S;
  using SC = typename PT<PC>::S;
  if (skip(sc)) return;
  const SF* rf = static_cast<const SF*>(rfv);
  SC* bc = static_cast<SC*>(bcv);
  const double s = dscale ? scale * (*dscale) : scale;
  const int kc = blockIdx.z;
  if ((int)blockIdx.y < gy_in) {
    const int ic = 2 + blockIdx.x * kTX + threadIdx.x;
    const int jc = 2 + blockIdx.y * kTY + threadIdx.y;
    if (ic > c.nx - 3 || jc > c.ny - 3) return;
    restrict3_node<SF, SC, false>(f, rf, c, bc, ic, jc, kc, s);
  } else {
    const long long r = ((long long)(blockIdx.y - gy_in) * gridDim.x + blockIdx.x) * (kTX * kTY) +
                        threadIdx.y * kTX + threadIdx.x;
    int ic, jc;
    if (!rim_node(c.nx, c.ny, r, &ic, &jc)) return;
    restrict3_node<SF, SC, true>(f, rf, c, bc, ic, jc, kc, s);
  }
}

// restrict3_node<..., RIM = true> for a node next to an x or y face (ic or jc in {1, n-2}), with
// the same operations in the same order, its fine values loaded first: each folded in-plane
// weighting Q(k) reads the (up to) 2 x 2 fw_clip windows of its fold lists at once, clamped into
// the grid, instead of one dependent load group per window and row.  The march's per-node
// workgroups serve the rim with it.
template <typename SF, typename SC>
__device__ __forceinline__ void restrict_rim_node(const DevOp& f, const SF* __restrict__ rf,
                                                  const DevOp& c, SC* __restrict__ bc, int ic,
                                                  int jc, int kc, double s) {
  const long long out = ((long long)kc * c.ny + jc) * c.nx + ic;
  if (ic == 0 || jc == 0 || kc == 0 || ic == c.nx - 1 || jc == c.ny - 1 || kc == c.nz - 1) {
    bc[out] = store_as<SC>(0.0);
    return;
  }
  if (!(ic == 1 || jc == 1 || ic == c.nx - 2 || jc == c.ny - 2)) {
    restrict3_node<SF, SC, true>(f, rf, c, bc, ic, jc, kc, s);
    return;
  }
  const long long fsxy = (long long)f.nx * f.ny;
  // fold lists (fold_list; at most two entries per axis, as the coarse sides are at least 5)
  const int xi[2] = {ic, ic == 1 ? 0 : c.nx - 1}, yi[2] = {jc, jc == 1 ? 0 : c.ny - 1};
  const double xw[2] = {1.0, ic == 1 ? c.theta[0] : c.theta[1]};
  const double yw[2] = {1.0, jc == 1 ? c.theta[2] : c.theta[3]};
  const int mx = ic == 1 || ic == c.nx - 2 ? 2 : 1, my = jc == 1 || jc == c.ny - 2 ? 2 : 1;
  // restrict_value's Q: sum over the fold lists of w * fw_clip(2 xi, 2 yi, k)
  auto Q = [&](int fk) -> double {
    const SF* __restrict__ pl = rf + fk * fsxy;
    double v[2][2][3][3];   // [y entry][x entry][row fj-1 .. fj+1][column fi-1 .. fi+1]
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int fi = 2 * xi[q], fj = 2 * yi[b];
#pragma unroll
        for (int dj = 0; dj < 3; ++dj) {
          const long long row = (long long)min(max(fj + dj - 1, 0), f.ny - 1) * f.nx;
#pragma unroll
          for (int di = 0; di < 3; ++di)
            v[b][q][dj][di] = ldd(pl, row + min(max(fi + di - 1, 0), f.nx - 1));
        }
      }
    const double w1[3] = {0.25, 0.5, 0.25};
    double acc = 0.0;
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        if (b >= my || q >= mx) continue;
        const double w = yw[b] * xw[q];
        if (w == 0.0) continue;
        const int fi = 2 * xi[q], fj = 2 * yi[b];
        double clip = 0.0;   // fw_clip(f, r, fi, fj, fk)
#pragma unroll
        for (int dj = -1; dj <= 1; ++dj) {
          const int j = fj + dj;
          if (j < 0 || j >= f.ny) continue;
          double a = 0.5 * v[b][q][dj + 1][1];
          if (fi > 0) a = fma(0.25, v[b][q][dj + 1][0], a);
          if (fi < f.nx - 1) a = fma(0.25, v[b][q][dj + 1][2], a);
          clip = fma(w1[dj + 1], a, clip);
        }
        acc = fma(w, clip, acc);
      }
    return acc;
  };
  const int fk = 2 * kc;
  const double ql = Q(fk - 1), qm = Q(fk), qh = Q(fk + 1);
  double v = fma(0.25, qh, fma(0.5, qm, 0.25 * ql));
  if (kc == 1 && c.theta[4] != 0.0) v = fma(c.theta[4], fma(0.25, ql, 0.5 * Q(0)), v);
  if (kc == c.nz - 2 && c.theta[5] != 0.0) v = fma(c.theta[5], fma(0.5, Q(fk + 2), 0.25 * qh), v);
  bc[out] = store_as<SC>(v * s);
}

// 3D restriction as a plane march over the interior box [2, nx-3] x [2, ny-3] of every interior
// coarse plane; the rim (restrict_rim_node) and the boundary planes take the per-node path in
// the workgroups past the march's.  A wave owns kRMRows coarse rows of
// kRMCols = 63 coarse columns (lane l: coarse column ic, fine columns 2ic and 2ic+1; lane 0 is the
// west halo column whose 2ic+1 value its east neighbour needs, passed by a DPP wave shift), a
// workgroup kRMWaves such waves, and a chunk of coarse planes.  Each fine plane of the chunk is
// loaded once (two values per lane per fine row, the 2 kRMRows + 1 fine rows under the wave's
// coarse rows), weighted in the plane (fw_row3 per fine row, fw_col3 per coarse row: fw_in
// operation for operation) and carried: coarse plane kc's low fine plane 2kc-1 is the previous
// plane's high one.  The z combination and the z folds are restrict_value's, so every coarse
// value equals the per-node kernel's bit for bit.
constexpr int kRMLanes = 64;
constexpr int kRMRows = 4;
constexpr int kRMWaves = 4;
constexpr int kRMCols = kRMLanes - 1;

// lane l gets lane l-1's v (lane 0: 0, a halo lane whose result is never stored)
__device__ __forceinline__ double dpp_west_d(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const int lo = __builtin_amdgcn_mov_dpp(static_cast<int>(u & 0xffffffffu), 0x138, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp(static_cast<int>(u >> 32), 0x138, 0xf, 0xf, true);
  return __builtin_bit_cast(double, (static_cast<unsigned long long>(static_cast<unsigned>(hi)) << 32) |
                                        static_cast<unsigned>(lo));
}

template <Prec PF, Prec PC>
__global__ void __launch_bounds__(kRMLanes * kRMWaves)
k_restrict3m(DevOp f, const void* rfv, DevOp c, void* bcv, double scale, const double* dscale,
             const KScalars* sc, int ntx, int nty, int kchunk, int nmarch, long long nrim_items) {
  using SF = typename PT<PF>::S;
  using SC = typename PT<PC>::S;
  constexpr int R = kRMRows;
  constexpr int FR = 2 * R + 1;
  if (skip(sc)) return;
  const SF* __restrict__ rf = static_cast<const SF*>(rfv);
  SC* __restrict__ bc = static_cast<SC*>(bcv);
  const double s = dscale ? scale * (*dscale) : scale;
  const long long fsxy = (long long)f.nx * f.ny;
  const int nrb = static_cast<int>(gridDim.x) - nmarch;   // per-node workgroups, dispatched last
  int b = static_cast<int>(blockIdx.x) - nmarch;
  if (b >= 0) {
    // boundary planes (every node, zero), then the rim of each interior plane
    const long long plane = (long long)c.nx * c.ny;
    const int bx = c.nx - 4, by = c.ny - 4;
    const long long nrim = bx > 0 && by > 0 ? 4LL * c.nx + 4LL * by : plane;
    for (long long it = (long long)b * blockDim.x + threadIdx.x; it < nrim_items;
         it += (long long)nrb * blockDim.x) {
      int ic, jc, kc;
      if (it < 2 * plane) {
        kc = it < plane ? 0 : c.nz - 1;
        const long long q = it % plane;
        jc = static_cast<int>(q / c.nx);
        ic = static_cast<int>(q % c.nx);
      } else {
        const long long q = it - 2 * plane;
        kc = 1 + static_cast<int>(q / nrim);
        if (!rim_node(c.nx, c.ny, q % nrim, &ic, &jc)) continue;
      }
      restrict_rim_node<SF, SC>(f, rf, c, bc, ic, jc, kc, s);
    }
    return;
  }
  b += nmarch;
  const int tx = b % ntx, rest = b / ntx, ty = rest % nty, cz = rest / nty;
  const int lane = static_cast<int>(threadIdx.x) & (kRMLanes - 1);
  const int wv = static_cast<int>(threadIdx.x) / kRMLanes;
  const int ic = 1 + tx * kRMCols + lane;
  const int icl = min(ic, c.nx - 2);          // lanes past the box load inside the fine row
  const int jc0 = 2 + ty * (R * kRMWaves) + wv * R;
  const int kc0 = 1 + cz * kchunk;
  const int kc1 = min(kc0 + kchunk, c.nz - 1);
  const bool col_out = lane >= 1 && ic <= c.nx - 3;
  long long roff[FR];
#pragma unroll
  for (int m = 0; m < FR; ++m) roff[m] = (long long)min(2 * jc0 - 1 + m, f.ny - 1) * f.nx + 2 * icl;
  // a fine plane's values under the wave (fine columns 2ic and 2ic+1 of each fine row), and
  // their fw_in at the wave's R coarse rows
  auto load_plane = [&](int fk, SF* w0, SF* w1) {
    const SF* __restrict__ pl = rf + fk * fsxy;
#pragma unroll
    for (int m = 0; m < FR; ++m) { w0[m] = pl[roff[m]]; w1[m] = pl[roff[m] + 1]; }
  };
  auto weigh = [&](const SF* w0, const SF* w1, double* q) {
    double a[FR];
#pragma unroll
    for (int m = 0; m < FR; ++m) {
      const double v0 = to_double(w0[m]), v1 = to_double(w1[m]);
      a[m] = fw_row3(dpp_west_d(v1), v0, v1);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) q[r] = fw_col3(a[2 * r], a[2 * r + 1], a[2 * r + 2]);
  };
  auto plane_q = [&](int fk, double* q) {
    SF w0[FR], w1[FR];
    load_plane(fk, w0, w1);
    weigh(w0, w1, q);
  };
  const double th_lo = c.theta[4], th_hi = c.theta[5];
  double ql[R], qm[R], qh[R], qx[R];
  // the next step's two fine planes load while the current step weighs its own
  SF cm0[FR], cm1[FR], ch0[FR], ch1[FR], nm0[FR], nm1[FR], nh0[FR], nh1[FR];
  load_plane(2 * kc0, cm0, cm1);
  load_plane(2 * kc0 + 1, ch0, ch1);
  plane_q(2 * kc0 - 1, ql);
  for (int kc = kc0; kc < kc1; ++kc) {
    const bool more = kc + 1 < kc1;   // workgroup-uniform
    if (more) {
      load_plane(2 * kc + 2, nm0, nm1);
      load_plane(2 * kc + 3, nh0, nh1);
    }
    weigh(cm0, cm1, qm);
    weigh(ch0, ch1, qh);
    double v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = fma(0.25, qh[r], fma(0.5, qm[r], 0.25 * ql[r]));
    if (kc == 1 && th_lo != 0.0) {   // z folds (workgroup-uniform branches)
      plane_q(0, qx);
#pragma unroll
      for (int r = 0; r < R; ++r) v[r] = fma(th_lo, fma(0.25, ql[r], 0.5 * qx[r]), v[r]);
    }
    if (kc == c.nz - 2 && th_hi != 0.0) {
      plane_q(2 * kc + 2, qx);
#pragma unroll
      for (int r = 0; r < R; ++r) v[r] = fma(th_hi, fma(0.5, qx[r], 0.25 * qh[r]), v[r]);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int jc = jc0 + r;
      if (col_out && jc <= c.ny - 3) bc[((long long)kc * c.ny + jc) * c.nx + ic] = store_as<SC>(v[r] * s);
      ql[r] = qh[r];
    }
    if (more) {
#pragma unroll
      for (int m = 0; m < FR; ++m) { cm0[m] = nm0[m]; cm1[m] = nm1[m]; ch0[m] = nh0[m]; ch1[m] = nh1[m]; }
    }
  }
}

// 3D prolongation, kProlongZ coarse intervals per thread: fine planes 2kc .. 2kc+2*kProlongZ-1
// of one (i, j) column (interior planes only) and the coarse planes kc .. kc+kProlongZ they lie
// between.  Every load is issued before any arithmetic (plane indices clamped in range; a
// value that is not needed is not stored), so a thread keeps 2*kProlongZ + 4*(kProlongZ+1)
// loads in flight instead of 10.  The arithmetic of each fine node is the one-interval form.
constexpr int kProlongZ = 4;
template <Prec PC, Prec PF>
__global__ void __launch_bounds__(kTX * kTY) k_prolong3(DevOp c, const void* ucv, DevOp f,
                                                       void* ufv, double factor,
                                                       const double* dscale,
                                                       const KScalars* sc) {
  using SC = typename PT<PC>::S;
  using SF = typename PT<PF>::S;
  if (skip(sc)) return;
  const int i = blockIdx.x * kTX + threadIdx.x;
  const int j = blockIdx.y * kTY + threadIdx.y;
  const int kc0 = blockIdx.z * kProlongZ;
  if (i < 1 || j < 1 || i > f.nx - 2 || j > f.ny - 2) return;
  const SC* uc = static_cast<const SC*>(ucv);
  SF* uf = static_cast<SF*>(ufv);
  const long long csxy = (long long)c.nx * c.ny;
  const long long fsxy = (long long)f.nx * f.ny;
  const long long col = (long long)j * f.nx + i;
  const ProlongCol w = prolong_col(c, i, j);
  // loads: coarse corners of planes kc0 .. kc0+kProlongZ (through the z closure extension),
  // fine values of planes 2kc0 .. 2kc0+2kProlongZ-1
  double l00[kProlongZ + 1], l01[kProlongZ + 1], l10[kProlongZ + 1], l11[kProlongZ + 1];
  double wz[kProlongZ + 1];
#pragma unroll
  for (int q = 0; q <= kProlongZ; ++q) {
    const int kk = prolong_zplane(c, min(kc0 + q, c.nz - 1), &wz[q]);
    const SC* pl = uc + kk * csxy;
    l00[q] = ldd(pl + (long long)w.y0 * c.nx, w.x0);
    l01[q] = ldd(pl + (long long)w.y0 * c.nx, w.x1);
    l10[q] = ldd(pl + (long long)w.y1 * c.nx, w.x0);
    l11[q] = ldd(pl + (long long)w.y1 * c.nx, w.x1);
  }
  double fv[2 * kProlongZ];
#pragma unroll
  for (int q = 0; q < 2 * kProlongZ; ++q) {
    const int k = min(max(2 * kc0 + q, 1), f.nz - 2);
    fv[q] = ldd(uf, k * fsxy + col);
  }
  const double s = dscale ? factor * (*dscale) : factor;
  double p_lo = prolong_plane(w, l00[0], l01[0], l10[0], l11[0], wz[0]);
#pragma unroll
  for (int q = 0; q < kProlongZ; ++q) {
    const double p_hi = prolong_plane(w, l00[q + 1], l01[q + 1], l10[q + 1], l11[q + 1], wz[q + 1]);
    // fine plane 2kc + o sits at coarse kc (o = 0) or halfway to kc + 1 (o = 1)
    const int kc = kc0 + q;
    const int ka = max(1, 2 * kc), kb = min(2 * kc + 2, f.nz - 1);   // [ka, kb)
    if (ka < kb) {
      const int o0 = ka & 1;      // 1 only for kc = 0 (ka = 1)
      const double fa = o0 ? fv[2 * q + 1] : fv[2 * q];   // fine plane ka = 2kc + o0
      uf[ka * fsxy + col] = store_as<SF>(prolong_fine(fa, s, o0 != 0, p_lo, p_hi));
      if (kb - ka == 2) uf[(ka + 1) * fsxy + col] = store_as<SF>(prolong_fine(fv[2 * q + 1], s, true, p_lo, p_hi));
    }
    p_lo = p_hi;
  }
}

// fine interior += factor * trilinear interpolation of the coarse correction, whose face nodes
// are taken as the homogeneous closure extension: theta times the inward neighbour per axis
// (edges and corners the product), which is what extending x, then y, then z faces composes
// to.
template <Prec PC, Prec PF>
__global__ void __launch_bounds__(kTX * kTY) k_prolong(DevOp c, const void* ucv, DevOp f,
                                                      void* ufv, double factor,
                                                      const double* dscale, int kZC,
                                                      const KScalars* sc) {
  using SC = typename PT<PC>::S;
  using SF = typename PT<PF>::S;
  if (skip(sc)) return;
  const int i = blockIdx.x * kTX + threadIdx.x;
  const int j = blockIdx.y * kTY + threadIdx.y;
  if (i < 1 || j < 1 || i > f.nx - 2 || j > f.ny - 2) return;
  const bool d3 = f.nz > 1;
  const int k0 = d3 ? max(1, (int)blockIdx.z * kZC) : 0;
  const int k1 = d3 ? min((int)blockIdx.z * kZC + kZC, f.nz - 1) : 1;
  if (k0 >= k1) return;
  const SC* uc = static_cast<const SC*>(ucv);
  SF* uf = static_cast<SF*>(ufv);
  const double s = dscale ? factor * (*dscale) : factor;
  const long long csxy = (long long)c.nx * c.ny;
  const long long fsxy = (long long)f.nx * f.ny;
  // in-plane bilinear stencil (columns x0/x1, rows y0/y1) through the closure extension
  const int ic = i >> 1, oi = i & 1, jc = j >> 1, oj = j & 1;
  int x0 = ic, x1 = ic + 1, y0 = jc, y1 = jc + 1;
  double wx0 = 1.0, wx1 = 1.0, wy0 = 1.0, wy1 = 1.0;
  if (x0 == 0) { wx0 = c.theta[0]; x0 = 1; }
  if (x1 == c.nx - 1) { wx1 = c.theta[1]; x1 = c.nx - 2; }
  if (y0 == 0) { wy0 = c.theta[2]; y0 = 1; }
  if (y1 == c.ny - 1) { wy1 = c.theta[3]; y1 = c.ny - 2; }
  const double hx = oi ? 0.5 : 1.0, hy = oj ? 0.5 : 1.0;
  if (!oi) wx1 = 0.0;
  if (!oj) wy1 = 0.0;
  // bilinear value of coarse plane kc (mapped through the z extension)
  auto P = [&](int kc) -> double {
    double wz = 1.0;
    if (d3) {
      if (kc == 0) { wz = c.theta[4]; kc = 1; } else if (kc == c.nz - 1) { wz = c.theta[5]; kc = c.nz - 2; }
    }
    if (wz == 0.0) return 0.0;
    const SC* pl = uc + kc * csxy;
    double a = 0.0;
    if (wy0 != 0.0) {
      const SC* r = pl + y0 * c.nx;
      double v = wx0 != 0.0 ? wx0 * ldd(r, x0) : 0.0;
      if (wx1 != 0.0) v += wx1 * ldd(r, x1);
      a += wy0 * (v * hx);
    }
    if (wy1 != 0.0) {
      const SC* r = pl + y1 * c.nx;
      double v = wx0 != 0.0 ? wx0 * ldd(r, x0) : 0.0;
      if (wx1 != 0.0) v += wx1 * ldd(r, x1);
      a += wy1 * (v * hx);
    }
    return wz * (a * hy);
  };
  int kc_have = -1;
  double p_lo = 0.0, p_hi = 0.0;
  for (int k = k0; k < k1; ++k) {
    const int kc = k >> 1, ok = d3 ? (k & 1) : 0;
    if (kc != kc_have) {
      p_lo = (kc == kc_have + 1 && kc_have >= 0) ? p_hi : P(kc);
      p_hi = d3 ? P(kc + 1) : 0.0;
      kc_have = kc;
    }
    const double v = ok ? 0.5 * (p_lo + p_hi) : p_lo;
    const long long id = k * fsxy + (long long)j * f.nx + i;
    uf[id] = store_as<SF>(ldd(uf, id) + s * v);
  }
}

// One wavefront per row of the dense inverse: x[map[r]] = sum_q Ainv[r][q] * b[map[q]].
template <Prec P>
__global__ void __launch_bounds__(kBlock) k_coarse_dense(const double* __restrict__ ainv,
                                                        const int* __restrict__ map, int m,
                                                        const void* bv, void* xv,
                                                        const KScalars* sc) {
  using S = typename PT<P>::S;
  if (skip(sc)) return;
  const S* b = static_cast<const S*>(bv);
  S* x = static_cast<S*>(xv);
  const int lane = threadIdx.x & (kWave - 1);
  const int row = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  if (row >= m) return;
  double acc = 0.0;
  for (int q = lane; q < m; q += kWave) acc = fma(ainv[(long long)row * m + q], ldd(b, map[q]), acc);
  acc = wave_sum(acc);
  if (lane == 0) x[map[row]] = store_as<S>(acc);
}

// kappa restriction for the hierarchy (full weighting inside, injection on the boundary).
__global__ void k_restrict_kappa(DevOp f, const double* kf, DevOp c, double* kc) {
  const bool d3 = c.nz > 1;
  for (long long cid = blockIdx.x * (long long)blockDim.x + threadIdx.x; cid < c.n;
       cid += (long long)gridDim.x * blockDim.x) {
    const int ic = static_cast<int>(cid % c.nx);
    const int jc = static_cast<int>((cid / c.nx) % c.ny);
    const int kcc = static_cast<int>(cid / c.sxy);
    const int fi = 2 * ic, fj = 2 * jc, fk = d3 ? 2 * kcc : 0;
    const long long fid = ((long long)fk * f.ny + fj) * f.nx + fi;
    if (is_boundary2(c, ic, jc) || (d3 && (kcc == 0 || kcc == c.nz - 1))) { kc[cid] = kf[fid]; continue; }
    const double w1[3] = {0.25, 0.5, 0.25};
    double acc = 0.0;
    for (int dk = (d3 ? -1 : 0); dk <= (d3 ? 1 : 0); ++dk)
      for (int dj = -1; dj <= 1; ++dj)
        for (int di = -1; di <= 1; ++di)
          acc += w1[di + 1] * w1[dj + 1] * (d3 ? w1[dk + 1] : 1.0) *
                 kf[((long long)(fk + dk) * f.ny + (fj + dj)) * f.nx + fi + di];
    kc[cid] = acc;
  }
}

// ----------------------------------------------------------------------------- vectors
template <Prec P>
__global__ void k_fill(void* xv, long long n, double v) {
  using S = typename PT<P>::S;
  S* x = static_cast<S*>(xv);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    x[i] = store_as<S>(v);
}

template <Prec PI, Prec PO>
__global__ void k_convert(const void* inv, void* outv, long long n, double scale) {
  using SI = typename PT<PI>::S;
  using SO = typename PT<PO>::S;
  const SI* in = static_cast<const SI*>(inv);
  SO* out = static_cast<SO*>(outv);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    out[i] = store_as<SO>(ldd(in, i) * scale);
}

// rho = 1/kappa in FP32 (the conductivity form the FP32/FP16 operators store).
__global__ void k_recip_f32(const double* in, float* out, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    out[i] = static_cast<float>(1.0 / in[i]);
}

template <Prec P>
__global__ void __launch_bounds__(kBlock) k_dot(const void* av, const void* bv, long long n,
                                               double* partials, const KScalars* sc) {
  using S = typename PT<P>::S;
  __shared__ double lds[kBlock / kWave];
  if (skip(sc)) return;
  const S* a = static_cast<const S*>(av);
  const S* b = static_cast<const S*>(bv);
  double acc = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    acc += ldd(a, i) * ldd(b, i);
  const double s = block_sum(acc, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

template <Prec P>
__global__ void __launch_bounds__(kBlock) k_cg_update(void* xv, void* rv, const void* pvv,
                                                     const void* qv, long long n, double* partials,
                                                     double* maxpart, const KScalars* sc,
                                                     const int* fresh) {
  using T = typename PT<P>::T;
  using S = typename PT<P>::S;
  __shared__ double lds[kBlock / kWave];
  if (skip(sc)) return;
  S* x = static_cast<S*>(xv);
  S* r = static_cast<S*>(rv);
  const S* p = static_cast<const S*>(pvv);
  const S* q = static_cast<const S*>(qv);
  const T alpha = T(sc->alpha);
  const bool void_x = fresh && *fresh;   // x holds a correction already folded in: take it as 0
  double acc = 0.0, mx = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const T x0 = void_x ? T(0) : T(ldd(x, i));
    const T xi = x0 + alpha * T(ldd(p, i));
    const T ri = T(ldd(r, i)) - alpha * T(ldd(q, i));
    x[i] = store_as<S>(xi);
    r[i] = store_as<S>(ri);
    acc += static_cast<double>(ri) * static_cast<double>(ri);
    mx = nmax(mx, fabs(static_cast<double>(ri)));
  }
  const double s = block_sum(acc, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
  if (maxpart) {
    const double m = block_max(mx, lds);
    if (threadIdx.x == 0) maxpart[blockIdx.x] = m;
  }
}

template <Prec P>
__global__ void k_xpay(const void* zv, void* pvv, long long n, const KScalars* sc) {
  using T = typename PT<P>::T;
  using S = typename PT<P>::S;
  if (skip(sc)) return;
  const S* z = static_cast<const S*>(zv);
  S* p = static_cast<S*>(pvv);
  const T beta = T(sc->beta);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = store_as<S>(fma(beta, T(ldd(p, i)), T(ldd(z, i))));   // as the fused apply forms it
}

template <Prec P>
__global__ void k_axpy_dev(void* yv, const void* xv, long long n, const double* a, double sign,
                           const KScalars* sc) {
  using T = typename PT<P>::T;
  using S = typename PT<P>::S;
  if (skip(sc)) return;
  S* y = static_cast<S*>(yv);
  const S* x = static_cast<const S*>(xv);
  const T av = T(sign * (a ? *a : 1.0));
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    y[i] = store_as<S>(T(ldd(y, i)) + av * T(ldd(x, i)));
}

// Stage 2 of every reduction: one block sums the partials in a fixed order (deterministic
// for a fixed launch shape) and updates the solve's scalars.
__global__ void __launch_bounds__(1024) k_finalize(KScalars* sc, const double* partials, int np,
                                                  int op, double tol, int tol_mode, double* slot) {
  __shared__ double lds[1024 / kWave];
  if (sc && (sc->done | sc->breakdown) && op != FIN_STORE && op != FIN_MAX) return;
  const bool is_max = (op == FIN_MAX || op == FIN_RELAX);
  double acc = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x)
    acc = is_max ? nmax(acc, partials[i]) : acc + partials[i];
  const double s = is_max ? block_max(acc, lds) : block_sum(acc, lds);
  if (threadIdx.x != 0) return;
  switch (op) {
    case FIN_CG_ALPHA:
      sc->pq = s;
      if (fabs(s) < sc->eps) { sc->breakdown = 1; break; }
      sc->alpha = sc->rz / s;
      break;
    case FIN_CG_RR:
      sc->rr = s;
      sc->metric = sqrt(s);
      sc->iter += 1;
      if (nonfinite(s)) { sc->breakdown = 1; break; }
      if (sc->metric < sc->target) sc->done = 1;
      break;
    case FIN_CG_BETA:
      sc->rz_old = sc->rz;
      sc->beta = s / sc->rz;
      sc->rz = s;
      break;
    case FIN_CG_INIT:
      sc->rz = s;
      break;
    case FIN_R0: {
      const double r0 = sqrt(s);
      if (nonfinite(s)) { sc->breakdown = 1; break; }
      sc->r0 = r0;
      sc->metric = r0;
      sc->target = tol_mode == PDEHIP_TOL_REL ? tol * r0 : tol;
      if (r0 < sc->target || r0 == 0.0) sc->done = 1;
      break;
    }
    case FIN_STORE:
      *slot = s;
      break;
    case FIN_MAX:
      *slot = s;
      break;
    case FIN_RELAX:
      if (slot) *slot = s;
      sc->metric = s;
      sc->iter += 1;
      if (s < tol) sc->done = 1;
      break;
  }
}

}  // namespace

// ============================================================================== launchers
template <typename S>
void closures_nr(const Launch& L, const DevOp& op, const FaceData& fd, S* u) {
  const int axes = op.nz > 1 ? 3 : 2;
  for (int a = 0; a < axes; ++a) {
    if (fd.kind[2 * a] == PDEHIP_BC_DIRICHLET && fd.kind[2 * a + 1] == PDEHIP_BC_DIRICHLET) continue;
    hipLaunchKernelGGL(k_closure_nr<S>, dim3(grid1(face_size(op, a), 1024)), dim3(kBlock), 0,
                       L.stream, op, fd, a, u, L.sc);
  }
}

template <typename S>
void closures_dirichlet(const Launch& L, const DevOp& op, const FaceData& fd, S* u) {
  const int axes = op.nz > 1 ? 3 : 2;
  for (int a = 0; a < axes; ++a) {
    if (fd.kind[2 * a] != PDEHIP_BC_DIRICHLET && fd.kind[2 * a + 1] != PDEHIP_BC_DIRICHLET) continue;
    hipLaunchKernelGGL(k_closure_d<S>, dim3(grid1(face_size(op, a), 1024)), dim3(kBlock), 0,
                       L.stream, op, fd, a, u, L.sc);
  }
}
template void closures_nr<double>(const Launch&, const DevOp&, const FaceData&, double*);
template void closures_nr<float>(const Launch&, const DevOp&, const FaceData&, float*);
template void closures_dirichlet<double>(const Launch&, const DevOp&, const FaceData&, double*);
template void closures_dirichlet<float>(const Launch&, const DevOp&, const FaceData&, float*);

#define PDEHIP_BY_PREC(p, KER, ...)                                                   \
  switch (p) {                                                                        \
    case P64: hipLaunchKernelGGL(KER<P64>, __VA_ARGS__); break;                       \
    case P32: hipLaunchKernelGGL(KER<P32>, __VA_ARGS__); break;                       \
    default: hipLaunchKernelGGL(KER<P16>, __VA_ARGS__); break;                        \
  }

namespace {
bool g_restrict_march = true;

// The plane-march restriction: chunks of at least 4 coarse planes, about two march workgroups
// per CU; the per-node workgroups after them, at most 1024.
template <Prec PF, Prec PC>
void launch_restrict3m(const Launch& L, const DevOp& fine, const void* rf, const DevOp& coarse,
                       void* bc, double scale, const double* dscale) {
  const int bx = coarse.nx - 4, by = coarse.ny - 4;
  const int ntx = (bx + kRMCols - 1) / kRMCols;
  const int nty = (by + kRMRows * kRMWaves - 1) / (kRMRows * kRMWaves);
  const int nzi = coarse.nz - 2;
  const long long txy = (long long)ntx * nty;
  const long long want = std::max(1LL, (2LL * device_cus() + txy - 1) / txy);
  const int kchunk = std::max(4, static_cast<int>((nzi + want - 1) / want));
  const int chunks = (nzi + kchunk - 1) / kchunk;
  const int nmarch = static_cast<int>(txy * chunks);
  const long long plane = (long long)coarse.nx * coarse.ny;
  const long long nrim_items = 2 * plane + (long long)nzi * (4LL * coarse.nx + 4LL * by);
  const int nrb = static_cast<int>(std::min<long long>((nrim_items + 255) / 256, 1024));
  hipLaunchKernelGGL((k_restrict3m<PF, PC>), dim3(nmarch + nrb), dim3(kRMLanes * kRMWaves), 0,
                     L.stream, fine, rf, coarse, bc, scale, dscale, L.sc, ntx, nty, kchunk, nmarch,
                     nrim_items);
}
}  // namespace

void set_restrict_march(bool on) { g_restrict_march = on; }
bool restrict_march() { return g_restrict_march; }

void restrict_fw(const Launch& L, const DevOp& fine, Prec pf, const void* rf, const DevOp& coarse,
                 Prec pc, void* bc, double scale, const double* dev_scale) {
  const int gx = (coarse.nx + kTX - 1) / kTX, gy = (coarse.ny + kTY - 1) / kTY;
  if (g_restrict_march && coarse.nz >= 3 && coarse.nx >= 5 && coarse.ny >= 5 &&
      2 * coarse.nx - 1 == fine.nx && 2 * coarse.ny - 1 == fine.ny && 2 * coarse.nz - 1 == fine.nz) {
#define RM_(A, B) launch_restrict3m<A, B>(L, fine, rf, coarse, bc, scale, dev_scale)
    if (pf == P64 && pc == P64) RM_(P64, P64);
    else if (pf == P32 && pc == P32) RM_(P32, P32);
    else if (pf == P32 && pc == P16) RM_(P32, P16);
    else if (pf == P16 && pc == P16) RM_(P16, P16);
    else if (pf == P64 && pc == P32) RM_(P64, P32);
    else RM_(P16, P32);
#undef RM_
    return;
  }
  if (coarse.nz > 1) {
    // interior box blocks, then rim blocks (rim_node numbering) stacked above them in y
    const int bx = coarse.nx - 4, by = coarse.ny - 4;
    const bool box = bx > 0 && by > 0;
    const int gxi = box ? (bx + kTX - 1) / kTX : 1;
    const int gy_in = box ? (by + kTY - 1) / kTY : 0;
    const long long rim = box ? 4LL * coarse.nx + 4LL * by : (long long)coarse.nx * coarse.ny;
    const long long per_row = (long long)gxi * kTX * kTY;
    const int gy_rim = static_cast<int>((rim + per_row - 1) / per_row);
    const dim3 g3(gxi, gy_in + gy_rim, coarse.nz);
    const dim3 b3(kTX, kTY);
#define R3_(A, B) hipLaunchKernelGGL((k_restrict3<A, B>), g3, b3, 0, L.stream, fine, rf, coarse, bc, scale, dev_scale, L.sc, gy_in)
    if (pf == P64 && pc == P64) R3_(P64, P64);
    else if (pf == P32 && pc == P32) R3_(P32, P32);
    else if (pf == P32 && pc == P16) R3_(P32, P16);
    else if (pf == P16 && pc == P16) R3_(P16, P16);
    else if (pf == P64 && pc == P32) R3_(P64, P32);
    else R3_(P16, P32);
#undef R3_
    return;
  }
  const int zc = coarse.nz > 1 ? zchunk_for(gx, gy, coarse.nz) : 1;
  const dim3 g(gx, gy, coarse.nz > 1 ? (coarse.nz + zc - 1) / zc : 1);
  const dim3 blk(kTX, kTY);
#define R_(A, B) hipLaunchKernelGGL((k_restrict<A, B>), g, blk, 0, L.stream, fine, rf, coarse, bc, scale, dev_scale, zc, L.sc)
  if (pf == P64 && pc == P64) R_(P64, P64);
  else if (pf == P32 && pc == P32) R_(P32, P32);
  else if (pf == P32 && pc == P16) R_(P32, P16);
  else if (pf == P16 && pc == P16) R_(P16, P16);
  else if (pf == P64 && pc == P32) R_(P64, P32);
  else R_(P16, P32);
#undef R_
}

void prolong_add(const Launch& L, const DevOp& coarse, Prec pc, const void* uc, const DevOp& fine,
                 Prec pf, void* uf, double factor, const double* dev_scale) {
  const int gx = (fine.nx + kTX - 1) / kTX, gy = (fine.ny + kTY - 1) / kTY;
  if (fine.nz > 1) {
    const int ncz = (fine.nz + 1) / 2;   // coarse intervals (the last one holds no interior plane)
    const dim3 g3(gx, gy, (ncz + kProlongZ - 1) / kProlongZ);
    const dim3 b3(kTX, kTY);
#define P3_(A, B) hipLaunchKernelGGL((k_prolong3<A, B>), g3, b3, 0, L.stream, coarse, uc, fine, uf, factor, dev_scale, L.sc)
    if (pc == P64 && pf == P64) P3_(P64, P64);
    else if (pc == P32 && pf == P32) P3_(P32, P32);
    else if (pc == P16 && pf == P32) P3_(P16, P32);
    else if (pc == P16 && pf == P16) P3_(P16, P16);
    else if (pc == P32 && pf == P64) P3_(P32, P64);
    else P3_(P32, P16);
#undef P3_
    return;
  }
  const int zc = fine.nz > 1 ? zchunk_for(gx, gy, fine.nz) : 1;
  const dim3 g(gx, gy, fine.nz > 1 ? (fine.nz + zc - 1) / zc : 1);
  const dim3 blk(kTX, kTY);
#define P_(A, B) hipLaunchKernelGGL((k_prolong<A, B>), g, blk, 0, L.stream, coarse, uc, fine, uf, factor, dev_scale, zc, L.sc)
  if (pc == P64 && pf == P64) P_(P64, P64);
  else if (pc == P32 && pf == P32) P_(P32, P32);
  else if (pc == P16 && pf == P32) P_(P16, P32);
  else if (pc == P16 && pf == P16) P_(P16, P16);
  else if (pc == P32 && pf == P64) P_(P32, P64);
  else P_(P32, P16);
#undef P_
}

void coarse_dense(const Launch& L, Prec p, const double* ainv, const int* map, int m, const void* b,
                  void* x) {
  const int rows_per_block = kBlock / kWave;
  const dim3 g((m + rows_per_block - 1) / rows_per_block);
  PDEHIP_BY_PREC(p, k_coarse_dense, g, dim3(kBlock), 0, L.stream, ainv, map, m, b, x, L.sc);
}

void restrict_kappa(hipStream_t s, const DevOp& fine, const double* kf, const DevOp& coarse,
                    double* kc) {
  hipLaunchKernelGGL(k_restrict_kappa, dim3(grid1(coarse.n)), dim3(kBlock), 0, s, fine, kf, coarse, kc);
}

int vec_blocks(long long n) { return static_cast<int>(grid1(n, 2048)); }

void fill(hipStream_t s, Prec p, void* x, long long n, double v) {
  const dim3 g(grid1(n));
  PDEHIP_BY_PREC(p, k_fill, g, dim3(kBlock), 0, s, x, n, v);
}

void convert(hipStream_t s, Prec pin, const void* in, Prec pout, void* out, long long n, double scale) {
  const dim3 g(grid1(n));
#define C_(A, B) hipLaunchKernelGGL((k_convert<A, B>), g, dim3(kBlock), 0, s, in, out, n, scale)
  if (pin == P64 && pout == P32) C_(P64, P32);
  else if (pin == P32 && pout == P64) C_(P32, P64);
  else if (pin == P64 && pout == P64) C_(P64, P64);
  else if (pin == P32 && pout == P32) C_(P32, P32);
  else if (pin == P64 && pout == P16) C_(P64, P16);
  else C_(P16, P64);
#undef C_
}

void reciprocal_f32(hipStream_t s, const double* in, float* out, long long n) {
  hipLaunchKernelGGL(k_recip_f32, dim3(grid1(n)), dim3(kBlock), 0, s, in, out, n);
}

void dot(const Launch& L, Prec p, const void* a, const void* b, long long n, double* partials) {
  const dim3 g(vec_blocks(n));
  PDEHIP_BY_PREC(p, k_dot, g, dim3(kBlock), 0, L.stream, a, b, n, partials, L.sc);
}

void finalize(const Launch& L, KScalars* sc, const double* partials, int nparts, int op, double tol,
              int tol_mode, double* out_slot) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, L.stream, sc, partials, nparts, op, tol,
                     tol_mode, out_slot);
}

void cg_update(const Launch& L, Prec p, void* x, void* r, const void* pv, const void* q, long long n,
               double* partials, double* maxpart, const int* fresh) {
  const dim3 g(vec_blocks(n));
  PDEHIP_BY_PREC(p, k_cg_update, g, dim3(kBlock), 0, L.stream, x, r, pv, q, n, partials, maxpart,
                 L.sc, fresh);
}

void cg_xpay(const Launch& L, Prec p, const void* z, void* pv, long long n) {
  const dim3 g(vec_blocks(n));
  PDEHIP_BY_PREC(p, k_xpay, g, dim3(kBlock), 0, L.stream, z, pv, n, L.sc);
}

void axpy_dev(const Launch& L, Prec p, void* y, const void* x, long long n, const double* a,
              double sign) {
  const dim3 g(vec_blocks(n));
  PDEHIP_BY_PREC(p, k_axpy_dev, g, dim3(kBlock), 0, L.stream, y, x, n, a, sign, L.sc);
}

}  // namespace pdehip
