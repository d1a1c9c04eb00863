// k_krylov.hip — BiCGStab and GMRES vector kernels with device-resident scalars.
//
// BiCGStab follows the reference CPU algorithm (src/backends/cpu/solvers/krylov.cpp:82-185),
// including its recomputation r = b - A x every iteration, with an optional right
// preconditioner.  GMRES follows krylov.cpp:187-333 (restarted, Givens rotations with the CPU
// sign convention) but orthogonalises by classical Gram-Schmidt applied twice (CGS2): each
// pass is one multi-dot and one multi-axpy launch instead of 2(j+1) launches of modified
// Gram-Schmidt; the resulting basis agrees with MGS to rounding.
#include "kernels_krylov.hpp"

namespace pdehip {
namespace {

__device__ __forceinline__ bool skip(const KScalars* sc) { return sc->done | sc->breakdown; }

int blocks_for(long long n) {
  long long b = (n + kBlock - 1) / kBlock;
  if (b < 1) b = 1;
  return static_cast<int>(b > 1024 ? 1024 : b);
}

// ----------------------------------------------------------------------------- BiCGStab
// Vector element type S is double (FP64 path) or float (mixed path, FP32 Krylov vectors);
// arithmetic and reductions are in double.  In the mixed path the solution stays FP64 and the
// x updates take a scale (the FP32 vectors hold the residual divided by it).
// x += scale*alpha*p_hat when ||s|| already met the target (krylov.cpp:143-148).
template <typename S>
__global__ void k_bicg_xs(double* __restrict__ x, const S* __restrict__ ph, long long n,
                          const KScalars* sc, const double* scale) {
  if (skip(sc) || !sc->flag_s) return;
  const double alpha = sc->alpha * (scale ? *scale : 1.0);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    x[i] += alpha * double(ph[i]);
}

template <typename S>
__global__ void __launch_bounds__(kBlock) k_dot2(const S* __restrict__ a,
                                                const S* __restrict__ b,
                                                const S* __restrict__ c, long long n,
                                                double* p1, double* p2, const KScalars* sc) {
  __shared__ double lds[kBlock / kWave];
  if (skip(sc)) return;
  double ab = 0.0, ac = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    ab += double(a[i]) * double(b[i]);
    ac += double(a[i]) * double(c[i]);
  }
  const double s1 = block_sum(ab, lds);
  const double s2 = block_sum(ac, lds);
  if (threadIdx.x == 0) { p1[blockIdx.x] = s1; p2[blockIdx.x] = s2; }
}

template <typename S>
__global__ void k_bicg_x(double* __restrict__ x, const S* __restrict__ ph,
                         const S* __restrict__ sh, long long n, const KScalars* sc,
                         const double* scale) {
  if (skip(sc)) return;
  const double f = scale ? *scale : 1.0;
  const double alpha = sc->alpha * f, omega = sc->omega * f;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    x[i] += alpha * double(ph[i]) + omega * double(sh[i]);
}

// Mixed path: r = s - omega t (the recurrence), partials of r.r.
__global__ void __launch_bounds__(kBlock) k_bicg_r(const float* __restrict__ sv,
                                                  const float* __restrict__ t,
                                                  float* __restrict__ r, long long n,
                                                  double* partials, const KScalars* sc) {
  __shared__ double lds[kBlock / kWave];
  if (skip(sc)) return;
  const double omega = sc->omega;
  double acc = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float ri = static_cast<float>(double(sv[i]) - omega * double(t[i]));
    r[i] = ri;
    acc += double(ri) * double(ri);
  }
  const double bs = block_sum(acc, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = bs;
}

// Jacobi preconditioner: out = in / diag on interior rows (folded diagonal), identity on
// boundary rows.
template <int DIM, int KIND, typename S>
__global__ void k_jacobi_pc(DevOp op, const S* __restrict__ in, S* __restrict__ out,
                            const KScalars* sc) {
  if (sc && skip(sc)) return;
  for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < op.n;
       id += (long long)gridDim.x * blockDim.x) {
    const IJK c = coords(op, id);
    const int i = c.i, j = c.j, k = c.k;
    if (is_boundary2(op, i, j) || (DIM == 3 && (k == 0 || k == op.nz - 1))) { out[id] = in[id]; continue; }
    W7<double> w = row_weights<DIM, KIND, double, double>(op, id);
    fold_center<DIM, double>(op, i, j, k, w);
    out[id] = static_cast<S>(double(in[id]) / w.C);
  }
}

// ----------------------------------------------------------------------------- GMRES
// partials[i*nb + block] = V_i . w for i < cnt (V rows of length ld).
template <typename S>
__global__ void __launch_bounds__(kBlock) k_multidot(const S* __restrict__ V, long long ld,
                                                    int cnt, const S* __restrict__ w,
                                                    long long n, double* partials,
                                                    const KScalars* sc) {
  __shared__ double lds[kBlock / kWave];
  if (skip(sc) || sc->cycle_done) return;
  // block (x, y): element blocks x of basis vectors 8y .. 8y+7, so every group of eight runs
  // in parallel (gridDim.y = ceil(cnt / 8))
  const int nb = gridDim.x;
  const int i0 = blockIdx.y * 8;
  if (i0 >= cnt) return;
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const int m = min(8, cnt - i0);
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x) {
    const double wv = double(w[q]);
#pragma unroll
    for (int t = 0; t < 8; ++t)
      if (t < m) acc[t] += double(V[(long long)(i0 + t) * ld + q]) * wv;
  }
  for (int t = 0; t < m; ++t) {
    const double s = block_sum(acc[t], lds);
    if (threadIdx.x == 0) partials[(long long)(i0 + t) * nb + blockIdx.x] = s;
  }
}

// w -= sum_i V_i h_i (h from device memory); also used with sign +1 for x updates.
template <typename S>
__global__ void k_multiaxpy(const S* __restrict__ V, long long ld, int cnt,
                            const double* __restrict__ h, S* __restrict__ w, long long n,
                            double sign, int check_cycle, const KScalars* sc) {
  if (skip(sc) || (check_cycle && sc->cycle_done)) return;
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int i = 0; i < cnt; ++i) acc += h[i] * double(V[(long long)i * ld + q]);
    w[q] = static_cast<S>(double(w[q]) + sign * acc);
  }
}

// Sum the multi-dot partials into hcol[i] (accumulate when add != 0).
// One block per dot product (blockIdx.x = i), each summing its nb partials in a fixed order.
__global__ void __launch_bounds__(256) k_fin_multi(const double* partials, int nb, int cnt,
                                                  double* hcol, int add, const KScalars* sc) {
  __shared__ double lds[256 / kWave];
  if (skip(sc) || sc->cycle_done) return;
  const int i = blockIdx.x;
  if (i >= cnt) return;
  double acc = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) acc += partials[(long long)i * nb + b];
  const double s = block_sum(acc, lds);
  if (threadIdx.x == 0) hcol[i] = add ? hcol[i] + s : s;
}

// H[j+1][j] = ||w||, V_{j+1} = w / H[j+1][j]
template <typename S>
__global__ void k_gm_normalize(S* __restrict__ vnext, const S* __restrict__ w,
                               long long n, const double* hval, const KScalars* sc) {
  if (skip(sc) || sc->cycle_done) return;
  const double h = *hval;
  if (!(h > 0.0)) return;
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x)
    vnext[q] = static_cast<S>(double(w[q]) / h);
}

__global__ void __launch_bounds__(256) k_fin_norm(const double* partials, int nb, double* out,
                                                 const KScalars* sc) {
  __shared__ double lds[256 / kWave];
  if (skip(sc) || sc->cycle_done) return;
  double acc = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) acc += partials[b];
  const double s = block_sum(acc, lds);
  if (threadIdx.x == 0) *out = sqrt(s);
}

// Givens step j (krylov.cpp:262-295): H is (m+1) x m row-major with leading dimension m.
__global__ void k_givens(double* H, double* cs, double* sn, double* g, int j, int m,
                         KScalars* sc) {
  if (skip(sc) || sc->cycle_done) return;
  for (int i = 0; i < j; ++i) {
    const double v1 = H[i * m + j], v2 = H[(i + 1) * m + j];
    H[i * m + j] = cs[i] * v1 - sn[i] * v2;
    H[(i + 1) * m + j] = sn[i] * v1 + cs[i] * v2;
  }
  const double h1 = H[(j + 1) * m + j], h0 = H[j * m + j];
  const double denom = sqrt(h0 * h0 + h1 * h1);
  if (denom < 1e-30) { sc->breakdown = 1; return; }
  cs[j] = h0 / denom;
  sn[j] = -h1 / denom;
  H[j * m + j] = cs[j] * h0 - sn[j] * h1;
  H[(j + 1) * m + j] = 0.0;
  const double g0 = g[j], g1 = g[j + 1];
  g[j] = cs[j] * g0 - sn[j] * g1;
  g[j + 1] = sn[j] * g0 + cs[j] * g1;
  sc->steps = j + 1;
  sc->iter += 1;
  sc->metric = fabs(g[j + 1]);
  if (fabs(g[j + 1]) < sc->target) sc->cycle_done = 1;
}

// Back substitution over the steps of the current cycle.
__global__ void k_backsub(const double* H, const double* g, double* y, int m, KScalars* sc) {
  if (skip(sc)) return;
  const int dim = sc->steps;
  for (int i = dim - 1; i >= 0; --i) {
    double sum = g[i];
    for (int q = i + 1; q < dim; ++q) sum -= H[i * m + q] * y[q];
    if (fabs(H[i * m + i]) < 1e-30) { sc->breakdown = 1; return; }
    y[i] = sum / H[i * m + i];
  }
  for (int i = dim; i < m; ++i) y[i] = 0.0;
}

// New cycle: V_0 = r / beta, g = [beta, 0, ...]; clears the cycle flags.
template <typename S>
__global__ void k_gm_start(S* __restrict__ v0, const double* __restrict__ r, long long n,
                           double* g, int m, KScalars* sc) {
  if (skip(sc)) return;
  const double beta = sc->metric;
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x)
    v0[q] = static_cast<S>(r[q] / beta);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    // cycle flags are reset by k_gm_reset before this kernel's first use
    g[0] = beta;
    for (int i = 1; i <= m; ++i) g[i] = 0.0;
  }
}

// H[i][j] = hc[i] + hc[stride + i]: the two CGS passes' coefficients staged stride apart.
__global__ void k_gm_setcol(double* H, const double* hc, int stride, int j, int m, int cnt,
                            const KScalars* sc) {
  if (skip(sc) || sc->cycle_done) return;
  for (int i = threadIdx.x; i < cnt; i += blockDim.x) H[i * m + j] = hc[i] + hc[stride + i];
}

__global__ void k_gm_reset(KScalars* sc) {
  if (threadIdx.x == 0) { sc->cycle_done = 0; sc->steps = 0; }
}

// The scalar step after the BiCGStab / GMRES reductions (s1, s2: the reduced sums).
__device__ void fin_apply(KScalars* sc, double s1, double s2, int op) {
  switch (op) {
    case KFIN_BI_RHO:
      sc->rho = s1;
      if (fabs(s1) < sc->eps) { sc->breakdown = 1; break; }
      if (sc->iter > 0) sc->beta = (s1 / sc->rho_old) * (sc->alpha / sc->omega);
      break;
    case KFIN_BI_ALPHA:
      if (fabs(s1) < sc->eps) { sc->breakdown = 1; break; }
      sc->alpha = sc->rho / s1;
      break;
    case KFIN_BI_S:
      sc->snorm = sqrt(s1);
      if (nonfinite(s1)) { sc->breakdown = 1; break; }
      if (sc->snorm < sc->target) sc->flag_s = 1;
      break;
    case KFIN_BI_OMEGA:
      sc->tt = s1;
      sc->ts = s2;
      if (fabs(s1) < sc->eps) { sc->breakdown = 1; break; }
      sc->omega = s2 / s1;
      break;
    case KFIN_BI_R:
      sc->metric = sqrt(s1);
      sc->iter += 1;
      if (nonfinite(s1)) { sc->breakdown = 1; break; }
      if (sc->metric < sc->target) { sc->done = 1; break; }
      if (fabs(sc->omega) < sc->eps) { sc->breakdown = 1; break; }
      sc->rho_old = sc->rho;
      break;
    case KFIN_GM_R:
      sc->metric = sqrt(s1);
      if (nonfinite(s1)) { sc->breakdown = 1; break; }
      if (sc->metric < sc->target) sc->done = 1;
      break;
  }
}

// Stage-2 reductions for BiCGStab and the true-residual checks.
__global__ void __launch_bounds__(1024) k_fin_k(KScalars* sc, const double* p1, const double* p2,
                                               int np, int op) {
  __shared__ double lds[1024 / kWave];
  if (skip(sc)) return;
  double a1 = 0.0, a2 = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) {
    a1 += p1[i];
    if (p2) a2 += p2[i];
  }
  const double s1 = block_sum(a1, lds);
  const double s2 = p2 ? block_sum(a2, lds) : 0.0;
  if (threadIdx.x == 0) fin_apply(sc, s1, s2, op);
}

// Jacobi scaling of one node: in / folded diagonal on interior rows, identity on boundary rows
// (the k_jacobi_pc rule).
template <int DIM, int KIND, typename S>
__device__ __forceinline__ S jacobi_node(const DevOp& op, long long id, S in) {
  const IJK c = coords(op, id);
  if (is_boundary2(op, c.i, c.j) || (DIM == 3 && (c.k == 0 || c.k == op.nz - 1))) return in;
  W7<double> w = row_weights<DIM, KIND, double, double>(op, id);
  fold_center<DIM, double>(op, c.i, c.j, c.k, w);
  return static_cast<S>(double(in) / w.C);
}

// p = r + beta (p - omega v) and, with the Jacobi preconditioner, ph = D^-1 p in the same pass.
template <int DIM, int KIND, bool PC, typename S>
__global__ void __launch_bounds__(kBlock) k_bicg_p_pc(DevOp op, const S* __restrict__ r, S* __restrict__ p,
                                                     const S* __restrict__ v, S* __restrict__ ph,
                                                     const KScalars* sc) {
  if (skip(sc)) return;
  const bool first = sc->iter == 0;
  const double beta = sc->beta, omega = sc->omega;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < op.n;
       i += (long long)gridDim.x * blockDim.x) {
    const S pi = static_cast<S>(first ? double(r[i]) : double(r[i]) + beta * (double(p[i]) - omega * double(v[i])));
    p[i] = pi;
    if constexpr (PC) ph[i] = jacobi_node<DIM, KIND, S>(op, i, pi);
  }
}

// s = r - alpha v with partials of s.s and, with the Jacobi preconditioner, sh = D^-1 s in the
// same pass.
template <int DIM, int KIND, bool PC, typename S>
__global__ void __launch_bounds__(kBlock) k_bicg_s_pc(DevOp op, const S* __restrict__ r,
                                                     const S* __restrict__ v, S* __restrict__ sv,
                                                     S* __restrict__ sh, double* partials, const KScalars* sc) {
  __shared__ double lds[kBlock / kWave];
  if (skip(sc)) return;
  const double alpha = sc->alpha;
  double acc = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < op.n;
       i += (long long)gridDim.x * blockDim.x) {
    const S si = static_cast<S>(double(r[i]) - alpha * double(v[i]));
    sv[i] = si;
    if constexpr (PC) sh[i] = jacobi_node<DIM, KIND, S>(op, i, si);
    acc += double(si) * double(si);
  }
  const double t = block_sum(acc, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

__global__ void k_bi_sdone(KScalars* sc) {
  if (threadIdx.x == 0 && !(sc->done | sc->breakdown) && sc->flag_s) {
    sc->metric = sc->snorm;
    sc->iter += 1;
    sc->done = 1;
  }
}

}  // namespace

int krylov_blocks(long long n) { return blocks_for(n); }

template <typename S>
void bicg_xs(hipStream_t s, double* x, const S* ph, long long n, const KScalars* sc, const double* scale) {
  hipLaunchKernelGGL(k_bicg_xs<S>, dim3(blocks_for(n)), dim3(kBlock), 0, s, x, ph, n, sc, scale);
}
template <typename S>
void dot2(hipStream_t s, const S* a, const S* b, const S* c, long long n, double* p1, double* p2, const KScalars* sc) {
  hipLaunchKernelGGL(k_dot2<S>, dim3(blocks_for(n)), dim3(kBlock), 0, s, a, b, c, n, p1, p2, sc);
}
template <typename S>
void bicg_x(hipStream_t s, double* x, const S* ph, const S* sh, long long n, const KScalars* sc, const double* scale) {
  hipLaunchKernelGGL(k_bicg_x<S>, dim3(blocks_for(n)), dim3(kBlock), 0, s, x, ph, sh, n, sc, scale);
}
template void bicg_xs<double>(hipStream_t, double*, const double*, long long, const KScalars*, const double*);
template void bicg_xs<float>(hipStream_t, double*, const float*, long long, const KScalars*, const double*);
template void dot2<double>(hipStream_t, const double*, const double*, const double*, long long, double*, double*, const KScalars*);
template void dot2<float>(hipStream_t, const float*, const float*, const float*, long long, double*, double*, const KScalars*);
template void bicg_x<double>(hipStream_t, double*, const double*, const double*, long long, const KScalars*, const double*);
template void bicg_x<float>(hipStream_t, double*, const float*, const float*, long long, const KScalars*, const double*);
void bicg_r(hipStream_t s, const float* sv, const float* t, float* r, long long n, double* partials, const KScalars* sc) {
  hipLaunchKernelGGL(k_bicg_r, dim3(blocks_for(n)), dim3(kBlock), 0, s, sv, t, r, n, partials, sc);
}
void bi_sdone(hipStream_t s, KScalars* sc) { hipLaunchKernelGGL(k_bi_sdone, dim3(1), dim3(64), 0, s, sc); }
void fin_k(hipStream_t s, KScalars* sc, const double* p1, const double* p2, int np, int op) {
  hipLaunchKernelGGL(k_fin_k, dim3(1), dim3(1024), 0, s, sc, p1, p2, np, op);
}

// Kernel variants by (dimension, operator kind) when the Jacobi scaling is fused in.
#define PC_DISPATCH(KER, ...)                                                                  \
  do {                                                                                         \
    const dim3 g_(blocks_for(op.n)), b_(kBlock);                                               \
    const bool d3_ = op.nz > 1;                                                                \
    if (!pc) hipLaunchKernelGGL((KER<2, PDEHIP_OP_CONST, false, S>), g_, b_, 0, s, __VA_ARGS__); \
    else if (op.kind == PDEHIP_OP_POINTWISE) {                                                 \
      if (d3_) hipLaunchKernelGGL((KER<3, PDEHIP_OP_POINTWISE, true, S>), g_, b_, 0, s, __VA_ARGS__); \
      else hipLaunchKernelGGL((KER<2, PDEHIP_OP_POINTWISE, true, S>), g_, b_, 0, s, __VA_ARGS__); \
    } else if (op.kind == PDEHIP_OP_FLUX) {                                                    \
      if (d3_) hipLaunchKernelGGL((KER<3, PDEHIP_OP_FLUX, true, S>), g_, b_, 0, s, __VA_ARGS__); \
      else hipLaunchKernelGGL((KER<2, PDEHIP_OP_FLUX, true, S>), g_, b_, 0, s, __VA_ARGS__);   \
    } else {                                                                                   \
      if (d3_) hipLaunchKernelGGL((KER<3, PDEHIP_OP_CONST, true, S>), g_, b_, 0, s, __VA_ARGS__); \
      else hipLaunchKernelGGL((KER<2, PDEHIP_OP_CONST, true, S>), g_, b_, 0, s, __VA_ARGS__);  \
    }                                                                                          \
  } while (0)
template <typename S>
void bicg_p_pc(hipStream_t s, const DevOp& op, const S* r, S* p, const S* v, S* ph, bool pc, const KScalars* sc) {
  PC_DISPATCH(k_bicg_p_pc, op, r, p, v, ph, sc);
}
template <typename S>
void bicg_s_pc(hipStream_t s, const DevOp& op, const S* r, const S* v, S* sv, S* sh, bool pc,
               double* partials, const KScalars* sc) {
  PC_DISPATCH(k_bicg_s_pc, op, r, v, sv, sh, partials, sc);
}
#undef PC_DISPATCH
#define INST_(S)                                                                                          \
  template void bicg_p_pc<S>(hipStream_t, const DevOp&, const S*, S*, const S*, S*, bool, const KScalars*); \
  template void bicg_s_pc<S>(hipStream_t, const DevOp&, const S*, const S*, S*, S*, bool, double*, const KScalars*);
INST_(double)
INST_(float)
#undef INST_

template <typename S>
void jacobi_pc(hipStream_t s, const DevOp& op, const S* in, S* out, const KScalars* sc) {
  long long b = (op.n + kBlock - 1) / kBlock;
  if (b > 4096) b = 4096;
  const bool d3 = op.nz > 1;
#define J_(D, K) hipLaunchKernelGGL((k_jacobi_pc<D, K, S>), dim3(b), dim3(kBlock), 0, s, op, in, out, sc)
  if (op.kind == PDEHIP_OP_POINTWISE) { if (d3) J_(3, PDEHIP_OP_POINTWISE); else J_(2, PDEHIP_OP_POINTWISE); }
  else if (op.kind == PDEHIP_OP_FLUX) { if (d3) J_(3, PDEHIP_OP_FLUX); else J_(2, PDEHIP_OP_FLUX); }
  else { if (d3) J_(3, PDEHIP_OP_CONST); else J_(2, PDEHIP_OP_CONST); }
#undef J_
}
template void jacobi_pc<double>(hipStream_t, const DevOp&, const double*, double*, const KScalars*);
template void jacobi_pc<float>(hipStream_t, const DevOp&, const float*, float*, const KScalars*);

template <typename S>
void multidot(hipStream_t s, const S* V, long long ld, int cnt, const S* w, long long n, double* partials, int nb, const KScalars* sc) {
  hipLaunchKernelGGL(k_multidot<S>, dim3(nb, std::max(1, (cnt + 7) / 8)), dim3(kBlock), 0, s, V, ld,
                     cnt, w, n, partials, sc);
}
template <typename S>
void multiaxpy(hipStream_t s, const S* V, long long ld, int cnt, const double* h, S* w, long long n, double sign, int check_cycle, const KScalars* sc) {
  hipLaunchKernelGGL(k_multiaxpy<S>, dim3(blocks_for(n)), dim3(kBlock), 0, s, V, ld, cnt, h, w, n, sign, check_cycle, sc);
}
template void multidot<double>(hipStream_t, const double*, long long, int, const double*, long long, double*, int, const KScalars*);
template void multidot<float>(hipStream_t, const float*, long long, int, const float*, long long, double*, int, const KScalars*);
template void multiaxpy<double>(hipStream_t, const double*, long long, int, const double*, double*, long long, double, int, const KScalars*);
template void multiaxpy<float>(hipStream_t, const float*, long long, int, const double*, float*, long long, double, int, const KScalars*);
void fin_multi(hipStream_t s, const double* partials, int nb, int cnt, double* hcol, int add, const KScalars* sc) {
  hipLaunchKernelGGL(k_fin_multi, dim3(std::max(1, cnt)), dim3(256), 0, s, partials, nb, cnt, hcol,
                     add, sc);
}
template <typename S>
void gm_normalize(hipStream_t s, S* vnext, const S* w, long long n, const double* hval, const KScalars* sc) {
  hipLaunchKernelGGL(k_gm_normalize<S>, dim3(blocks_for(n)), dim3(kBlock), 0, s, vnext, w, n, hval, sc);
}
template void gm_normalize<double>(hipStream_t, double*, const double*, long long, const double*, const KScalars*);
template void gm_normalize<float>(hipStream_t, float*, const float*, long long, const double*, const KScalars*);
void fin_norm(hipStream_t s, const double* partials, int nb, double* out, const KScalars* sc) {
  hipLaunchKernelGGL(k_fin_norm, dim3(1), dim3(256), 0, s, partials, nb, out, sc);
}
void givens(hipStream_t s, double* H, double* cs, double* sn, double* g, int j, int m, KScalars* sc) {
  hipLaunchKernelGGL(k_givens, dim3(1), dim3(1), 0, s, H, cs, sn, g, j, m, sc);
}
void backsub(hipStream_t s, const double* H, const double* g, double* y, int m, KScalars* sc) {
  hipLaunchKernelGGL(k_backsub, dim3(1), dim3(1), 0, s, H, g, y, m, sc);
}
void gm_setcol(hipStream_t s, double* H, const double* hc, int stride, int j, int m, int cnt,
               const KScalars* sc) {
  hipLaunchKernelGGL(k_gm_setcol, dim3(1), dim3(64), 0, s, H, hc, stride, j, m, cnt, sc);
}
template <typename S>
void gm_start(hipStream_t s, S* v0, const double* r, long long n, double* g, int m, KScalars* sc) {
  hipLaunchKernelGGL(k_gm_reset, dim3(1), dim3(64), 0, s, sc);
  hipLaunchKernelGGL(k_gm_start<S>, dim3(blocks_for(n)), dim3(kBlock), 0, s, v0, r, n, g, m, sc);
}
template void gm_start<double>(hipStream_t, double*, const double*, long long, double*, int, KScalars*);
template void gm_start<float>(hipStream_t, float*, const double*, long long, double*, int, KScalars*);

}  // namespace pdehip
