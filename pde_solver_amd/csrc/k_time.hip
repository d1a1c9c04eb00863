// k_time.hip — explicit time-step stage kernels, the reference residual-norm kernel and the
// single-workgroup lexicographic Gauss-Seidel used by the reference-semantics V-cycle.
#include "faces_dev.hpp"
#include "kernels_ext.hpp"

namespace pdehip {
namespace {

// Transcendental terms out of line: their code would otherwise raise every stage kernel's
// register count (and lower its occupancy) whether or not a problem uses them.
__device__ __noinline__ double eval_nl_other(int kind, double coeff, double u) {
  switch (kind) {
    case PDEHIP_NL_SIN: return coeff * sin(u);
    case PDEHIP_NL_COS: return coeff * cos(u);
    case PDEHIP_NL_EXP: return coeff * exp(u);
    default: return coeff * fabs(u);
  }
}

__device__ __forceinline__ double eval_nl(const NonlinearTerms& nl, double u) {
  double v = 0.0;
  for (int t = 0; t < nl.n; ++t) {
    const pdehip_nonlinear& term = nl.t[t];
    if (term.kind == PDEHIP_NL_POWER) v += term.coeff * ipow(u, term.power);
    else v += eval_nl_other(term.kind, term.coeff, u);
  }
  return v;
}

__device__ __forceinline__ double safe_div(double num, double den) {  // safe_math.h:48-58
  if (fabs(den) < 1e-12) return 0.0;
  const double r = num / den;
  return isfinite(r) ? r : 0.0;
}

template <typename T> __device__ __forceinline__ double ldv(const T* p, long long i) { return static_cast<double>(p[i]); }

// du/dt = -(A v + f + N(v))/ut on interior nodes (fill_dudt, reference solver.cpp:1182-1195),
// zero on the boundary; then one of the stage combinations below.
template <int KIND, typename S>
__device__ __forceinline__ double dudt_at(const DevOp& op, const S* v, long long id, int i, int j,
                                         const double* f, double f_const, const NonlinearTerms& nl,
                                         double ut, double inv_ut) {
  if (is_boundary2(op, i, j)) return 0.0;
  const W7<double> w = row_weights<2, KIND, double, double>(op, id);
  const double vc = ldv(v, id);
  const double opv = w.C * vc + (w.E * ldv(v, id + 1) + w.W * ldv(v, id - 1) +
                                 w.N * ldv(v, id + op.nx) + w.S * ldv(v, id - op.nx));
  const double rhs = opv + (f ? f[id] : f_const) + (nl.n ? eval_nl(nl, vc) : 0.0);
  // safe_div(-rhs, ut) with the reciprocal of the (uniform) u_t coefficient formed once
  if (fabs(ut) < 1e-12) return 0.0;
  const double r = -rhs * inv_ut;
  return isfinite(r) ? r : 0.0;
}

// Stage kernel.  Per node:  k = F(v)
//   mode 0 (FE / RK4 s1 / RK2 s1 / SSP s1): acc = k (if acc), out = un + c1*k
//   mode 1 (RK4 s2, s3):                      acc += 2k,        out = un + c1*k
//   mode 2 (RK4 s4):                          out = un + c1*(acc + k)
//   mode 3 (RK2 s2):                          out = un + c1*(acc + k)     (c1 = dt/2)
//   mode 4 (SSPRK2 s2):                       out = 0.5*un + 0.5*(v + dt*k)
//   mode 5 (SSPRK3 s2):                       out = 0.75*un + 0.25*(v + dt*k)
//   mode 6 (SSPRK3 s3):                       out = (1/3)*un + (2/3)*(v + dt*k)
template <int KIND, typename S>
__global__ void __launch_bounds__(kBlock) k_stage(DevOp op, const S* __restrict__ v,
                                                 const S* __restrict__ un, S* __restrict__ acc,
                                                 S* __restrict__ out, const double* f,
                                                 double f_const, NonlinearTerms nl, double ut,
                                                 double dt, double c1, int mode) {
  const double inv_ut = 1.0 / ut;
  for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < op.n;
       id += (long long)gridDim.x * blockDim.x) {
    const IJK c = coords(op, id);
    const int i = c.i, j = c.j;
    const double k = dudt_at<KIND, S>(op, v, id, i, j, f, f_const, nl, ut, inv_ut);
    const double u0 = ldv(un, id);
    double o;
    switch (mode) {
      case 0: if (acc) acc[id] = static_cast<S>(k); o = u0 + c1 * k; break;
      case 1: { const double a = ldv(acc, id) + 2.0 * k; acc[id] = static_cast<S>(a); o = u0 + c1 * k; break; }
      case 2: o = u0 + c1 * ((ldv(acc, id)) + k); break;
      case 3: o = u0 + c1 * (ldv(acc, id) + k); break;
      case 4: o = 0.5 * u0 + 0.5 * (ldv(v, id) + dt * k); break;
      case 5: o = 0.75 * u0 + 0.25 * (ldv(v, id) + dt * k); break;
      default: o = (1.0 / 3.0) * u0 + (2.0 / 3.0) * (ldv(v, id) + dt * k); break;
    }
    out[id] = static_cast<S>(o);
  }
}

// u_tt inline update (solver.cpp:1307-1311): vel += dt*accel, next = u + dt*vel.
template <int KIND>
__global__ void k_utt(DevOp op, const double* __restrict__ u, double* __restrict__ vel,
                      double* __restrict__ next, const double* f, double f_const, NonlinearTerms nl,
                      double ut, double utt, double dt) {
  for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < op.n;
       id += (long long)gridDim.x * blockDim.x) {
    const IJK c = coords(op, id);
    const int i = c.i, j = c.j;
    if (is_boundary2(op, i, j)) { next[id] = u[id]; continue; }
    const W7<double> w = row_weights<2, KIND, double, double>(op, id);
    const double opv = w.C * u[id] + (w.E * u[id + 1] + w.W * u[id - 1] + w.N * u[id + op.nx] +
                                      w.S * u[id - op.nx]);
    const double rhs = opv + (f ? f[id] : f_const) + (nl.n ? eval_nl(nl, u[id]) : 0.0);
    const double accel = safe_div(-(rhs + ut * vel[id]), utt);
    const double vn = vel[id] + dt * accel;
    vel[id] = vn;
    next[id] = u[id] + dt * vn;
  }
}

// ComputeResidualNorms (residual.cpp:153-583): per-node residual, block partial sum of r^2
// and max |r|.  2D includes boundary rows with the residual's closures (Neumann along the
// axis, Robin alpha+beta/h on every face); 3D interior rows only.
template <int DIM, int KIND>
__global__ void __launch_bounds__(kBlock) k_resnorm(DevOp op, FaceData fd, const double* __restrict__ u,
                                                   const double* f, double f_const,
                                                   NonlinearTerms nl, double* sums, double* maxs,
                                                   int* degenerate) {
  __shared__ double lds[kBlock / kWave];
  double acc = 0.0, mx = 0.0;
  for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < op.n;
       id += (long long)gridDim.x * blockDim.x) {
    const IJK c = coords(op, id);
    const int i = c.i, j = c.j, k = c.k;
    double r;
    if (DIM == 2 && is_boundary2(op, i, j)) {
      if (!resnorm_boundary2(op, fd, u, id, i, j, &r)) continue;
    } else {
      if (is_boundary2(op, i, j) || (DIM == 3 && (k == 0 || k == op.nz - 1))) continue;
      const W7<double> w = row_weights<DIM, KIND, double, double>(op, id);
      if (fabs(w.C) < 1e-12) { *degenerate = 1; continue; }
      double s = w.E * u[id + 1] + w.W * u[id - 1] + w.N * u[id + op.nx] + w.S * u[id - op.nx];
      if (DIM == 3) s = s + w.F * u[id + op.sxy] + w.B * u[id - op.sxy];
      const double nlv = (DIM == 2 && nl.n) ? eval_nl(nl, u[id]) : 0.0;
      r = w.C * u[id] + s + ((f ? f[id] : f_const) + nlv);
    }
    acc += r * r;
    mx = nmax(mx, fabs(r));
  }
  const double s = block_sum(acc, lds);
  const double m = block_max(mx, lds);
  if (threadIdx.x == 0) { sums[blockIdx.x] = s; maxs[blockIdx.x] = m; }
}

// 3D form of k_resnorm (interior rows only): blocks take whole x rows (k, j) in turn, so node
// coordinates come from one division per row instead of two 64-bit divisions per node.  (Four
// rows in flight per thread measured slower: the FP64 face divisions of four nodes cost the
// occupancy.)
template <int KIND>
__global__ void __launch_bounds__(kBlock) k_resnorm3(DevOp op, const double* __restrict__ u,
                                                    const double* f, double f_const, double* sums,
                                                    double* maxs, int* degenerate) {
  __shared__ double lds[kBlock / kWave];
  double acc = 0.0, mx = 0.0;
  const int rows_j = op.ny - 2;
  const long long rows = (long long)rows_j * (op.nz - 2);
  for (long long row = blockIdx.x; row < rows; row += gridDim.x) {
    const int j = 1 + (int)(row % rows_j);
    const int k = 1 + (int)(row / rows_j);
    const long long base = (long long)k * op.sxy + (long long)j * op.nx;
    for (int i = 1 + threadIdx.x; i <= op.nx - 2; i += blockDim.x) {
      const long long id = base + i;
      const W7<double> w = row_weights<3, KIND, double, double>(op, id);
      if (fabs(w.C) < 1e-12) { *degenerate = 1; continue; }
      double s = w.E * u[id + 1] + w.W * u[id - 1] + w.N * u[id + op.nx] + w.S * u[id - op.nx];
      s = s + w.F * u[id + op.sxy] + w.B * u[id - op.sxy];
      const double r = w.C * u[id] + s + (f ? f[id] : f_const);
      acc += r * r;
      mx = nmax(mx, fabs(r));
    }
  }
  const double s = block_sum(acc, lds);
  const double m = block_max(mx, lds);
  if (threadIdx.x == 0) { sums[blockIdx.x] = s; maxs[blockIdx.x] = m; }
}

// Lexicographic Gauss-Seidel of the reference V-cycle (multigrid.cpp:69-94) for one 2D level:
// one workgroup walks the anti-diagonals d = i + j in order, so every node sees its left and
// lower neighbours already updated and its right and upper neighbours not yet, exactly the
// lexicographic order.  ax, by, center: constant Poisson+e stencil.
__global__ void __launch_bounds__(1024) k_lexgs(int nx, int ny, double ax, double by, double center,
                                               const double* __restrict__ b, double* __restrict__ u,
                                               int iters) {
  for (int it = 0; it < iters; ++it) {
    for (int d = 2; d <= (nx - 2) + (ny - 2); ++d) {
      const int i_lo = max(1, d - (ny - 2));
      const int i_hi = min(nx - 2, d - 1);
      for (int i = i_lo + threadIdx.x; i <= i_hi; i += blockDim.x) {
        const int j = d - i;
        const long long id = (long long)j * nx + i;
        const double rhs = b[id] - ax * (u[id - 1] + u[id + 1]) - by * (u[id - nx] + u[id + nx]);
        u[id] = rhs / center;
      }
      __syncthreads();
    }
  }
}

// b = -f on interior rows, the boundary datum elsewhere (BuildRhs2D, solver.cpp:214-230).
__global__ void k_build_rhs(DevOp op, const double* f, double f_const, const double* bnd,
                            double* b) {
  const bool d3 = op.nz > 1;
  for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < op.n;
       id += (long long)gridDim.x * blockDim.x) {
    const IJK c = coords(op, id);
    const int i = c.i, j = c.j, k = c.k;
    const bool bd = is_boundary2(op, i, j) || (d3 && (k == 0 || k == op.nz - 1));
    b[id] = bd ? bnd[id] : -(f ? f[id] : f_const);
  }
}

// x64 += s * e32, s = *scale (host value when scale == null)
__global__ void k_add_f32(double* __restrict__ x, const float* __restrict__ e, long long n,
                          const double* scale, double hs) {
  const double s = scale ? *scale : hs;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    x[i] += s * static_cast<double>(e[i]);
}

// out32 = (float)(in64 * s), s = 1 / *norm
__global__ void k_to_f32_scaled(const double* __restrict__ in, float* __restrict__ out, long long n,
                                const double* norm, const KScalars* sc) {
  if (sc && (sc->done | sc->breakdown)) return;
  const double s = (*norm > 0.0) ? 1.0 / *norm : 1.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    out[i] = static_cast<float>(in[i] * s);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_maxabs(const T* __restrict__ a, long long n,
                                                  double* partials) {
  __shared__ double lds[kBlock / kWave];
  double m = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    m = nmax(m, fabs(static_cast<double>(a[i])));
  const double r = block_max(m, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

// scale = 1/max (inverse kept too); used to normalise the FP16 levels' input.
__global__ void __launch_bounds__(1024) k_fin_scale(const double* partials, int np, KScalars* sc) {
  __shared__ double lds[1024 / kWave];
  if (sc->done | sc->breakdown) return;
  double m = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) m = nmax(m, partials[i]);
  const double r = block_max(m, lds);
  if (threadIdx.x == 0) {
    sc->inv_scale = r > 0.0 ? r : 1.0;
    sc->scale = r > 0.0 ? 1.0 / r : 1.0;
  }
}

// x64 += s x32, x32 = 0, when a replacement is due (gate clear).
__global__ void k_flush_f32(double* __restrict__ x, float* __restrict__ e, long long n,
                            const double* scale, const KScalars* gate) {
  if (gate->done | gate->breakdown) return;
  const double s = *scale;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    x[i] += s * static_cast<double>(e[i]);
    e[i] = 0.0f;
  }
}

// Residual-replacement bookkeeping of the mixed CG (one block): the inner recurrence runs on
// r / snorm with snorm = |r0| fixed; a replacement is due once the recurrence residual has
// fallen by `delta` below the largest one since the last replacement, or claims the outer
// target.  Only a true (FP64) residual stops the solve.
__global__ void __launch_bounds__(1024) k_fin_ru(KScalars* si, KScalars* so, KScalars* gate,
                                                const double* partials, int np, int op,
                                                double delta) {
  __shared__ double lds[1024 / kWave];
  double acc = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) acc += partials[i];
  const double s = block_sum(acc, lds);
  if (threadIdx.x != 0) return;
  switch (op) {
    case FIN_RU_START:
      so->snorm = so->metric > 0.0 ? so->metric : 1.0;
      so->steps = 0;
      si->done = so->done | so->breakdown;
      gate->done = 1;
      break;
    case FIN_RU_CHECK:
      if (si->done | si->breakdown) { gate->done = 1; break; }
      si->rr = s;
      si->metric = sqrt(s);
      si->iter += 1;
      if (nonfinite(s)) { si->breakdown = 1; so->breakdown = 1; gate->done = 1; break; }
      si->rho = fmax(si->rho, si->metric);
      si->flag_s = (si->metric < delta * si->rho || si->metric * so->snorm < so->target) ? 1 : 0;
      gate->done = si->flag_s ? 0 : 1;
      break;
    case FIN_RU_APPLY:
      gate->done = 1;
      if ((si->done | si->breakdown) || !si->flag_s) break;
      so->metric = sqrt(s);
      so->steps += 1;
      so->iter = si->iter;
      si->flag_s = 0;
      if (nonfinite(s)) { so->breakdown = 1; si->breakdown = 1; break; }
      si->metric = so->metric / so->snorm;
      si->rho = si->metric;
      if (so->metric < so->target) { so->done = 1; si->done = 1; }
      break;
    default:
      break;
  }
}

}  // namespace

static long long misc_blocks(long long n) {
  long long b = (n + kBlock - 1) / kBlock;
  return b > 4096 ? 4096 : (b < 1 ? 1 : b);
}

void build_rhs(hipStream_t s, const DevOp& op, const double* f, double f_const, const double* bnd,
               double* b) {
  hipLaunchKernelGGL(k_build_rhs, dim3(misc_blocks(op.n)), dim3(kBlock), 0, s, op, f, f_const, bnd, b);
}

void add_f32(hipStream_t s, double* x, const float* e, long long n, const double* scale, double hs) {
  hipLaunchKernelGGL(k_add_f32, dim3(misc_blocks(n)), dim3(kBlock), 0, s, x, e, n, scale, hs);
}

void to_f32_scaled(hipStream_t s, const double* in, float* out, long long n, const double* norm) {
  hipLaunchKernelGGL(k_to_f32_scaled, dim3(misc_blocks(n)), dim3(kBlock), 0, s, in, out, n, norm,
                     static_cast<const KScalars*>(nullptr));
}

void to_f32_scaled_gated(hipStream_t s, const double* in, float* out, long long n,
                         const double* norm, const KScalars* gate) {
  hipLaunchKernelGGL(k_to_f32_scaled, dim3(misc_blocks(n)), dim3(kBlock), 0, s, in, out, n, norm,
                     gate);
}

void flush_f32(hipStream_t s, double* x64, float* x32, long long n, const double* scale,
               const KScalars* gate) {
  hipLaunchKernelGGL(k_flush_f32, dim3(misc_blocks(n)), dim3(kBlock), 0, s, x64, x32, n, scale,
                     gate);
}

void finalize_ru(hipStream_t s, KScalars* si, KScalars* so, KScalars* gate,
                 const double* partials, int np, int op, double delta) {
  hipLaunchKernelGGL(k_fin_ru, dim3(1), dim3(1024), 0, s, si, so, gate, partials, np, op, delta);
}

int maxabs_blocks(long long n) { return static_cast<int>(misc_blocks(n) > 1024 ? 1024 : misc_blocks(n)); }

void maxabs32(hipStream_t s, const float* a, long long n, double* partials) {
  hipLaunchKernelGGL((k_maxabs<float>), dim3(maxabs_blocks(n)), dim3(kBlock), 0, s, a, n, partials);
}

void fin_scale(hipStream_t s, const double* partials, int np, KScalars* sc) {
  hipLaunchKernelGGL(k_fin_scale, dim3(1), dim3(1024), 0, s, partials, np, sc);
}

void time_stage(hipStream_t s, const DevOp& op, const double* v, const double* un, double* acc,
                double* out, const double* f, double f_const, const NonlinearTerms& nl, double ut,
                double dt, double c1, int mode) {
  long long blocks = (op.n + kBlock - 1) / kBlock;
  if (blocks > 8192) blocks = 8192;
  if (op.kind == PDEHIP_OP_POINTWISE)
    hipLaunchKernelGGL((k_stage<PDEHIP_OP_POINTWISE, double>), dim3(blocks), dim3(kBlock), 0, s, op, v, un, acc, out, f, f_const, nl, ut, dt, c1, mode);
  else if (op.kind == PDEHIP_OP_FLUX)
    hipLaunchKernelGGL((k_stage<PDEHIP_OP_FLUX, double>), dim3(blocks), dim3(kBlock), 0, s, op, v, un, acc, out, f, f_const, nl, ut, dt, c1, mode);
  else
    hipLaunchKernelGGL((k_stage<PDEHIP_OP_CONST, double>), dim3(blocks), dim3(kBlock), 0, s, op, v, un, acc, out, f, f_const, nl, ut, dt, c1, mode);
}

void time_stage32(hipStream_t s, const DevOp& op, const float* v, const float* un, float* acc,
                  float* out, const double* f, double f_const, const NonlinearTerms& nl, double ut,
                  double dt, double c1, int mode) {
  long long blocks = (op.n + kBlock - 1) / kBlock;
  if (blocks > 8192) blocks = 8192;
  if (op.kind == PDEHIP_OP_POINTWISE)
    hipLaunchKernelGGL((k_stage<PDEHIP_OP_POINTWISE, float>), dim3(blocks), dim3(kBlock), 0, s, op, v, un, acc, out, f, f_const, nl, ut, dt, c1, mode);
  else if (op.kind == PDEHIP_OP_FLUX)
    hipLaunchKernelGGL((k_stage<PDEHIP_OP_FLUX, float>), dim3(blocks), dim3(kBlock), 0, s, op, v, un, acc, out, f, f_const, nl, ut, dt, c1, mode);
  else
    hipLaunchKernelGGL((k_stage<PDEHIP_OP_CONST, float>), dim3(blocks), dim3(kBlock), 0, s, op, v, un, acc, out, f, f_const, nl, ut, dt, c1, mode);
}

void time_utt(hipStream_t s, const DevOp& op, const double* u, double* vel, double* next,
              const double* f, double f_const, const NonlinearTerms& nl, double ut, double utt,
              double dt) {
  long long blocks = (op.n + kBlock - 1) / kBlock;
  if (blocks > 8192) blocks = 8192;
  if (op.kind == PDEHIP_OP_POINTWISE)
    hipLaunchKernelGGL((k_utt<PDEHIP_OP_POINTWISE>), dim3(blocks), dim3(kBlock), 0, s, op, u, vel, next, f, f_const, nl, ut, utt, dt);
  else if (op.kind == PDEHIP_OP_FLUX)
    hipLaunchKernelGGL((k_utt<PDEHIP_OP_FLUX>), dim3(blocks), dim3(kBlock), 0, s, op, u, vel, next, f, f_const, nl, ut, utt, dt);
  else
    hipLaunchKernelGGL((k_utt<PDEHIP_OP_CONST>), dim3(blocks), dim3(kBlock), 0, s, op, u, vel, next, f, f_const, nl, ut, utt, dt);
}

int resnorm_blocks(const DevOp& op) {
  long long blocks = (op.n + kBlock - 1) / kBlock;
  return static_cast<int>(blocks > 2048 ? 2048 : blocks);
}

void residual_norms(hipStream_t s, const DevOp& op, const FaceData& fd, const double* u,
                    const double* f, double f_const, const NonlinearTerms& nl, double* sums,
                    double* maxs, int* degenerate) {
  const dim3 g(resnorm_blocks(op));
  const bool d3 = op.nz > 1;
#define RN_(D, K) hipLaunchKernelGGL((k_resnorm<D, K>), g, dim3(kBlock), 0, s, op, fd, u, f, f_const, nl, sums, maxs, degenerate)
#define RN3_(K) hipLaunchKernelGGL((k_resnorm3<K>), g, dim3(kBlock), 0, s, op, u, f, f_const, sums, maxs, degenerate)
  if (d3 && op.nx >= 3 && op.ny >= 3 && op.nz >= 3) {
    if (op.kind == PDEHIP_OP_POINTWISE) RN3_(PDEHIP_OP_POINTWISE);
    else if (op.kind == PDEHIP_OP_FLUX) RN3_(PDEHIP_OP_FLUX);
    else RN3_(PDEHIP_OP_CONST);
  } else if (op.kind == PDEHIP_OP_POINTWISE) { if (d3) RN_(3, PDEHIP_OP_POINTWISE); else RN_(2, PDEHIP_OP_POINTWISE); }
  else if (op.kind == PDEHIP_OP_FLUX) { if (d3) RN_(3, PDEHIP_OP_FLUX); else RN_(2, PDEHIP_OP_FLUX); }
  else { if (d3) RN_(3, PDEHIP_OP_CONST); else RN_(2, PDEHIP_OP_CONST); }
#undef RN3_
#undef RN_
}

void lex_gs(hipStream_t s, int nx, int ny, double ax, double by, double center, const double* b,
            double* u, int iters) {
  const int diag = (nx - 2 < ny - 2) ? nx - 2 : ny - 2;
  int threads = 64;
  while (threads < diag && threads < 1024) threads *= 2;
  hipLaunchKernelGGL(k_lexgs, dim3(1), dim3(threads), 0, s, nx, ny, ax, by, center, b, u, iters);
}

}  // namespace pdehip
