// k_time.hip — explicit time-step stage kernels, the reference residual-norm kernel and the
// single-workgroup lexicographic Gauss-Seidel used by the reference-semantics V-cycle.
#include <algorithm>
#include <type_traits>

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

// Which nonlinear terms a problem has, fixed per launch: none, integer powers 0..3 only
// (evaluated as one cubic), or any kind.
enum NlMode { NL_NONE = 0, NL_CUBIC = 1, NL_ANY = 2 };

int nl_mode(const NonlinearTerms& nl) {
  if (nl.n <= 0) return NL_NONE;
  for (int t = 0; t < nl.n; ++t)
    if (nl.t[t].kind != PDEHIP_NL_POWER || nl.t[t].power < 0 || nl.t[t].power > 3) return NL_ANY;
  return NL_CUBIC;
}

// An NL_CUBIC problem's terms as the cubic a0 + a1 u + a2 u^2 + a3 u^3 (coefficients of equal
// powers summed), evaluated by Horner's rule: three FMAs per node instead of a per-term
// selection.  It rounds differently from the term-by-term sum (by an ulp or so of the terms),
// the same way in every kernel that takes this path.
struct Cubic {
  double a[4];
};

Cubic cubic_of(const NonlinearTerms& nl) {
  Cubic c{};
  for (int t = 0; t < nl.n; ++t) c.a[nl.t[t].power] += nl.t[t].coeff;
  return c;
}

__device__ __forceinline__ double eval_cubic(const Cubic& c, double u) {
  return fma(fma(fma(c.a[3], u, c.a[2]), u, c.a[1]), u, c.a[0]);
}

// Any terms, in order, summed from 0.
__device__ __forceinline__ double eval_nl(const NonlinearTerms& nl, double u) {
  double v = 0.0;
#pragma unroll
  for (int t = 0; t < PDEHIP_MAX_NONLINEAR; ++t) {
    if (t >= nl.n) break;
    const pdehip_nonlinear& term = nl.t[t];
    if (term.kind == PDEHIP_NL_POWER) v += term.coeff * ipow(u, term.power);
    else v += eval_nl_other(term.kind, term.coeff, u);
  }
  return v;
}

// N(u) under the launch's mode.
template <int NLM>
__device__ __forceinline__ double eval_terms(const NonlinearTerms& nl, const Cubic& cb, double u) {
  if constexpr (NLM == NL_NONE) return 0.0;
  else if constexpr (NLM == NL_CUBIC) return eval_cubic(cb, u);
  else return eval_nl(nl, u);
}

// Launch f(KIND, NLM) with both as compile-time constants.
template <typename F>
void by_kind_nl(const DevOp& op, const NonlinearTerms& nl, F&& f) {
  using PW = std::integral_constant<int, PDEHIP_OP_POINTWISE>;
  using FX = std::integral_constant<int, PDEHIP_OP_FLUX>;
  using CT = std::integral_constant<int, PDEHIP_OP_CONST>;
  const int m = nl_mode(nl);
  auto with_kind = [&](auto K) {
    if (m == NL_NONE) f(K, std::integral_constant<int, NL_NONE>{});
    else if (m == NL_CUBIC) f(K, std::integral_constant<int, NL_CUBIC>{});
    else f(K, std::integral_constant<int, NL_ANY>{});
  };
  if (op.kind == PDEHIP_OP_POINTWISE) with_kind(PW{});
  else if (op.kind == PDEHIP_OP_FLUX) with_kind(FX{});
  else with_kind(CT{});
}

__device__ __forceinline__ double safe_div(double num, double den) {  // safe_math.h:48-58
  if (fabs(den) < 1e-12) return 0.0;
  const double r = num / den;
  return isfinite(r) ? r : 0.0;
}

template <typename T> __device__ __forceinline__ double ldv(const T* p, long long i) { return static_cast<double>(p[i]); }

// du/dt = -(A v + f + N(v))/ut at interior node id from its value and its four neighbours'
// (fill_dudt, reference solver.cpp:1182-1195).  One definition for the per-stage kernel and the
// fused step, so both round alike.
template <int KIND, int NLM>
__device__ __forceinline__ double dudt_values(const DevOp& op, long long id, double vc, double ve,
                                             double vw, double vn, double vs, const double* f,
                                             double f_const, const NonlinearTerms& nl,
                                             const Cubic& cb, double ut, double inv_ut) {
  const W7<double> w = row_weights<2, KIND, double, double>(op, id);
  const double opv = w.C * vc + (w.E * ve + w.W * vw + w.N * vn + w.S * vs);
  const double rhs = opv + (f ? f[id] : f_const) + eval_terms<NLM>(nl, cb, vc);
  // safe_div(-rhs, ut) with the reciprocal of the (uniform) u_t coefficient formed once
  if (fabs(ut) < 1e-12) return 0.0;
  const double r = -rhs * inv_ut;
  return isfinite(r) ? r : 0.0;
}

// The same, zero on the boundary, with the values read from v.
template <int KIND, int NLM, typename S>
__device__ __forceinline__ double dudt_at(const DevOp& op, const S* v, long long id, int i, int j,
                                         const double* f, double f_const, const NonlinearTerms& nl,
                                         const Cubic& cb, double ut, double inv_ut) {
  if (is_boundary2(op, i, j)) return 0.0;
  return dudt_values<KIND, NLM>(op, id, ldv(v, id), ldv(v, id + 1), ldv(v, id - 1), ldv(v, id + op.nx),
                           ldv(v, id - op.nx), f, f_const, nl, cb, ut, inv_ut);
}

// RK4 stage combinations (k: the stage's du/dt, u0: u_n, a: the accumulator) with explicit
// fused multiply-adds, shared by the per-stage kernel and the fused step.
__device__ __forceinline__ double stage_out(double u0, double c1, double k) { return fma(c1, k, u0); }
__device__ __forceinline__ double stage_acc(double a, double k) { return fma(2.0, k, a); }
__device__ __forceinline__ double stage_last(double u0, double c1, double a, double k) { return fma(c1, a + k, u0); }

// Stage kernel.  Per node:  k = F(v)
//   mode 0 (FE / RK4 s1 / RK2 s1 / SSP s1): acc = k (if acc), out = un + c1*k
//   mode 1 (RK4 s2, s3):                      acc += 2k,        out = un + c1*k
//   mode 2 (RK4 s4):                          out = un + c1*(acc + k)
//   mode 3 (RK2 s2):                          out = un + c1*(acc + k)     (c1 = dt/2)
//   mode 4 (SSPRK2 s2):                       out = 0.5*un + 0.5*(v + dt*k)
//   mode 5 (SSPRK3 s2):                       out = 0.75*un + 0.25*(v + dt*k)
//   mode 6 (SSPRK3 s3):                       out = (1/3)*un + (2/3)*(v + dt*k)
template <int KIND, int NLM, typename S>
__global__ void __launch_bounds__(kBlock) k_stage(DevOp op, const S* __restrict__ v,
                                                 const S* __restrict__ un, S* __restrict__ acc,
                                                 S* __restrict__ out, const double* f,
                                                 double f_const, NonlinearTerms nl, Cubic cb,
                                                 double ut, double dt, double c1, int mode) {
  const double inv_ut = 1.0 / ut;
  for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < op.n;
       id += (long long)gridDim.x * blockDim.x) {
    const IJK c = coords(op, id);
    const int i = c.i, j = c.j;
    const double k = dudt_at<KIND, NLM, S>(op, v, id, i, j, f, f_const, nl, cb, ut, inv_ut);
    const double u0 = ldv(un, id);
    double o;
    switch (mode) {
      case 0: if (acc) acc[id] = static_cast<S>(k); o = stage_out(u0, c1, k); break;
      case 1: { const double a = stage_acc(ldv(acc, id), k); acc[id] = static_cast<S>(a); o = stage_out(u0, c1, k); break; }
      case 2: o = stage_last(u0, c1, ldv(acc, id), k); break;
      case 3: o = stage_last(u0, c1, ldv(acc, id), k); break;
      case 4: o = 0.5 * u0 + 0.5 * (ldv(v, id) + dt * k); break;
      case 5: o = 0.75 * u0 + 0.25 * (ldv(v, id) + dt * k); break;
      default: o = (1.0 / 3.0) * u0 + (2.0 / 3.0) * (ldv(v, id) + dt * k); break;
    }
    out[id] = static_cast<S>(o);
  }
}

// One whole RK4 step of a 2D time series per launch (temporal blocking).  A workgroup owns a
// TX x TY tile (TX = 54) and holds u on the tile plus a 5-node ring, EX = 64 columns by EY
// rows, in LDS.  Wave w (of 4) owns rows w, w + 4, ... and lane l column l: the same nodes in
// every stage, so u_n and the RK4 accumulator stay in registers and LDS holds only the stage
// input and output.  Stage r runs on the extended tile shrunk by r nodes per side, so no stage
// needs another workgroup's values: per node, u_n is read once and u_{n+1} written once
// instead of the 17 accesses of the per-stage launches.  Stage values are rounded to the state
// type S where the per-stage kernels store them (k1 and the accumulator, each stage input, the
// result), with the same shared arithmetic, so the result is theirs bit for bit.  The closures
// that follow a step (Neumann / Robin x faces, then y faces, then Dirichlet x and y faces;
// reference solver.cpp:1300-1329) run in LDS.  u_{n+1} is formed on the tile plus one ring
// because a face node's inward neighbour can sit in the next tile (a grid side of 1 + a
// multiple of the tile width); the x-face closure runs on that ring as well, since the y-face
// closure reads its results.
template <int KIND, int NLM, typename S, int TY>
__global__ void __launch_bounds__(256) k_rk4_step(DevOp op, FaceData fd, const S* __restrict__ u,
                                                 S* __restrict__ out, const double* f, double f_const,
                                                 NonlinearTerms nl, Cubic cb, double ut, double h2,
                                                 double h1, double h6, int tiles_x) {
  constexpr int H = 5, EX = 64, TX = EX - 2 * H, EY = TY + 2 * H, KR = EY / 4;
  static_assert(EY % 4 == 0, "the rows are dealt to four waves");
  __shared__ S L0[EX * EY], L1[EX * EY];
  const int nx = op.nx, ny = op.ny;
  const int lane = static_cast<int>(threadIdx.x & 63);
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const int i0 = static_cast<int>(blockIdx.x % tiles_x) * TX - H;
  const int j0 = static_cast<int>(blockIdx.x / tiles_x) * TY - H;
  const int i = i0 + lane;
  const bool col_in = i >= 0 && i < nx;
  const bool col_face = i == 0 || i == nx - 1;
  S u0[KR], acc[KR];
#pragma unroll
  for (int k = 0; k < KR; ++k) {
    const int ey = wave + 4 * k, j = j0 + ey;
    const S v = (col_in && j >= 0 && j < ny) ? u[(long long)j * nx + i] : S(0);
    u0[k] = v;
    acc[k] = S(0);
    L0[ey * EX + lane] = v;
  }
  __syncthreads();
  const double inv_ut = 1.0 / ut;
  // stage r reads L0, L1, L0, L1 (u_n, T1, T2, T3) and writes L1, L0, L1, L0 (T1, T2, T3,
  // u_{n+1}); u_n itself comes from registers after stage 1
#pragma unroll 1
  for (int r = 1; r <= 4; ++r) {
    const S* vin = (r & 1) ? L0 : L1;
    S* vout = (r & 1) ? L1 : L0;
    const double c1 = r == 3 ? h1 : h2;
    const bool lane_on = col_in && lane >= r && lane < EX - r;
#pragma unroll
    for (int k = 0; k < KR; ++k) {
      const int ey = wave + 4 * k, j = j0 + ey;
      if (ey < r || ey >= EY - r || j < 0 || j >= ny) continue;   // the same for the whole wave
      if (!lane_on) continue;
      const int q = ey * EX + lane;
      double kv = 0.0;
      if (!col_face && j != 0 && j != ny - 1)
        kv = dudt_values<KIND, NLM>(op, (long long)j * nx + i, double(vin[q]), double(vin[q + 1]),
                                    double(vin[q - 1]), double(vin[q + EX]), double(vin[q - EX]),
                                    f, f_const, nl, cb, ut, inv_ut);
      const double un = double(u0[k]);
      if (r == 1) {
        acc[k] = static_cast<S>(kv);
        vout[q] = static_cast<S>(stage_out(un, c1, kv));
      } else if (r < 4) {
        acc[k] = static_cast<S>(stage_acc(double(acc[k]), kv));
        vout[q] = static_cast<S>(stage_out(un, c1, kv));
      } else {
        vout[q] = static_cast<S>(stage_last(un, h6, double(acc[k]), kv));
      }
    }
    __syncthreads();
  }
  // closures on L0 = u_{n+1}: phase 0 (x faces, Neumann / Robin) on the tile plus one ring,
  // the later phases on the tile; only workgroups whose tile or ring holds a face node
  const bool on_face = i0 + H - 1 <= 0 || i0 + EX - H >= nx - 1 || j0 + H - 1 <= 0 ||
                       j0 + EY - H >= ny - 1;
#pragma unroll 1
  for (int phase = 0; on_face && phase < 4; ++phase) {
    const int axis = phase & 1;
    const bool dirichlet = phase >= 2;
    const int lo = phase == 0 ? H - 1 : H;     // region [lo, EX - lo) x [lo, EY - lo)
    const bool lane_on = col_in && lane >= lo && lane < EX - lo;
#pragma unroll
    for (int k = 0; k < KR; ++k) {
      const int ey = wave + 4 * k, j = j0 + ey;
      if (ey < lo || ey >= EY - lo || j < 0 || j >= ny) continue;
      if (!lane_on) continue;
      const int pos = axis == 0 ? i : j, last = axis == 0 ? nx - 1 : ny - 1;
      if (pos != 0 && pos != last) continue;
      const int side = pos == 0 ? 0 : 1, face = 2 * axis + side;
      const long long fq = axis == 0 ? j : i;
      const int q = ey * EX + lane;
      if (dirichlet) {
        if (fd.kind[face] == PDEHIP_BC_DIRICHLET) L0[q] = static_cast<S>(fd.value[face][fq]);
      } else if (fd.kind[face] != PDEHIP_BC_DIRICHLET) {
        const int inner = axis == 0 ? q + (side ? -1 : 1) : q + (side ? -EX : EX);
        double v;
        if (closure_nr_value(fd, face, side, fq, double(L0[inner]), &v)) L0[q] = static_cast<S>(v);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < KR; ++k) {
    const int ey = wave + 4 * k, j = j0 + ey;
    if (ey < H || ey >= EY - H || j >= ny) continue;
    if (lane >= H && lane < EX - H && i < nx) out[(long long)j * nx + i] = L0[ey * EX + lane];
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

// out32 = (float)(in64 / *norm), zero_out = 0, partials of out.out and max |out|
__global__ void __launch_bounds__(kBlock) k_to_f32_scaled_stats(const double* __restrict__ in,
                                                               float* __restrict__ out, long long n,
                                                               const double* norm,
                                                               float* __restrict__ zero_out,
                                                               double* sums, double* maxs) {
  __shared__ double lds[kBlock / kWave];
  const double s = (*norm > 0.0) ? 1.0 / *norm : 1.0;
  double acc = 0.0, m = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float v = static_cast<float>(in[i] * s);
    out[i] = v;
    if (zero_out) zero_out[i] = 0.0f;
    acc += static_cast<double>(v) * static_cast<double>(v);
    m = nmax(m, fabs(static_cast<double>(v)));
  }
  const double t = block_sum(acc, lds);
  if (threadIdx.x == 0) sums[blockIdx.x] = t;
  __syncthreads();   // lds reused
  const double r = block_max(m, lds);
  if (threadIdx.x == 0) maxs[blockIdx.x] = r;
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
                            const double* scale, const KScalars* gate, int skip_fresh) {
  if (gate->done | gate->breakdown) return;
  if (skip_fresh && gate->fresh) return;
  const double s = *scale;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    x[i] = fma(s, static_cast<double>(e[i]), x[i]);   // as the fused replacement forms it
    e[i] = 0.0f;
  }
}

// dst = src when the solve's current solution is the second buffer (so->xsel)
__global__ void k_copy_if_xsel(double* __restrict__ dst, const double* __restrict__ src,
                               long long n, const KScalars* so) {
  if (!so->xsel) return;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// The final true residual is skipped when the solve converged at a replacement, which left
// r = b - A x of the final x and its partials.
__global__ void k_fresh_gate(const KScalars* so, const KScalars* si, KScalars* gate) {
  gate->done = so->done && !so->breakdown && !si->breakdown;
  gate->breakdown = 0;
}

// Residual-replacement bookkeeping of the mixed CG (one block): the inner recurrence runs on
// r / snorm with snorm = |r0| fixed; a replacement is due once the recurrence residual has
// fallen by `delta` below the largest one since the last replacement, or claims the outer
// target.  Only a true (FP64) residual stops the solve.
__device__ void fin_ru_step(KScalars* si, KScalars* so, KScalars* gate, double s, int op,
                            double delta);

// With maxp: afterwards, as k_fin_scale on si, the input normalisation from the max partials.
__global__ void __launch_bounds__(1024) k_fin_ru(KScalars* si, KScalars* so, KScalars* gate,
                                                const double* partials, int np, int op,
                                                double delta, const double* maxp, int nm) {
  __shared__ double lds[1024 / kWave];
  double acc = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) acc += partials[i];
  const double s = block_sum(acc, lds);
  double mx = 0.0;
  if (maxp) {
    for (int i = threadIdx.x; i < nm; i += blockDim.x) mx = nmax(mx, maxp[i]);
    mx = block_max(mx, lds);
  }
  if (threadIdx.x != 0) return;
  fin_ru_step(si, so, gate, s, op, delta);
  if (maxp && !(si->done | si->breakdown)) {
    si->inv_scale = mx > 0.0 ? mx : 1.0;
    si->scale = mx > 0.0 ? 1.0 / mx : 1.0;
  }
}

__device__ void fin_ru_step(KScalars* si, KScalars* so, KScalars* gate, double s, int op,
                            double delta) {
  switch (op) {
    case FIN_RU_START:
      so->snorm = so->metric > 0.0 ? so->metric : 1.0;
      so->steps = 0;
      si->done = so->done | so->breakdown;
      gate->done = 1;
      break;
    case FIN_RU_CHECK:
      if (si->done | si->breakdown) { gate->done = 1; break; }
      so->fresh = 0;   // the CG update before this check has restarted the FP32 correction
      si->rr = s;
      si->metric = sqrt(s);
      si->iter += 1;
      if (nonfinite(s)) { si->breakdown = 1; so->breakdown = 1; gate->done = 1; break; }
      si->rho = fmax(si->rho, si->metric);
      si->flag_s = (si->metric < delta * si->rho || si->metric * so->snorm < so->target) ? 1 : 0;
      gate->done = si->flag_s ? 0 : 1;
      break;
    case FIN_RU_APPLY:
    case FIN_RU_APPLY_PP:
      gate->done = 1;
      if ((si->done | si->breakdown) || !si->flag_s) break;
      if (op == FIN_RU_APPLY_PP) {
        // the fused replacement ran: its output buffer is now the solution, and the FP32
        // correction it folded in is void until the next CG update
        so->xsel ^= 1;
        so->fresh = 1;
      }
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

void to_f32_scaled_stats(hipStream_t s, const double* in, float* out, long long n,
                         const double* norm, float* zero_out, double* sums, double* maxs) {
  hipLaunchKernelGGL(k_to_f32_scaled_stats, dim3(vec_blocks(n)), dim3(kBlock), 0, s, in, out, n,
                     norm, zero_out, sums, maxs);
}

void to_f32_scaled_gated(hipStream_t s, const double* in, float* out, long long n,
                         const double* norm, const KScalars* gate) {
  hipLaunchKernelGGL(k_to_f32_scaled, dim3(misc_blocks(n)), dim3(kBlock), 0, s, in, out, n, norm,
                     gate);
}

void flush_f32(hipStream_t s, double* x64, float* x32, long long n, const double* scale,
               const KScalars* gate, bool skip_fresh) {
  hipLaunchKernelGGL(k_flush_f32, dim3(misc_blocks(n)), dim3(kBlock), 0, s, x64, x32, n, scale,
                     gate, skip_fresh ? 1 : 0);
}

void copy_if_xsel(hipStream_t s, double* dst, const double* src, long long n, const KScalars* so) {
  hipLaunchKernelGGL(k_copy_if_xsel, dim3(misc_blocks(n)), dim3(kBlock), 0, s, dst, src, n, so);
}

void fresh_gate(hipStream_t s, const KScalars* so, const KScalars* si, KScalars* gate) {
  hipLaunchKernelGGL(k_fresh_gate, dim3(1), dim3(1), 0, s, so, si, gate);
}

void finalize_ru(hipStream_t s, KScalars* si, KScalars* so, KScalars* gate,
                 const double* partials, int np, int op, double delta, const double* maxp,
                 int nmax) {
  hipLaunchKernelGGL(k_fin_ru, dim3(1), dim3(1024), 0, s, si, so, gate, partials, np, op, delta,
                     maxp, nmax);
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
  const Cubic cb = cubic_of(nl);
  by_kind_nl(op, nl, [&](auto K, auto M) {
    hipLaunchKernelGGL((k_stage<decltype(K)::value, decltype(M)::value, double>), dim3(blocks),
                       dim3(kBlock), 0, s, op, v, un, acc, out, f, f_const, nl, cb, ut, dt, c1, mode);
  });
}

void time_stage32(hipStream_t s, const DevOp& op, const float* v, const float* un, float* acc,
                  float* out, const double* f, double f_const, const NonlinearTerms& nl, double ut,
                  double dt, double c1, int mode) {
  long long blocks = (op.n + kBlock - 1) / kBlock;
  if (blocks > 8192) blocks = 8192;
  const Cubic cb = cubic_of(nl);
  by_kind_nl(op, nl, [&](auto K, auto M) {
    hipLaunchKernelGGL((k_stage<decltype(K)::value, decltype(M)::value, float>), dim3(blocks),
                       dim3(kBlock), 0, s, op, v, un, acc, out, f, f_const, nl, cb, ut, dt, c1, mode);
  });
}

namespace {
bool g_fused_steps = true;

// Tiles of 54 x TY nodes (64-column extended tiles); TY = 22 keeps each workgroup's LDS at
// 16 KB (FP32) / 32 KB (FP64) and a 2049^2 grid at 3572 workgroups.
template <typename S, int TY>
void rk4_step_launch(hipStream_t s, const DevOp& op, const FaceData& fd, const S* u, S* out,
                     const double* f, double fc, const NonlinearTerms& nl, double ut, double dt) {
  constexpr int TX = 54;
  const int tx = (op.nx + TX - 1) / TX, ty = (op.ny + TY - 1) / TY;
  const dim3 g(static_cast<unsigned>(tx * ty));
  const double h2 = 0.5 * dt, h6 = dt / 6.0;
  const Cubic cb = cubic_of(nl);
  by_kind_nl(op, nl, [&](auto K, auto M) {
    hipLaunchKernelGGL((k_rk4_step<decltype(K)::value, decltype(M)::value, S, TY>), g,
                       dim3(256), 0, s, op, fd, u, out, f, fc, nl, cb, ut, h2, dt, h6, tx);
  });
}
}  // namespace

void set_fused_steps(bool on) { g_fused_steps = on; }
bool fused_steps() { return g_fused_steps; }

void time_rk4_step(hipStream_t s, const DevOp& op, const FaceData& fd, const double* u, double* out,
                   const double* f, double fc, const NonlinearTerms& nl, double ut, double dt) {
  rk4_step_launch<double, 22>(s, op, fd, u, out, f, fc, nl, ut, dt);
}

void time_rk4_step(hipStream_t s, const DevOp& op, const FaceData& fd, const float* u, float* out,
                   const double* f, double fc, const NonlinearTerms& nl, double ut, double dt) {
  rk4_step_launch<float, 22>(s, op, fd, u, out, f, fc, nl, ut, dt);
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

// 3D grids take the plane march (resnorm_plane) when plane stencils are on.  The partials
// count of a 3D grid covers both kernels (the plane march zeroes the ones past its grid), so it
// does not depend on the switch.
static bool resnorm_3d_plane(const DevOp& op) {
  return op.nz >= 3 && op.nx >= 3 && op.ny >= 3 && plane_stencils();
}

int resnorm_blocks(const DevOp& op) {
  long long blocks = (op.n + kBlock - 1) / kBlock;
  const int legacy = static_cast<int>(blocks > 2048 ? 2048 : blocks);
  if (op.nz >= 3 && op.nx >= 3 && op.ny >= 3) return std::max(legacy, plane_blocks(op));
  return legacy;
}

void residual_norms(hipStream_t s, const DevOp& op, const FaceData& fd, const double* u,
                    const double* f, double f_const, const NonlinearTerms& nl, double* sums,
                    double* maxs, int* degenerate) {
  const dim3 g(resnorm_blocks(op));
  const bool d3 = op.nz > 1;
#define RN_(D, K) hipLaunchKernelGGL((k_resnorm<D, K>), g, dim3(kBlock), 0, s, op, fd, u, f, f_const, nl, sums, maxs, degenerate)
#define RN3_(K) hipLaunchKernelGGL((k_resnorm3<K>), g, dim3(kBlock), 0, s, op, u, f, f_const, sums, maxs, degenerate)
  if (resnorm_3d_plane(op) && resnorm_plane(s, op, u, f, f_const, sums, maxs, degenerate)) {
    // launched
  } else if (d3 && op.nx >= 3 && op.ny >= 3 && op.nz >= 3) {
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
