// k_stencil.hip — stencil kernels for gfx950: operator apply, residual, red-black Gauss-Seidel
// and the reference-semantics Jacobi / SOR sweeps.
//
// Mapping.  2D levels: one thread per node, flattened index, grid-stride.  3D levels: one
// thread per (i,j) column of a z-chunk of kZChunk planes ("column march"): consecutive
// threads own consecutive x, so every plane access is a coalesced row segment, and the z
// neighbours of the running node stay in registers (u[k-1], u[k], u[k+1]) instead of being
// re-read.  x/y neighbours come through L1/L2 (the same rows are read by the adjacent
// threads of the block).  Each launch has >= 8 x #CU workgroups for the 257^3 level.
#include <algorithm>

#include <type_traits>

#include "kernels.hpp"

namespace pdehip {
namespace {

constexpr int kZChunk = 16;
bool g_plane = true;

template <Prec P> struct PT;
template <> struct PT<P64> { using T = double; using S = double; using K = double; };
template <> struct PT<P32> { using T = float; using S = float; using K = float; };
template <> struct PT<P16> { using T = float; using S = __half; using K = float; };

template <typename S> __device__ __forceinline__ auto ldT(const S* p, long long i);
template <> __device__ __forceinline__ auto ldT<double>(const double* p, long long i) { return p[i]; }
template <> __device__ __forceinline__ auto ldT<float>(const float* p, long long i) { return p[i]; }
template <> __device__ __forceinline__ auto ldT<__half>(const __half* p, long long i) { return __half2float(p[i]); }

__device__ __forceinline__ bool skip(const KScalars* sc) { return sc && (sc->done | sc->breakdown); }

// FP32/FP16 applies of operators whose rows sum to the reaction term use the difference form
// (apply_diff): the sum form loses ~(8/h^2)/lambda of its FP32 digits on smooth vectors, which
// makes the Krylov recursion drift from the true residual.  FP64 keeps the reference's sum.
template <typename T, int KIND>
constexpr bool kDiff = !std::is_same<T, double>::value &&
                       (KIND == PDEHIP_OP_CONST || KIND == PDEHIP_OP_FLUX);

// ----------------------------------------------------------------------------- apply / residual
// MODE 0: y = A x (partials of x.y).  MODE 1: r = b - A x (partials of r.r).
template <int DIM, int KIND, Prec P, int MODE>
__global__ void __launch_bounds__(kBlock) k_stencil(DevOp op, const void* __restrict__ bv,
                                                   const void* __restrict__ xv,
                                                   void* __restrict__ yv, double* partials,
                                                   const KScalars* sc) {
  using T = typename PT<P>::T;
  using S = typename PT<P>::S;
  using K = typename PT<P>::K;
  __shared__ double lds[kBlock / kWave];
  if (skip(sc)) return;
  const S* __restrict__ x = static_cast<const S*>(xv);
  const S* __restrict__ b = static_cast<const S*>(bv);
  S* __restrict__ y = static_cast<S*>(yv);
  double acc = 0.0;
  if constexpr (DIM == 2) {
    const long long n = op.n;
    for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < n;
         id += (long long)gridDim.x * blockDim.x) {
      const IJK c = coords(op, id);
      const int i = c.i, j = c.j;
      const T xc = ldT(x, id);
      T out;
      if (is_boundary2(op, i, j)) {
        out = (MODE == 0) ? xc : (ldT(b, id) - xc);
      } else {
        W7<T> w = row_weights<2, KIND, T, K>(op, id);
        T ax;
        if constexpr (kDiff<T, KIND>) {
          ax = apply_diff<2, T>(op, w, T(op.e), i, j, 0, xc, ldT(x, id - 1), ldT(x, id + 1),
                                ldT(x, id - op.nx), ldT(x, id + op.nx), T(0), T(0));
        } else {
          fold_center<2, T>(op, i, j, 0, w);
          ax = w.C * xc + w.E * ldT(x, id + 1) + w.W * ldT(x, id - 1) +
               w.N * ldT(x, id + op.nx) + w.S * ldT(x, id - op.nx);
        }
        out = (MODE == 0) ? ax : (ldT(b, id) - ax);
      }
      y[id] = store_as<S>(static_cast<double>(out));
      if (partials) acc += (MODE == 0) ? static_cast<double>(xc) * static_cast<double>(out)
                                       : static_cast<double>(out) * static_cast<double>(out);
    }
  } else {
    const long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (p < op.sxy) {
      const int j = static_cast<int>(static_cast<unsigned>(p) / static_cast<unsigned>(op.nx));
      const int i = static_cast<int>(p) - j * op.nx;
      const bool bxy = is_boundary2(op, i, j);
      const int k0 = blockIdx.y * kZChunk;
      const int k1 = min(k0 + kZChunk, op.nz);
      long long id = k0 * op.sxy + p;
      T xm = (k0 > 0) ? ldT(x, id - op.sxy) : T(0);
      T xc = ldT(x, id);
      for (int k = k0; k < k1; ++k, id += op.sxy) {
        const T xp = (k + 1 < op.nz) ? ldT(x, id + op.sxy) : T(0);
        T out;
        if (bxy || k == 0 || k == op.nz - 1) {
          out = (MODE == 0) ? xc : (ldT(b, id) - xc);
        } else {
          W7<T> w = row_weights<3, KIND, T, K>(op, id);
          T ax;
          if constexpr (kDiff<T, KIND>) {
            ax = apply_diff<3, T>(op, w, T(op.e), i, j, k, xc, ldT(x, id - 1), ldT(x, id + 1),
                                  ldT(x, id - op.nx), ldT(x, id + op.nx), xm, xp);
          } else {
            fold_center<3, T>(op, i, j, k, w);
            ax = w.C * xc + w.E * ldT(x, id + 1) + w.W * ldT(x, id - 1) +
                 w.N * ldT(x, id + op.nx) + w.S * ldT(x, id - op.nx) + w.F * xp + w.B * xm;
          }
          out = (MODE == 0) ? ax : (ldT(b, id) - ax);
        }
        y[id] = store_as<S>(static_cast<double>(out));
        if (partials) acc += (MODE == 0) ? static_cast<double>(xc) * static_cast<double>(out)
                                         : static_cast<double>(out) * static_cast<double>(out);
        xm = xc;
        xc = xp;
      }
    }
  }
  if (partials) {
    const double s = block_sum(acc, lds);
    if (threadIdx.x == 0) partials[blockIdx.y * gridDim.x + blockIdx.x] = s;
  }
}

// ----------------------------------------------------------------------------- red-black GS
template <int DIM, int KIND, Prec P>
__global__ void __launch_bounds__(kBlock) k_rbgs(DevOp op, const void* __restrict__ bv,
                                                void* __restrict__ xv, int color,
                                                const KScalars* sc) {
  using T = typename PT<P>::T;
  using S = typename PT<P>::S;
  using K = typename PT<P>::K;
  if (skip(sc)) return;
  const S* __restrict__ b = static_cast<const S*>(bv);
  S* __restrict__ x = static_cast<S*>(xv);
  if constexpr (DIM == 2) {
    // one thread per node of the colour: row j holds (nx-2)/2 or so colour nodes
    const int half = (op.nx - 1) / 2;  // colour nodes per interior row (upper bound)
    const long long total = (long long)half * (op.ny - 2);
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
      const int j = 1 + static_cast<int>(t / half);
      const int q = static_cast<int>(t % half);
      const int i = 1 + ((1 + j + color) & 1) + 2 * q;
      if (i > op.nx - 2) continue;
      const long long id = (long long)j * op.nx + i;
      W7<T> w = row_weights<2, KIND, T, K, true>(op, id);
      fold_center<2, T>(op, i, j, 0, w);
      const T v = gs_value<2, T>(w, ldT(b, id), ldT(x, id + 1), ldT(x, id - 1), ldT(x, id + op.nx),
                                 ldT(x, id - op.nx), T(0), T(0));
      x[id] = store_as<S>(static_cast<double>(v));
    }
  } else {
    const long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (p >= op.sxy) return;
    const int j = static_cast<int>(static_cast<unsigned>(p) / static_cast<unsigned>(op.nx));
    const int i = static_cast<int>(p) - j * op.nx;
    if (is_boundary2(op, i, j)) return;
    const int k0 = max(1, (int)blockIdx.y * kZChunk);
    const int k1 = min((int)blockIdx.y * kZChunk + kZChunk, op.nz - 1);
    // first k of this colour
    int k = k0 + ((i + j + k0 + color) & 1);
    for (; k < k1; k += 2) {
      const long long id = k * op.sxy + p;
      W7<T> w = row_weights<3, KIND, T, K, true>(op, id);
      fold_center<3, T>(op, i, j, k, w);
      const T v = gs_value<3, T>(w, ldT(b, id), ldT(x, id + 1), ldT(x, id - 1), ldT(x, id + op.nx),
                                 ldT(x, id - op.nx), ldT(x, id + op.sxy), ldT(x, id - op.sxy));
      x[id] = store_as<S>(static_cast<double>(v));
    }
  }
}

// ----------------------------------------------------------------------------- reference relaxation
__device__ __forceinline__ double eval_nonlinear(const NonlinearTerms& nl, double u) {
  double v = 0.0;
  for (int t = 0; t < nl.n; ++t) {
    const pdehip_nonlinear& term = nl.t[t];
    switch (term.kind) {
      case PDEHIP_NL_POWER: v += term.coeff * ipow(u, term.power); break;
      case PDEHIP_NL_SIN: v += term.coeff * sin(u); break;
      case PDEHIP_NL_COS: v += term.coeff * cos(u); break;
      case PDEHIP_NL_EXP: v += term.coeff * exp(u); break;
      default: v += term.coeff * fabs(u); break;
    }
  }
  return v;
}

// Jacobi (reference relaxation.cpp:437-545, solver.cpp:827-959): next = (-(f+N(u)) - sum)/C,
// block max |next - u|.
template <int DIM, int KIND>
__global__ void __launch_bounds__(kBlock) k_jacobi_ref(DevOp op, const double* __restrict__ f,
                                                      double f_const, NonlinearTerms nl,
                                                      const double* __restrict__ x,
                                                      double* __restrict__ next, double* partials,
                                                      int* degenerate, const KScalars* sc) {
  __shared__ double lds[kBlock / kWave];
  if (skip(sc)) return;
  double md = 0.0;
  const long long n = op.n;
  for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < n;
       id += (long long)gridDim.x * blockDim.x) {
    const IJK c = coords(op, id);
    const int i = c.i, j = c.j, k = c.k;
    if (is_boundary2(op, i, j) || (DIM == 3 && (k == 0 || k == op.nz - 1))) continue;
    const W7<double> w = row_weights<DIM, KIND, double, double>(op, id);
    if (fabs(w.C) < 1e-12) { *degenerate = 1; continue; }
    double s = w.E * x[id + 1] + w.W * x[id - 1] + w.N * x[id + op.nx] + w.S * x[id - op.nx];
    if constexpr (DIM == 3) s = s + w.F * x[id + op.sxy] + w.B * x[id - op.sxy];
    const double fv = f ? f[id] : f_const;
    const double nlv = nl.n ? eval_nonlinear(nl, x[id]) : 0.0;
    const double v = (-(fv + nlv) - s) / w.C;
    next[id] = v;
    md = nmax(md, fabs(v - x[id]));
  }
  const double m = block_max(md, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = m;
}

// One colour of red-black SOR with the reference update (relaxation.cpp:281-286).
template <int DIM, int KIND>
__global__ void __launch_bounds__(kBlock) k_sor_rb_ref(DevOp op, const double* __restrict__ f,
                                                      double f_const, NonlinearTerms nl,
                                                      double* __restrict__ x, double omega,
                                                      int color, double* partials, int* degenerate,
                                                      const KScalars* sc) {
  __shared__ double lds[kBlock / kWave];
  if (skip(sc)) return;
  double md = 0.0;
  const long long n = op.n;
  for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < n;
       id += (long long)gridDim.x * blockDim.x) {
    const IJK c = coords(op, id);
    const int i = c.i, j = c.j, k = c.k;
    if (((i + j + k) & 1) != color) continue;
    if (is_boundary2(op, i, j) || (DIM == 3 && (k == 0 || k == op.nz - 1))) continue;
    const W7<double> w = row_weights<DIM, KIND, double, double>(op, id);
    if (fabs(w.C) < 1e-12) { *degenerate = 1; continue; }
    double s = w.E * x[id + 1] + w.W * x[id - 1] + w.N * x[id + op.nx] + w.S * x[id - op.nx];
    if constexpr (DIM == 3) s = s + w.F * x[id + op.sxy] + w.B * x[id - op.sxy];
    const double fv = f ? f[id] : f_const;
    const double old = x[id];
    const double nlv = nl.n ? eval_nonlinear(nl, old) : 0.0;
    const double jac = (-(fv + nlv) - s) / w.C;
    const double upd = (1.0 - omega) * old + omega * jac;
    x[id] = upd;
    md = nmax(md, fabs(upd - old));
  }
  const double m = block_max(md, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = nmax(m, color ? partials[blockIdx.x] : 0.0);
}

// ----------------------------------------------------------------------------- dispatch
dim3 grid_for(const DevOp& op) {
  if (op.nz > 1) {
    const unsigned gx = static_cast<unsigned>((op.sxy + kBlock - 1) / kBlock);
    const unsigned gy = static_cast<unsigned>((op.nz + kZChunk - 1) / kZChunk);
    return dim3(gx, gy, 1);
  }
  long long blocks = (op.n + kBlock - 1) / kBlock;
  if (blocks > kMaxPartials) blocks = kMaxPartials;
  return dim3(static_cast<unsigned>(blocks), 1, 1);
}

template <int MODE, int DIM, int KIND>
void launch_stencil_p(const Launch& L, const DevOp& op, Prec p, const void* b, const void* x,
                      void* y, double* partials) {
  const dim3 g = grid_for(op);
  switch (p) {
    case P64: hipLaunchKernelGGL((k_stencil<DIM, KIND, P64, MODE>), g, dim3(kBlock), 0, L.stream, op, b, x, y, partials, L.sc); break;
    case P32: hipLaunchKernelGGL((k_stencil<DIM, KIND, P32, MODE>), g, dim3(kBlock), 0, L.stream, op, b, x, y, partials, L.sc); break;
    default: hipLaunchKernelGGL((k_stencil<DIM, KIND, P16, MODE>), g, dim3(kBlock), 0, L.stream, op, b, x, y, partials, L.sc); break;
  }
}

template <int MODE>
void launch_stencil(const Launch& L, const DevOp& op, Prec p, const void* b, const void* x, void* y,
                    double* partials) {
  const bool d3 = op.nz > 1;
  switch (op.kind) {
    case PDEHIP_OP_CONST:
      if (d3) launch_stencil_p<MODE, 3, PDEHIP_OP_CONST>(L, op, p, b, x, y, partials);
      else launch_stencil_p<MODE, 2, PDEHIP_OP_CONST>(L, op, p, b, x, y, partials);
      break;
    case PDEHIP_OP_POINTWISE:
      if (d3) launch_stencil_p<MODE, 3, PDEHIP_OP_POINTWISE>(L, op, P64, b, x, y, partials);
      else launch_stencil_p<MODE, 2, PDEHIP_OP_POINTWISE>(L, op, P64, b, x, y, partials);
      break;
    default:
      if (d3) launch_stencil_p<MODE, 3, PDEHIP_OP_FLUX>(L, op, p, b, x, y, partials);
      else launch_stencil_p<MODE, 2, PDEHIP_OP_FLUX>(L, op, p, b, x, y, partials);
  }
}

template <int DIM, int KIND>
void launch_rbgs_p(const Launch& L, const DevOp& op, Prec p, const void* b, void* x, int color) {
  dim3 g;
  if (DIM == 3) {
    g = grid_for(op);
  } else {
    const long long total = (long long)((op.nx - 1) / 2) * (op.ny - 2);
    long long blocks = (total + kBlock - 1) / kBlock;
    if (blocks < 1) blocks = 1;
    if (blocks > 65535) blocks = 65535;
    g = dim3(static_cast<unsigned>(blocks));
  }
  switch (p) {
    case P64: hipLaunchKernelGGL((k_rbgs<DIM, KIND, P64>), g, dim3(kBlock), 0, L.stream, op, b, x, color, L.sc); break;
    case P32: hipLaunchKernelGGL((k_rbgs<DIM, KIND, P32>), g, dim3(kBlock), 0, L.stream, op, b, x, color, L.sc); break;
    default: hipLaunchKernelGGL((k_rbgs<DIM, KIND, P16>), g, dim3(kBlock), 0, L.stream, op, b, x, color, L.sc); break;
  }
}

}  // namespace

// 3D levels use the plane-march kernels (k_plane.hip); 2D levels the grid-stride ones here.
int partial_blocks(const DevOp& op) {
  if (op.nz > 1 && g_plane) return plane_blocks(op);
  const dim3 g = grid_for(op);
  return static_cast<int>(g.x * g.y);
}

void apply(const Launch& L, const DevOp& op, Prec p, const void* x, void* y, double* partials) {
  if (op.nz > 1 && g_plane) apply_plane(L, op, p, x, y, partials);
  else launch_stencil<0>(L, op, p, nullptr, x, y, partials);
}

void residual(const Launch& L, const DevOp& op, Prec p, const void* b, const void* x, void* r,
              double* partials) {
  if (op.nz > 1 && g_plane) residual_plane(L, op, p, b, x, r, partials);
  else launch_stencil<1>(L, op, p, b, x, r, partials);
}

int partial_blocks_max(const DevOp& op) {
  const dim3 g = grid_for(op);
  const int legacy = static_cast<int>(g.x * g.y);
  return op.nz > 1 ? std::max(legacy, plane_blocks(op)) : legacy;
}

void set_plane_stencils(bool on) { g_plane = on; }
bool plane_stencils() { return g_plane; }

void rbgs(const Launch& L, const DevOp& op, Prec p, const void* b, void* x, int color) {
  const bool d3 = op.nz > 1;
  switch (op.kind) {
    case PDEHIP_OP_CONST:
      if (d3) launch_rbgs_p<3, PDEHIP_OP_CONST>(L, op, p, b, x, color);
      else launch_rbgs_p<2, PDEHIP_OP_CONST>(L, op, p, b, x, color);
      break;
    case PDEHIP_OP_POINTWISE:
      if (d3) launch_rbgs_p<3, PDEHIP_OP_POINTWISE>(L, op, P64, b, x, color);
      else launch_rbgs_p<2, PDEHIP_OP_POINTWISE>(L, op, P64, b, x, color);
      break;
    default:
      if (d3) launch_rbgs_p<3, PDEHIP_OP_FLUX>(L, op, p, b, x, color);
      else launch_rbgs_p<2, PDEHIP_OP_FLUX>(L, op, p, b, x, color);
  }
}

int relax_blocks(const DevOp& op) {
  long long blocks = (op.n + kBlock - 1) / kBlock;
  if (blocks > 2048) blocks = 2048;
  return static_cast<int>(blocks);
}

void jacobi_ref(const Launch& L, const DevOp& op, const double* f, double f_const,
                const NonlinearTerms& nl, const double* x, double* next, double* partials,
                int* degenerate) {
  const dim3 g(relax_blocks(op));
  const bool d3 = op.nz > 1;
  if (op.kind == PDEHIP_OP_POINTWISE) {
    if (d3) hipLaunchKernelGGL((k_jacobi_ref<3, PDEHIP_OP_POINTWISE>), g, dim3(kBlock), 0, L.stream, op, f, f_const, nl, x, next, partials, degenerate, L.sc);
    else hipLaunchKernelGGL((k_jacobi_ref<2, PDEHIP_OP_POINTWISE>), g, dim3(kBlock), 0, L.stream, op, f, f_const, nl, x, next, partials, degenerate, L.sc);
  } else if (op.kind == PDEHIP_OP_FLUX) {
    if (d3) hipLaunchKernelGGL((k_jacobi_ref<3, PDEHIP_OP_FLUX>), g, dim3(kBlock), 0, L.stream, op, f, f_const, nl, x, next, partials, degenerate, L.sc);
    else hipLaunchKernelGGL((k_jacobi_ref<2, PDEHIP_OP_FLUX>), g, dim3(kBlock), 0, L.stream, op, f, f_const, nl, x, next, partials, degenerate, L.sc);
  } else {
    if (d3) hipLaunchKernelGGL((k_jacobi_ref<3, PDEHIP_OP_CONST>), g, dim3(kBlock), 0, L.stream, op, f, f_const, nl, x, next, partials, degenerate, L.sc);
    else hipLaunchKernelGGL((k_jacobi_ref<2, PDEHIP_OP_CONST>), g, dim3(kBlock), 0, L.stream, op, f, f_const, nl, x, next, partials, degenerate, L.sc);
  }
}

void sor_rb_ref(const Launch& L, const DevOp& op, const double* f, double f_const,
                const NonlinearTerms& nl, double* x, double omega, int color, double* partials,
                int* degenerate) {
  const dim3 g(relax_blocks(op));
  const bool d3 = op.nz > 1;
  if (op.kind == PDEHIP_OP_POINTWISE) {
    if (d3) hipLaunchKernelGGL((k_sor_rb_ref<3, PDEHIP_OP_POINTWISE>), g, dim3(kBlock), 0, L.stream, op, f, f_const, nl, x, omega, color, partials, degenerate, L.sc);
    else hipLaunchKernelGGL((k_sor_rb_ref<2, PDEHIP_OP_POINTWISE>), g, dim3(kBlock), 0, L.stream, op, f, f_const, nl, x, omega, color, partials, degenerate, L.sc);
  } else if (op.kind == PDEHIP_OP_FLUX) {
    if (d3) hipLaunchKernelGGL((k_sor_rb_ref<3, PDEHIP_OP_FLUX>), g, dim3(kBlock), 0, L.stream, op, f, f_const, nl, x, omega, color, partials, degenerate, L.sc);
    else hipLaunchKernelGGL((k_sor_rb_ref<2, PDEHIP_OP_FLUX>), g, dim3(kBlock), 0, L.stream, op, f, f_const, nl, x, omega, color, partials, degenerate, L.sc);
  } else {
    if (d3) hipLaunchKernelGGL((k_sor_rb_ref<3, PDEHIP_OP_CONST>), g, dim3(kBlock), 0, L.stream, op, f, f_const, nl, x, omega, color, partials, degenerate, L.sc);
    else hipLaunchKernelGGL((k_sor_rb_ref<2, PDEHIP_OP_CONST>), g, dim3(kBlock), 0, L.stream, op, f, f_const, nl, x, omega, color, partials, degenerate, L.sc);
  }
}

}  // namespace pdehip
