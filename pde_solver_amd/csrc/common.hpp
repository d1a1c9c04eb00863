// common.hpp — device-side definitions shared by every HIP translation unit of libpdehip.
//
// Grid layout (all levels, all precisions): full vertex-centred grid including boundary
// nodes, x fastest, idx = (k*ny + j)*nx + i (reference cpu_utils.h:13-19).  Boundary rows are
// identity; interior rows are a 5-point (2D) or 7-point (3D) stencil.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pdehip.h"

namespace pdehip {

constexpr int kWave = 64;          // CDNA wavefront
constexpr int kBlock = 256;        // default workgroup (4 waves)
constexpr int kMaxPartials = 4096; // stage-1 reduction partials per dot product

// ------------------------------------------------------------------ operator descriptor
// One level's operator, passed to kernels by value.
struct DevOp {
  int nx, ny, nz;        // nz == 1: 2D
  int kind;              // pdehip_op_kind
  long long sxy;         // nx*ny
  long long n;           // nx*ny*nz
  double w[7];           // CONST weights: C, W, E, S, N, B(k-1), F(k+1)
  const double* pw;      // POINTWISE weights, structure of arrays [7][n]
  const void* kappa;     // FLUX: kappa per node (FP64 levels) or 1/kappa (FP32/FP16 levels)
  double fx, fy, fz;     // FLUX: a/h^2 per axis
  double e;              // FLUX: constant reaction term
  double theta[6];       // folded closure factors (0 = not folded / Dirichlet)
  double phi[6];         // 1 - theta, formed in FP64 (1 = Dirichlet / not folded)
};

template <typename T>
struct W7 {
  T C, W, E, S, N, B, F;
};

template <typename T> __device__ __forceinline__ T ld(const T* p, long long i) { return p[i]; }
__device__ __forceinline__ float ldf(const __half* p, long long i) { return __half2float(p[i]); }
__device__ __forceinline__ float ldf(const float* p, long long i) { return p[i]; }
__device__ __forceinline__ double ldf(const double* p, long long i) { return p[i]; }

template <typename T> struct Conv;
template <> struct Conv<double> {
  __device__ static double from(double v) { return v; }
  __device__ static double from(float v) { return static_cast<double>(v); }
  __device__ static double from(__half v) { return static_cast<double>(__half2float(v)); }
};
template <> struct Conv<float> {
  __device__ static float from(double v) { return static_cast<float>(v); }
  __device__ static float from(float v) { return v; }
  __device__ static float from(__half v) { return __half2float(v); }
};

template <typename S> __device__ __forceinline__ S store_as(double v);
template <> __device__ __forceinline__ double store_as<double>(double v) { return v; }
template <> __device__ __forceinline__ float store_as<float>(double v) { return static_cast<float>(v); }
template <> __device__ __forceinline__ __half store_as<__half>(double v) { return __float2half(static_cast<float>(v)); }
template <typename S> __device__ __forceinline__ S store_f(float v);
template <> __device__ __forceinline__ double store_f<double>(float v) { return v; }
template <> __device__ __forceinline__ float store_f<float>(float v) { return v; }
template <> __device__ __forceinline__ __half store_f<__half>(float v) { return __float2half(v); }

// Reciprocal for the smoothers: the hardware estimate (float, 1 ulp) or the estimate refined by
// two Newton steps (double).  The FP64 operator itself (apply, residual) divides exactly; a
// smoother only has to be the same in every kernel that applies it.
template <typename T> __device__ __forceinline__ T recip(T x);
template <> __device__ __forceinline__ float recip<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <> __device__ __forceinline__ double recip<double>(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  r = fma(r, fma(-x, r, 1.0), r);
  return r;
}

// Face coefficient: the axis factor a/h^2 times the harmonic mean of the two node
// conductivities.  FP64 levels store kappa and divide exactly.  FP32/FP16 levels store
// rho = 1/kappa, where the harmonic mean is 2/(rho_c + rho_n): one add and one hardware
// reciprocal (kappa = 0 gives rho = inf and a zero face).  Both forms are symmetric bit for bit
// in their two arguments, so a face computed from either side agrees.
template <typename T> __device__ __forceinline__ T face_coef(T kc, T kn, T f);
template <> __device__ __forceinline__ double face_coef<double>(double kc, double kn, double f) {
  const double s = kc + kn;
  return s > 0.0 ? f * (2.0 * kc * kn) / s : 0.0;
}
template <> __device__ __forceinline__ float face_coef<float>(float rc, float rn, float f) {
  // rounded here: a kernel that sums faces straight from registers must not fuse this product
  // into the sum (an FMA skips its rounding) where a kernel that keeps faces in LDS cannot, or
  // the two round the same operator differently whenever f is not a power of two
#pragma clang fp contract(off)
  return (f + f) * __builtin_amdgcn_rcpf(rc + rn);
}

// Smoother face coefficient: FP64 with the division replaced by recip(); FP32 as face_coef.
template <typename T> __device__ __forceinline__ T face_smooth(T kc, T kn, T f);
template <> __device__ __forceinline__ double face_smooth<double>(double kc, double kn, double f) {
  const double s = kc + kn;
  return s > 0.0 ? f * (2.0 * kc * kn) * recip<double>(s) : 0.0;
}
template <> __device__ __forceinline__ float face_smooth<float>(float rc, float rn, float f) {
  return face_coef<float>(rc, rn, f);
}

// Weights of interior row (i,j,k) before folding.  SMOOTH: faces by face_smooth (the
// smoothers' operator), else by face_coef (the exact operator).
template <int DIM, int KIND, typename T, typename TK, bool SMOOTH = false>
__device__ __forceinline__ W7<T> row_weights(const DevOp& op, long long id) {
  W7<T> w;
  if constexpr (KIND == PDEHIP_OP_CONST) {
    w.C = T(op.w[0]); w.W = T(op.w[1]); w.E = T(op.w[2]); w.S = T(op.w[3]); w.N = T(op.w[4]);
    w.B = T(op.w[5]); w.F = T(op.w[6]);
  } else if constexpr (KIND == PDEHIP_OP_POINTWISE) {
    const double* p = op.pw;
    const long long n = op.n;
    w.C = T(p[id]); w.W = T(p[n + id]); w.E = T(p[2 * n + id]); w.S = T(p[3 * n + id]);
    w.N = T(p[4 * n + id]);
    if constexpr (DIM == 3) { w.B = T(p[5 * n + id]); w.F = T(p[6 * n + id]); }
    else { w.B = T(0); w.F = T(0); }
  } else {
    const TK* k = static_cast<const TK*>(op.kappa);
    const T kc = Conv<T>::from(k[id]);
    const T fx = T(op.fx), fy = T(op.fy);
    auto face = [](T a, T b, T f) { return SMOOTH ? face_smooth<T>(a, b, f) : face_coef<T>(a, b, f); };
    w.W = face(kc, Conv<T>::from(k[id - 1]), fx);
    w.E = face(kc, Conv<T>::from(k[id + 1]), fx);
    w.S = face(kc, Conv<T>::from(k[id - op.nx]), fy);
    w.N = face(kc, Conv<T>::from(k[id + op.nx]), fy);
    T sum = w.W + w.E + w.S + w.N;
    if constexpr (DIM == 3) {
      const T fz = T(op.fz);
      w.B = face(kc, Conv<T>::from(k[id - op.sxy]), fz);
      w.F = face(kc, Conv<T>::from(k[id + op.sxy]), fz);
      sum += w.B + w.F;
    } else {
      w.B = T(0); w.F = T(0);
    }
    w.C = T(op.e) - sum;
  }
  return w;
}

// Folding: an interior row next to a folded face adds w_nb*theta_f to its centre (explicit FMA,
// so every kernel that folds a row rounds it the same way).
template <int DIM, typename T>
__device__ __forceinline__ void fold_center(const DevOp& op, int i, int j, int k, W7<T>& w) {
  if (i == 1) w.C = fma(w.W, T(op.theta[0]), w.C);
  if (i == op.nx - 2) w.C = fma(w.E, T(op.theta[1]), w.C);
  if (j == 1) w.C = fma(w.S, T(op.theta[2]), w.C);
  if (j == op.ny - 2) w.C = fma(w.N, T(op.theta[3]), w.C);
  if constexpr (DIM == 3) {
    if (k == 1) w.C = fma(w.B, T(op.theta[4]), w.C);
    if (k == op.nz - 2) w.C = fma(w.F, T(op.theta[5]), w.C);
  }
}

// A x on one interior row in difference form, for operators whose rows sum to the reaction
// term e (constant coefficients without variable reaction, flux form):
//   A x = e x_c - sum_nb w_nb (x_c - x_nb),
// where a neighbour on a folded face contributes x_c - x_nb -> phi x_c - x_nb (phi = 1 - theta;
// x_nb is that boundary node's stored value).  The same operator as C' x_c + sum w x with the
// folded centre C', but for smooth x every term is small, so the FP32 product keeps its relative
// accuracy where the sum form cancels ~ (8/h^2) / lambda of it away.  `w` holds the unfolded
// neighbour weights.
template <int DIM, typename T>
__device__ __forceinline__ T apply_diff(const DevOp& op, const W7<T>& w, T e, int i, int j, int k,
                                        T xc, T xw, T xe, T xs, T xn, T xb, T xf) {
  T cw = xc, ce = xc, cs = xc, cn = xc, cb = xc, cf = xc;
  if (i == 1) cw = T(op.phi[0]) * xc;
  if (i == op.nx - 2) ce = T(op.phi[1]) * xc;
  if (j == 1) cs = T(op.phi[2]) * xc;
  if (j == op.ny - 2) cn = T(op.phi[3]) * xc;
  T a = e * xc;
  a = fma(-w.E, ce - xe, a);
  a = fma(-w.W, cw - xw, a);
  a = fma(-w.N, cn - xn, a);
  a = fma(-w.S, cs - xs, a);
  if constexpr (DIM == 3) {
    if (k == 1) cb = T(op.phi[4]) * xc;
    if (k == op.nz - 2) cf = T(op.phi[5]) * xc;
    a = fma(-w.F, cf - xf, a);
    a = fma(-w.B, cb - xb, a);
  }
  return a;
}

// A x on one interior row as a fixed FMA chain (centre, E, W, N, S, F, B), shared by every
// kernel that forms the operator from weights at hand, so they round it the same way.
template <typename T>
__device__ __forceinline__ T apply_row(const W7<T>& w, T xc, T xe, T xw, T xn, T xs, T xf, T xb) {
  T a = w.C * xc;
  a = fma(w.E, xe, a);
  a = fma(w.W, xw, a);
  a = fma(w.N, xn, a);
  a = fma(w.S, xs, a);
  a = fma(w.F, xf, a);
  a = fma(w.B, xb, a);
  return a;
}

// Gauss-Seidel value of one interior row: (b - sum_nb w_nb u_nb) * recip(C), the neighbour sum
// as a fixed FMA chain (E, W, N, S, F, B).  Every smoother kernel uses this, so the fused and
// the in-place sweeps agree bit for bit whatever the compiler's contraction choices.
template <int DIM, typename T>
__device__ __forceinline__ T gs_sum(const W7<T>& w, T ue, T uw, T un, T us, T uf, T ub) {
  T s = w.E * ue;
  s = fma(w.W, uw, s);
  s = fma(w.N, un, s);
  s = fma(w.S, us, s);
  if constexpr (DIM == 3) {
    s = fma(w.F, uf, s);
    s = fma(w.B, ub, s);
  }
  return s;
}
template <int DIM, typename T>
__device__ __forceinline__ T gs_value(const W7<T>& w, T b, T ue, T uw, T un, T us, T uf, T ub) {
  return (b - gs_sum<DIM, T>(w, ue, uw, un, us, uf, ub)) * recip<T>(w.C);
}

__device__ __forceinline__ bool is_boundary2(const DevOp& op, int i, int j) {
  return i == 0 || j == 0 || i == op.nx - 1 || j == op.ny - 1;
}

// ------------------------------------------------------------------ launch shaping (host)
// z-chunk count for a plane-march launch: `tiles` (i,j) tiles each split into chunks of
// ceil(nz / chunks) planes plus `prologue` warm-up planes; `capacity` workgroups run at once.
// Minimises rounds x planes-per-workgroup, so the last round is not a thin tail.
inline int plane_chunk(long long tiles, int nz, int capacity, int prologue, int min_planes) {
  int best_kc = nz;
  long long best = -1;
  for (int chunks = 1; chunks <= nz; ++chunks) {
    const int kc = (nz + chunks - 1) / chunks;
    if (kc < min_planes && chunks > 1) break;
    const long long n = tiles * ((nz + kc - 1) / kc);
    const long long rounds = (n + capacity - 1) / capacity;
    const long long cost = rounds * (kc + prologue);
    if (best < 0 || cost < best) { best = cost; best_kc = kc; }
  }
  return best_kc;
}

// ------------------------------------------------------------------ reductions
// u^p for an integer p by repeated squaring (|p| <= 16; pow beyond): u^1 and u^2 are the
// correctly rounded values pow gives, higher powers differ by a few ulp at most.
__device__ __forceinline__ double ipow(double u, int p) {
  if (p > 16 || p < -16) return pow(u, static_cast<double>(p));
  const bool inv = p < 0;
  unsigned e = static_cast<unsigned>(inv ? -p : p);
  double r = 1.0, b = u;
  while (e) {
    if (e & 1u) r *= b;
    b *= b;
    e >>= 1;
  }
  return inv ? 1.0 / r : r;
}

// Grid coordinates (i, j, k) of node id.  Grids of fewer than 2^31 nodes (every 2D grid up to
// 46341^2, 3D up to 1290^3) use 32-bit unsigned divisions, a few instructions where the 64-bit
// ones cost dozens per node; larger grids take the 64-bit path.
struct IJK { int i, j, k; };
__device__ __forceinline__ IJK coords(const DevOp& op, long long id) {
  IJK c;
  if (op.n < 2147483648LL) {
    const unsigned u = static_cast<unsigned>(id);
    const unsigned nx = static_cast<unsigned>(op.nx), ny = static_cast<unsigned>(op.ny);
    const unsigned row = u / nx;
    c.i = static_cast<int>(u - row * nx);
    c.j = static_cast<int>(row % ny);
    c.k = static_cast<int>(row / ny);
  } else {
    const long long row = id / op.nx;
    c.i = static_cast<int>(id - row * op.nx);
    c.j = static_cast<int>(row % op.ny);
    c.k = static_cast<int>(row / op.ny);
  }
  return c;
}

// Maximum that propagates NaN (fmax drops it), so a non-finite value in a max-reduction
// reaches the stop rule instead of reading as a small norm.
template <typename T>
__device__ __forceinline__ T nmax(T a, T b) { return (a != a || a > b) ? a : b; }

// Non-finite reduction result (NaN or +-inf).
__device__ __forceinline__ bool nonfinite(double v) { return !(fabs(v) <= 1.7976931348623157e308); }

// Sum over the 64 lanes of a wavefront.
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v = nmax(v, __shfl_xor(v, off, kWave));
  return v;
}

// Block sum: wave shuffle, then one LDS slot per wave, then the first wave.  Result valid in
// thread 0.  `lds` must hold blockDim.x/64 elements.
template <typename T>
__device__ __forceinline__ T block_sum(T v, T* lds) {
  v = wave_sum(v);
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  const int nw = (blockDim.x + kWave - 1) / kWave;
  T r = (threadIdx.x < nw) ? lds[threadIdx.x] : T(0);
  if (wid == 0) r = wave_sum(r);
  __syncthreads();
  return r;
}
template <typename T>
__device__ __forceinline__ T block_max(T v, T* lds) {
  v = wave_max(v);
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  const int nw = (blockDim.x + kWave - 1) / kWave;
  T r = (threadIdx.x < nw) ? lds[threadIdx.x] : T(0);
  if (wid == 0) r = wave_max(r);
  __syncthreads();
  return r;
}

// ------------------------------------------------------------------ device scalars
// Scalars of one Krylov solve, resident in device memory so iterations never wait on the host.
struct KScalars {
  double rz, rz_old, pq, rr, alpha, beta, omega, rho, rho_old, tt, ts;
  double r0, target, metric;
  double eps;        // breakdown threshold on denominators (reference 1e-30)
  double scale;      // mixed precision: normalisation of the preconditioner input
  double inv_scale;
  double snorm;
  int done;          // 1 once the stop rule holds; later kernels return immediately
  int breakdown;     // 1 when a denominator fell below eps
  int iter;          // completed iterations
  int flag_s;        // BiCGStab: ||s|| met the target, x += alpha*p_hat pending
  int cycle_done;    // GMRES: the current restart cycle is finished
  int steps;         // GMRES: Arnoldi steps of the current cycle
  int pad0, pad1;
  double best;       // mixed BiCGStab: smallest true residual so far ...
  int best_iter;     // ... reached at this iteration
  int stalled;       // 1 once it has not halved for the stall window
  int xsel;          // mixed CG, fused replacement: which of the two FP64 solution buffers is current
  int fresh;         // ... 1 from a replacement until the next FP32 update: the FP32 correction is void
};

}  // namespace pdehip
