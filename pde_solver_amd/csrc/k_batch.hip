// k_batch.hip — batched kernels for many independent 2D constant-coefficient problems.
//
// Every launch walks a list of work items; an item belongs to one problem (BItem.p) and is
// processed by one wavefront (stencil, transfer and smoother items) or one workgroup (vector
// items).  Per-problem data comes from a device table (BEnt per problem and level, BVec per
// problem), so problems of any size share a launch.  Items of a problem whose scalars say
// "done" return at once.  Reductions write one partial per item; items of a problem are
// contiguous, and the segmented finalize sums each problem's range in a fixed order.
#include <type_traits>

#include "batch.hpp"
#include "faces_dev.hpp"

namespace pdehip {
namespace bk {
namespace {

constexpr int kLanes = 64;
// Strips: each lane holds two adjacent columns, so a wavefront covers 128 columns, and the
// 124 columns away from the 2-column ring on each side are output.
constexpr int kTileX = 2 * kLanes - 4;
constexpr int kWavesPB = 4;          // wavefront items per 256-thread workgroup
constexpr int kAhead = 1;            // row marches: rows loaded ahead of the row in use

template <Prec P> struct PT;
template <> struct PT<P64> { using T = double; using S = double; };
template <> struct PT<P32> { using T = float; using S = float; };
template <> struct PT<P16> { using T = float; using S = __half; };

// Buffer access.  Pointers read from the problem tables are generic (flat) to the compiler, and
// flat loads count against the LDS/scalar counter as well as the vector-memory one, so every
// shuffle or scalar wait would also wait for the rows in flight.  Every batch buffer is device
// memory: loads and stores go through the global address space.
#define PDEHIP_GAS __attribute__((address_space(1)))
template <typename S> __device__ __forceinline__ auto ldS(const S* p, long long i);
template <> __device__ __forceinline__ auto ldS<double>(const double* p, long long i) {
  return ((const PDEHIP_GAS double*)p)[i];
}
template <> __device__ __forceinline__ auto ldS<float>(const float* p, long long i) {
  return ((const PDEHIP_GAS float*)p)[i];
}
template <> __device__ __forceinline__ auto ldS<__half>(const __half* p, long long i) {
  return __half2float(__ushort_as_half(((const PDEHIP_GAS unsigned short*)p)[i]));
}
template <> __device__ __forceinline__ auto ldS<int>(const int* p, long long i) {
  return ((const PDEHIP_GAS int*)p)[i];
}
// Two adjacent elements p[i], p[i+1] in one global load / store (dwordx4, x2 or x1): two-lane
// vectors with the element's own alignment, which global accesses of this width accept.
template <typename S> struct V2;
template <> struct V2<double> { typedef double type __attribute__((ext_vector_type(2))) __attribute__((aligned(8))); };
template <> struct V2<float> { typedef float type __attribute__((ext_vector_type(2))) __attribute__((aligned(4))); };
template <> struct V2<__half> {
  typedef unsigned short type __attribute__((ext_vector_type(2))) __attribute__((aligned(2)));
};
template <typename S> struct Pair2 { S a, b; };
template <typename S> __device__ __forceinline__ Pair2<S> ld2(const S* p, long long i) {
  const typename V2<S>::type v = *((const PDEHIP_GAS typename V2<S>::type*)(p + i));
  if constexpr (std::is_same<S, __half>::value) return Pair2<S>{__ushort_as_half(v.x), __ushort_as_half(v.y)};
  else return Pair2<S>{v.x, v.y};
}
template <typename S> __device__ __forceinline__ void st2(S* p, long long i, S a, S b) {
  typename V2<S>::type v;
  if constexpr (std::is_same<S, __half>::value) { v.x = __half_as_ushort(a); v.y = __half_as_ushort(b); }
  else { v.x = a; v.y = b; }
  *((PDEHIP_GAS typename V2<S>::type*)(p + i)) = v;
}
__device__ __forceinline__ void stS(double* p, long long i, double v) { ((PDEHIP_GAS double*)p)[i] = v; }
__device__ __forceinline__ void stS(float* p, long long i, float v) { ((PDEHIP_GAS float*)p)[i] = v; }
__device__ __forceinline__ void stS(__half* p, long long i, __half v) {
  ((PDEHIP_GAS unsigned short*)p)[i] = __half_as_ushort(v);
}

template <typename S, typename T> __device__ __forceinline__ T rounded(T v);
template <> __device__ __forceinline__ double rounded<double, double>(double v) { return v; }
template <> __device__ __forceinline__ float rounded<float, float>(float v) { return v; }
template <> __device__ __forceinline__ float rounded<__half, float>(float v) { return __half2float(__float2half(v)); }

__device__ __forceinline__ bool done(const KScalars* sc) {
  return sc && (ldS(&sc->done, 0) | ldS(&sc->breakdown, 0));
}

// Wavefront item index.  Workgroup slot b is dispatched to XCD b % 8; numbering the
// workgroups XCD-contiguously ((b % 8) * per + b / 8, gridDim.x = 8 * per) puts consecutive
// items (neighbouring strips and row chunks of one problem, which share halo rows and cache
// lines) in the same XCD's L2.  Indices past the list return in the callers (w >= n).
// The index is wave-uniform (readfirstlane), so the item and table reads are scalar loads.
__device__ __forceinline__ int wave_item() {
  const int per = static_cast<int>(gridDim.x >> 3);
  const int t = static_cast<int>(blockIdx.x & 7) * per + static_cast<int>(blockIdx.x >> 3);
  return __builtin_amdgcn_readfirstlane(t * kWavesPB + static_cast<int>(threadIdx.x / kLanes));
}

template <typename T>
__device__ __forceinline__ W7<T> const_w(const DevOp& op) {
  W7<T> w;
  w.C = T(op.w[0]); w.W = T(op.w[1]); w.E = T(op.w[2]); w.S = T(op.w[3]); w.N = T(op.w[4]);
  w.B = T(0); w.F = T(0);
  return w;
}

// Neighbour lanes by DPP wave shifts with zero fill (one VALU op, or two for a double, instead
// of an LDS permute): lane l gets lane l-1's value (from_west; lane 0: 0) or lane l+1's
// (from_east; lane 63: 0).  Lanes 0 and 63 hold a strip's halo columns, which are never stored.
__device__ __forceinline__ float from_west(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x138,
                                                            0xf, 0xf, true));
}
__device__ __forceinline__ float from_east(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x130,
                                                            0xf, 0xf, true));
}
__device__ __forceinline__ double dpp_d(double v, int ctrl_west) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const int lo = static_cast<int>(u & 0xffffffffu), hi = static_cast<int>(u >> 32);
  const int rl = ctrl_west ? __builtin_amdgcn_mov_dpp(lo, 0x138, 0xf, 0xf, true)
                           : __builtin_amdgcn_mov_dpp(lo, 0x130, 0xf, 0xf, true);
  const int rh = ctrl_west ? __builtin_amdgcn_mov_dpp(hi, 0x138, 0xf, 0xf, true)
                           : __builtin_amdgcn_mov_dpp(hi, 0x130, 0xf, 0xf, true);
  return __builtin_bit_cast(double, (static_cast<unsigned long long>(static_cast<unsigned>(rh)) << 32) |
                                        static_cast<unsigned>(rl));
}
__device__ __forceinline__ double from_west(double v) { return dpp_d(v, 1); }
__device__ __forceinline__ double from_east(double v) { return dpp_d(v, 0); }

template <typename T, typename S> __device__ __forceinline__ T cvt(S v);
template <> __device__ __forceinline__ double cvt<double, double>(double v) { return v; }
template <> __device__ __forceinline__ float cvt<float, float>(float v) { return v; }
template <> __device__ __forceinline__ float cvt<float, __half>(__half v) { return __half2float(v); }

// Columns c0, c0+1 (c0 even, possibly outside the grid) of the row at offset rowoff, clamped into
// the grid, from one paired load of columns lo, lo+1 with lo = clamp(c0, 0, nx-2).
template <typename T, typename S>
__device__ __forceinline__ void ld_cols(const S* p, long long rowoff, int c0, int nx, T& v0, T& v1) {
  const int lo = min(max(c0, 0), nx - 2);
  const Pair2<S> q = ld2(p, rowoff + lo);
  const T a = cvt<T, S>(q.a), b = cvt<T, S>(q.b);
  v0 = min(max(c0, 0), nx - 1) == lo ? a : b;
  v1 = min(max(c0 + 1, 0), nx - 1) == lo ? a : b;
}

// Strip march over item (p, strip a, rows [b, c)): lane L holds columns c0 = a*124-2+2L and
// c0+1; x rows j-1, j, j+1 (and the right-hand side row j) sit in registers, the next rows are
// loaded kAhead steps ahead, the x-neighbours across lanes come by shuffles.
// f(i, j, id, xc, xw, xe, xs, xn, bj) runs on the 124 output columns (lanes 1..62) inside the
// grid.  Loads are clamped to the grid.
template <typename T, typename S, typename SB, typename F>
__device__ __forceinline__ void strip_march(const S* __restrict__ x, const SB* __restrict__ b,
                                            int nx, int ny, const BItem& it, F&& f) {
  const int lane = threadIdx.x & (kLanes - 1);
  const int c0 = it.a * kTileX - 2 + 2 * lane;
  const bool out = lane >= 1 && lane < kLanes - 1;
  const bool out0 = out && c0 < nx, out1 = out && c0 + 1 < nx;
  auto row = [&](int j) -> long long { return (long long)min(max(j, 0), ny - 1) * nx; };
  // X0/X1[k] = x row j-1+k (columns c0, c0+1), B0/B1[k] = b row j+k
  T X0[3 + kAhead], X1[3 + kAhead], B0[1 + kAhead], B1[1 + kAhead];
#pragma unroll
  for (int k = 0; k < 2 + kAhead; ++k) ld_cols<T>(x, row(it.b - 1 + k), c0, nx, X0[k], X1[k]);
#pragma unroll
  for (int k = 0; k < kAhead; ++k) {
    if constexpr (std::is_same<SB, void>::value) { B0[k] = T(0); B1[k] = T(0); }
    else ld_cols<T>(b, row(it.b + k), c0, nx, B0[k], B1[k]);
  }
  for (int j = it.b; j < it.c; ++j) {
    ld_cols<T>(x, row(j + 1 + kAhead), c0, nx, X0[2 + kAhead], X1[2 + kAhead]);
    if constexpr (std::is_same<SB, void>::value) { B0[kAhead] = T(0); B1[kAhead] = T(0); }
    else ld_cols<T>(b, row(j + kAhead), c0, nx, B0[kAhead], B1[kAhead]);
    const T w0 = from_west(X1[1]), e1 = from_east(X0[1]);
    const long long id0 = (long long)j * nx + c0;
    if (out0) f(c0, j, id0, X0[1], w0, X1[1], X0[0], X0[2], B0[0]);
    if (out1) f(c0 + 1, j, id0 + 1, X1[1], X0[1], e1, X1[0], X1[2], B1[0]);
#pragma unroll
    for (int k = 0; k < 2 + kAhead; ++k) { X0[k] = X0[k + 1]; X1[k] = X1[k + 1]; }
#pragma unroll
    for (int k = 0; k < kAhead; ++k) { B0[k] = B0[k + 1]; B1[k] = B1[k + 1]; }
  }
}

// ------------------------------------------------------------------------------- smoother
// One red-black sweep of item (p, strip a, rows [b, c)), both colours, out of place: first
// colour on row p, second colour on row p-1 from the first colour's values.  Lane L holds
// columns c0 = a*124-2+2L (even) and c0+1, which have opposite colours, so on every row exactly
// one of the two updates and every lane does useful work: the element of colour c on row p is
// (c + p) & 1, the same for all lanes.  First colour: every element but the outermost two
// (their neighbours are outside the wavefront); second colour and output: lanes 1..62.
template <Prec P, bool ZERO>
__global__ void __launch_bounds__(kLanes * kWavesPB)
kb_sweep(const BEnt* __restrict__ tab, const BItem* __restrict__ items, int n, int c1, int dir) {
  using T = typename PT<P>::T;
  using S = typename PT<P>::S;
  const int w = wave_item();
  if (w >= n) return;
  const BItem it = items[w];
  const BEnt& e = tab[it.p];
  if (done(e.sc)) return;
  const DevOp& op = e.op;
  const S* __restrict__ b = static_cast<const S*>(e.b);
  const S* __restrict__ xin = static_cast<const S*>(dir ? e.alt : e.u);
  S* __restrict__ xout = static_cast<S*>(dir ? e.u : e.alt);
  const int lane = threadIdx.x & (kLanes - 1);
  const int c0 = it.a * kTileX - 2 + 2 * lane;
  const int j0 = it.b, j1 = it.c;
  const int ny = op.ny, nx = op.nx;
  const int c2 = 1 - c1;
  const bool int0 = c0 >= 1 && c0 <= nx - 2, int1 = c0 + 1 >= 1 && c0 + 1 <= nx - 2;
  const bool r1_0 = lane >= 1 && int0, r1_1 = lane <= kLanes - 2 && int1;
  const bool mid = lane >= 1 && lane < kLanes - 1;
  const bool r2_0 = mid && int0, r2_1 = mid && int1;
  const bool out0 = mid && c0 < nx, out1 = mid && c0 + 1 < nx;
  auto row = [&](int j) -> long long { return (long long)min(max(j, 0), ny - 1) * nx; };
  auto ld_u = [&](int j, T& a, T& bb) {
    if constexpr (ZERO) { a = T(0); bb = T(0); }
    else ld_cols<T>(xin, row(j), c0, nx, a, bb);
  };
  const W7<T> w0 = const_w<T>(op);
  // row queues: U*[k] = x row p-2+k, B*[k] = b row p-1+k (k < 4 / 3 in use); the rest are
  // kAhead rows in flight
  T U0[4 + kAhead], U1[4 + kAhead], B0[3 + kAhead], B1[3 + kAhead];
  U0[0] = U0[1] = U1[0] = U1[1] = T(0);
#pragma unroll
  for (int k = 2; k < 4 + kAhead; ++k) ld_u(j0 - 4 + k, U0[k], U1[k]);
  B0[0] = B0[1] = B1[0] = B1[1] = T(0);
#pragma unroll
  for (int k = 2; k < 3 + kAhead; ++k) ld_cols<T>(b, row(j0 - 3 + k), c0, nx, B0[k], B1[k]);
  // GS update of column i on row j from its four neighbours
  auto upd = [&](int i, int j, T bj, T ue, T uw, T un, T us) -> T {
    W7<T> wt = w0;
    fold_center<2, T>(op, i, j, 0, wt);
    return rounded<S, T>(gs_value<2, T>(wt, bj, ue, uw, un, us, T(0), T(0)));
  };
  for (int p = j0 - 1; p <= j1; ++p) {
#pragma unroll
    for (int k = 0; k < 3 + kAhead; ++k) { U0[k] = U0[k + 1]; U1[k] = U1[k + 1]; }
    ld_u(p + 1 + kAhead, U0[3 + kAhead], U1[3 + kAhead]);
#pragma unroll
    for (int k = 0; k < 2 + kAhead; ++k) { B0[k] = B0[k + 1]; B1[k] = B1[k + 1]; }
    ld_cols<T>(b, row(p + 1 + kAhead), c0, nx, B0[2 + kAhead], B1[2 + kAhead]);
    if (p >= 1 && p <= ny - 2) {
      if (((c1 + p) & 1) == 0) {   // column c0 has colour c1 on row p
        const T uw = from_west(U1[2]);
        if (r1_0) U0[2] = upd(c0, p, B0[1], U1[2], uw, U0[3], U0[1]);
      } else {
        const T ue = from_east(U0[2]);
        if (r1_1) U1[2] = upd(c0 + 1, p, B1[1], ue, U0[2], U1[3], U1[1]);
      }
    }
    const int q = p - 1;
    if (q >= j0 && q < j1) {
      if (q >= 1 && q <= ny - 2) {
        if (((c2 + q) & 1) == 0) {
          const T uw = from_west(U1[1]);
          if (r2_0) U0[1] = upd(c0, q, B0[0], U1[1], uw, U0[2], U0[0]);
        } else {
          const T ue = from_east(U0[1]);
          if (r2_1) U1[1] = upd(c0 + 1, q, B1[0], ue, U0[1], U1[2], U1[0]);
        }
      }
      const long long id0 = (long long)q * nx + c0;
      if (out0 && out1) {
        st2(xout, id0, store_as<S>(static_cast<double>(U0[1])), store_as<S>(static_cast<double>(U1[1])));
      } else {
        if (out0) stS(xout, id0, store_as<S>(static_cast<double>(U0[1])));
        if (out1) stS(xout, id0 + 1, store_as<S>(static_cast<double>(U1[1])));
      }
    }
  }
}

// ------------------------------------------------------------------------------- residual
// r = b - A u over strip item (p, strip, rows) of problem p (boundary rows b - u).
template <Prec P>
__global__ void __launch_bounds__(kLanes * kWavesPB)
kb_residual(const BEnt* __restrict__ tab, const BItem* __restrict__ items, int n) {
  using T = typename PT<P>::T;
  using S = typename PT<P>::S;
  const int w = wave_item();
  if (w >= n) return;
  const BItem it = items[w];
  const BEnt& e = tab[it.p];
  if (done(e.sc)) return;
  const DevOp& op = e.op;
  S* __restrict__ r = static_cast<S*>(e.r);
  const W7<T> w0 = const_w<T>(op);
  strip_march<T>(static_cast<const S*>(e.u), static_cast<const S*>(e.b), op.nx, op.ny, it,
                 [&](int i, int j, long long id, T uc, T uw, T ue, T us, T un, T bj) {
    T out;
    if (i == 0 || j == 0 || i == op.nx - 1 || j == op.ny - 1) {
      out = bj - uc;
    } else {
      W7<T> wt = w0;
      fold_center<2, T>(op, i, j, 0, wt);
      const T ax = wt.C * uc + wt.E * ue + wt.W * uw + wt.N * un + wt.S * us;
      out = bj - ax;
    }
    stS(r, id, store_as<S>(static_cast<double>(out)));
  });
}

// ------------------------------------------------------------------------------- transfers
__device__ __forceinline__ int fold_list(int c, int n, double th_lo, double th_hi, int* idx, double* w) {
  idx[0] = c; w[0] = 1.0;
  int m = 1;
  if (c == 1) { idx[m] = 0; w[m] = th_lo; ++m; }
  if (c == n - 2) { idx[m] = n - 1; w[m] = th_hi; ++m; }
  return m;
}

template <typename SF>
__device__ __forceinline__ double fw2(const DevOp& f, const SF* rf, int fi, int fj) {
  const double w1[3] = {0.25, 0.5, 0.25};
  double acc = 0.0;
  for (int dj = -1; dj <= 1; ++dj) {
    const int j = fj + dj;
    if (j < 0 || j >= f.ny) continue;
    const SF* row = rf + (long long)j * f.nx;
    double a = 0.5 * static_cast<double>(ldS(row, fi));
    if (fi > 0) a += 0.25 * static_cast<double>(ldS(row, fi - 1));
    if (fi < f.nx - 1) a += 0.25 * static_cast<double>(ldS(row, fi + 1));
    acc += w1[dj + 1] * a;
  }
  return acc;
}

// Coarse rows [a, b) of problem p: FW of the fine residual with the folded closure transpose
// (coarse boundary nodes zero), times the problem's normalisation scale when scale0.
template <Prec PF, Prec PC>
__global__ void __launch_bounds__(kLanes * kWavesPB)
kb_restrict(const BEnt* __restrict__ ftab, const BEnt* __restrict__ ctab,
            const BItem* __restrict__ items, int n, int scale0) {
  using SF = typename PT<PF>::S;
  using SC = typename PT<PC>::S;
  const int w = wave_item();
  if (w >= n) return;
  const BItem it = items[w];
  const BEnt& fe = ftab[it.p];
  const BEnt& ce = ctab[it.p];
  if (done(fe.sc)) return;
  const DevOp& f = fe.op;
  const DevOp& c = ce.op;
  const SF* rf = static_cast<const SF*>(fe.r);
  SC* bc = static_cast<SC*>(ce.b);
  const double s = scale0 ? fe.sc->scale : 1.0;
  const int lane = threadIdx.x & (kLanes - 1);
  for (int jc = it.a; jc < it.b; ++jc) {
    SC* row = bc + (long long)jc * c.nx;
    if (jc == 0 || jc == c.ny - 1) {
      for (int ic = lane; ic < c.nx; ic += kLanes) stS(row, ic, store_as<SC>(0.0));
      continue;
    }
    int yi[3];
    double yw[3];
    const int my = fold_list(jc, c.ny, c.theta[2], c.theta[3], yi, yw);
#pragma unroll 2
    for (int ic = lane; ic < c.nx; ic += kLanes) {
      if (ic == 0 || ic == c.nx - 1) { stS(row, ic, store_as<SC>(0.0)); continue; }
      int xi[3];
      double xw[3];
      const int mx = fold_list(ic, c.nx, c.theta[0], c.theta[1], xi, xw);
      double acc = 0.0;
      for (int bb = 0; bb < my; ++bb)
        for (int q = 0; q < mx; ++q) {
          const double wt = yw[bb] * xw[q];
          if (wt != 0.0) acc += wt * fw2(f, rf, 2 * xi[q], 2 * yi[bb]);
        }
      stS(row, ic, store_as<SC>(acc * s));
    }
  }
}

// Fine rows [a, b) of problem p += pfac * (inv_scale when scale0) * bilinear interpolation of
// the closure-extended coarse correction.
template <Prec PC, Prec PF>
__global__ void __launch_bounds__(kLanes * kWavesPB)
kb_prolong(const BEnt* __restrict__ ctab, const BEnt* __restrict__ ftab,
           const BItem* __restrict__ items, int n, int scale0) {
  using SC = typename PT<PC>::S;
  using SF = typename PT<PF>::S;
  const int w = wave_item();
  if (w >= n) return;
  const BItem it = items[w];
  const BEnt& ce = ctab[it.p];
  const BEnt& fe = ftab[it.p];
  if (done(fe.sc)) return;
  const DevOp& c = ce.op;
  const DevOp& f = fe.op;
  const SC* uc = static_cast<const SC*>(ce.u);
  SF* uf = static_cast<SF*>(fe.u);
  const double s = fe.pfac * (scale0 ? fe.sc->inv_scale : 1.0);
  const int lane = threadIdx.x & (kLanes - 1);
  for (int j = max(it.a, 1); j < min(it.b, f.ny - 1); ++j) {
    const int jc = j >> 1, oj = j & 1;
    int y0 = jc, y1 = jc + 1;
    double wy0 = 1.0, wy1 = oj ? 1.0 : 0.0;
    if (y0 == 0) { wy0 = c.theta[2]; y0 = 1; }
    if (y1 == c.ny - 1) { wy1 *= c.theta[3]; y1 = c.ny - 2; }
    const double hy = oj ? 0.5 : 1.0;
    const SC* r0 = uc + (long long)y0 * c.nx;
    const SC* r1 = uc + (long long)y1 * c.nx;
#pragma unroll 4
    for (int i = 1 + lane; i < f.nx - 1; i += kLanes) {
      const int ic = i >> 1, oi = i & 1;
      int x0 = ic, x1 = ic + 1;
      double wx0 = 1.0, wx1 = oi ? 1.0 : 0.0;
      if (x0 == 0) { wx0 = c.theta[0]; x0 = 1; }
      if (x1 == c.nx - 1) { wx1 *= c.theta[1]; x1 = c.nx - 2; }
      const double hx = oi ? 0.5 : 1.0;
      double a = 0.0;
      if (wy0 != 0.0) {
        double v = wx0 != 0.0 ? wx0 * static_cast<double>(ldS(r0, x0)) : 0.0;
        if (wx1 != 0.0) v += wx1 * static_cast<double>(ldS(r0, x1));
        a += wy0 * (v * hx);
      }
      if (wy1 != 0.0) {
        double v = wx0 != 0.0 ? wx0 * static_cast<double>(ldS(r1, x0)) : 0.0;
        if (wx1 != 0.0) v += wx1 * static_cast<double>(ldS(r1, x1));
        a += wy1 * (v * hx);
      }
      const long long id = (long long)j * f.nx + i;
      stS(uf, id, store_as<SF>(static_cast<double>(ldS(uf, id)) + s * (a * hy)));
    }
  }
}

// Coarsest level: item (p, row r): x[map[r]] = sum_q Ainv[r][q] b[map[q]].
template <Prec P>
__global__ void __launch_bounds__(kLanes * kWavesPB)
kb_dense(const BEnt* __restrict__ tab, const BDense* __restrict__ dense,
         const BItem* __restrict__ items, int n) {
  using S = typename PT<P>::S;
  const int w = wave_item();
  if (w >= n) return;
  const BItem it = items[w];
  const BEnt& e = tab[it.p];
  if (done(e.sc)) return;
  const BDense d = dense[it.p];
  const S* b = static_cast<const S*>(e.b);
  S* x = static_cast<S*>(e.u);
  const int lane = threadIdx.x & (kLanes - 1);
  const int row = it.a;
  double acc = 0.0;
  for (int q = lane; q < d.m; q += kLanes)
    acc += ldS(d.ainv, (long long)row * d.m + q) * static_cast<double>(ldS(b, ldS(d.map, q)));
  acc = wave_sum(acc);
  if (lane == 0) stS(x, ldS(d.map, row), store_as<S>(acc));
}

// ------------------------------------------------------------------------------- level-0 Krylov
// Item (p, e0, e1) over the problem's level-0 vectors, one workgroup per item.
template <Prec P>
__global__ void __launch_bounds__(256)
kb_vec(const BVec* __restrict__ vecs, const BItem* __restrict__ items, int n, int op,
       int inner, double* partials, double* maxpart) {
  using T = typename PT<P>::T;
  using S = typename PT<P>::S;
  __shared__ double lds[256 / kWave];
  const int w = blockIdx.x;
  if (w >= n) return;
  const BItem it = items[w];
  const BVec& v = vecs[it.p];
  const KScalars* sc = inner ? v.si : v.so;
  double acc = 0.0, mx = 0.0;
  if (!done(sc)) {
    S* x = static_cast<S*>(v.x);
    S* r = static_cast<S*>(v.r);
    S* pp = static_cast<S*>(v.p);
    const S* q = static_cast<const S*>(v.q);
    const S* z = static_cast<const S*>(v.z);
    const long long i0 = it.a + threadIdx.x, i1 = it.b;
    const int st = blockDim.x;
    switch (op) {
      case V_CG_UPDATE: {
        const double alpha = ldS(&sc->alpha, 0);
        const T a = T(alpha);
        if (inner) {
          // the inner correction goes straight into the FP64 solution: x64 += |r64| alpha p
          const double ax = ldS(&v.so->snorm, 0) * alpha;
#pragma unroll 4
          for (long long i = i0; i < i1; i += st) {
            stS(v.x64, i, ldS(v.x64, i) + ax * static_cast<double>(ldS(pp, i)));
            const T ri = T(ldS(r, i)) - a * T(ldS(q, i));
            stS(r, i, store_as<S>(static_cast<double>(ri)));
            acc += static_cast<double>(ri) * static_cast<double>(ri);
            mx = nmax(mx, fabs(static_cast<double>(ri)));
          }
        } else {
#pragma unroll 4
          for (long long i = i0; i < i1; i += st) {
            stS(x, i, store_as<S>(static_cast<double>(T(ldS(x, i)) + a * T(ldS(pp, i)))));
            const T ri = T(ldS(r, i)) - a * T(ldS(q, i));
            stS(r, i, store_as<S>(static_cast<double>(ri)));
            acc += static_cast<double>(ri) * static_cast<double>(ri);
            mx = nmax(mx, fabs(static_cast<double>(ri)));
          }
        }
        break;
      }
      case V_DOT_RZ:
#pragma unroll 4
        for (long long i = i0; i < i1; i += st)
          acc += static_cast<double>(ldS(r, i)) * static_cast<double>(ldS(z, i));
        break;
      case V_XPAY: {
        const T bt = T(ldS(&sc->beta, 0));
#pragma unroll 4
        for (long long i = i0; i < i1; i += st)
          stS(pp, i, store_as<S>(static_cast<double>(T(ldS(z, i)) + bt * T(ldS(pp, i)))));
        break;
      }
      case V_COPY_ZP:
#pragma unroll 4
        for (long long i = i0; i < i1; i += st) stS(pp, i, store_as<S>(static_cast<double>(ldS(z, i))));
        break;
      case V_RR:
#pragma unroll 4
        for (long long i = i0; i < i1; i += st) {
          const double ri = static_cast<double>(ldS(r, i));
          acc += ri * ri;
          mx = nmax(mx, fabs(ri));
        }
        break;
      case V_ZERO_X:
        for (long long i = i0; i < i1; i += st) stS(x, i, store_as<S>(0.0));
        break;
      case V_TO32: {   // inner rhs = r64 / snorm (the fixed inner scale)
        const double nr = ldS(&v.so->snorm, 0);
        const double sinv = nr > 0.0 ? 1.0 / nr : 1.0;
#pragma unroll 4
        for (long long i = i0; i < i1; i += st) stS(r, i, store_as<S>(ldS(v.r64, i) * sinv));
        break;
      }
      case V_COPY_X0:
#pragma unroll 4
        for (long long i = i0; i < i1; i += st) stS(v.x64, i, ldS(v.x0, i));
        break;
      default:
        break;
    }
  }
  if (partials) {
    const double bs = block_sum(acc, lds);
    if (threadIdx.x == 0) partials[w] = bs;
  }
  if (maxpart) {
    const double m = block_max(mx, lds);
    if (threadIdx.x == 0) maxpart[w] = m;
  }
}

// Level-0 rows [a, b) of problem p.  Inner (OUTER false): q = A p with the level-0 operator of
// the Krylov precision (tab0) and partials of p.q.  Outer: r64 = b64 - A64 x64 with partials
// of r.r (the FP64 refinement residual).
template <Prec P, bool OUTER, bool ALL = false>
__global__ void __launch_bounds__(kLanes * kWavesPB)
kb_apply_dot(const BVec* __restrict__ vecs, const BEnt* __restrict__ tab0,
             const BItem* __restrict__ items, int n, double* partials) {
  using T = typename PT<P>::T;
  using S = typename PT<P>::S;
  const int w = wave_item();
  if (w >= n) return;
  const BItem it = items[w];
  const BVec& v = vecs[it.p];
  const int lane = threadIdx.x & (kLanes - 1);
  double acc = 0.0;
  const KScalars* sc = OUTER ? v.so : tab0[it.p].sc;
  if (ALL || !done(sc)) {
    const DevOp& op = OUTER ? v.op64 : tab0[it.p].op;
    const S* x = OUTER ? reinterpret_cast<const S*>(v.x64) : static_cast<const S*>(v.p);
    S* y = OUTER ? reinterpret_cast<S*>(v.r64) : static_cast<S*>(v.q);
    const W7<T> w0 = const_w<T>(op);
    const T e = T(op.e);
    auto body = [&](int i, int j, long long id, T xc, T xw, T xe, T xs, T xn, T bj) {
      T ax;
      if (i == 0 || j == 0 || i == op.nx - 1 || j == op.ny - 1) {
        ax = xc;
      } else {
        ax = apply_diff<2, T>(op, w0, e, i, j, 0, xc, xw, xe, xs, xn, T(0), T(0));
      }
      if constexpr (OUTER) {
        const T rv = bj - ax;
        stS(y, id, store_as<S>(static_cast<double>(rv)));
        acc += static_cast<double>(rv) * static_cast<double>(rv);
      } else {
        stS(y, id, store_as<S>(static_cast<double>(ax)));
        acc += static_cast<double>(xc) * static_cast<double>(ax);
      }
    };
    if constexpr (OUTER) strip_march<T>(x, reinterpret_cast<const S*>(v.b64), op.nx, op.ny, it, body);
    else strip_march<T>(x, static_cast<const void*>(nullptr), op.nx, op.ny, it, body);
  }
  acc = wave_sum(acc);
  if (lane == 0) partials[w] = acc;
}

// Residual replacement (mixed precision): for problems whose inner scalars flag it, the FP32
// Krylov residual is replaced by the true residual (b64 - A64 x64) / snorm, computed in FP64
// from the FP64 solution; partials are |b - A x|^2 in FP64 (zero for other problems).
__global__ void __launch_bounds__(kLanes * kWavesPB)
kb_replace(const BVec* __restrict__ vecs, const BItem* __restrict__ items, int n,
           double* partials) {
  const int w = wave_item();
  if (w >= n) return;
  const BItem it = items[w];
  const BVec& v = vecs[it.p];
  const int lane = threadIdx.x & (kLanes - 1);
  double acc = 0.0;
  const KScalars* si = v.si;
  if (!done(si) && si->flag_s) {
    const DevOp& op = v.op64;
    float* r = static_cast<float*>(v.r);
    const double inv = 1.0 / v.so->snorm;
    const W7<double> w0 = const_w<double>(op);
    const double e = op.e;
    strip_march<double>(v.x64, static_cast<const double*>(v.b64), op.nx, op.ny, it,
                        [&](int i, int j, long long id, double xc, double xw, double xe, double xs,
                            double xn, double bj) {
      double ax;
      if (i == 0 || j == 0 || i == op.nx - 1 || j == op.ny - 1) {
        ax = xc;
      } else {
        ax = apply_diff<2, double>(op, w0, e, i, j, 0, xc, xw, xe, xs, xn, 0.0, 0.0);
      }
      const double rv = bj - ax;
      stS(r, id, static_cast<float>(rv * inv));
      acc += rv * rv;
    });
  }
  acc = wave_sum(acc);
  if (lane == 0) partials[w] = acc;
}

// Segmented finalize: workgroup q sums problem q's partial range in a fixed order.
__global__ void __launch_bounds__(256)
kb_finalize(const BVec* __restrict__ vecs, const int* __restrict__ first,
            const int* __restrict__ cnt, int nprob, const double* __restrict__ partials,
            const double* __restrict__ maxpart, int op, int inner, double tol, int tol_mode) {
  __shared__ double lds[256 / kWave];
  const int q = blockIdx.x;
  if (q >= nprob) return;
  const BVec& v = vecs[q];
  KScalars* sc = inner ? v.si : v.so;
  if (done(sc)) return;
  double acc = 0.0, mx = 0.0;
  for (int t = threadIdx.x; t < cnt[q]; t += blockDim.x) {
    acc += partials[first[q] + t];
    if (maxpart) mx = nmax(mx, maxpart[first[q] + t]);
  }
  const double s = block_sum(acc, lds);
  const double m = maxpart ? block_max(mx, lds) : 0.0;
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
    case FIN_RU_CHECK: {   // inner scalars; tol = the replacement threshold delta
      sc->rr = s;
      sc->metric = sqrt(s);
      sc->iter += 1;
      if (nonfinite(s)) { sc->breakdown = 1; v.so->breakdown = 1; break; }
      sc->rho = fmax(sc->rho, sc->metric);   // largest residual since the last replacement
      v.so->pad0 = sc->iter;
      sc->flag_s = (sc->metric < tol * sc->rho || sc->metric * v.so->snorm < v.so->target) ? 1 : 0;
      break;
    }
    case FIN_RU_APPLY: {   // inner scalars; the outer ones take the true residual
      if (!sc->flag_s) break;
      KScalars* so = v.so;
      so->metric = sqrt(s);
      so->steps += 1;
      if (nonfinite(s)) { so->breakdown = 1; sc->breakdown = 1; break; }
      sc->metric = so->metric / so->snorm;
      sc->rho = sc->metric;
      sc->flag_s = 0;
      if (so->metric < so->target) { so->done = 1; sc->done = 1; }
      break;
    }
    default:
      break;
  }
  if (maxpart && !(sc->done | sc->breakdown)) {
    sc->inv_scale = m > 0.0 ? m : 1.0;
    sc->scale = m > 0.0 ? 1.0 / m : 1.0;
  }
}

// Inner scalars of every unfinished problem reset for a new refinement; the outer metric is
// kept as the correction's scale (snorm).
__global__ void kb_reset_inner(const BVec* __restrict__ vecs, int nprob, int outer_iter_inc) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nprob) return;
  const BVec& v = vecs[q];
  v.so->pad0 += v.si->iter;          // inner iterations of the finished refinement
  KScalars k{};
  k.eps = 1e-300;
  k.scale = 1.0;
  k.inv_scale = 1.0;
  k.done = (v.so->done | v.so->breakdown);
  *v.si = k;
  if (!k.done) {
    v.so->snorm = v.so->metric;
    v.so->steps += outer_iter_inc;   // refinements
  }
}

// Fresh scalars for a solve (outer and inner): breakdown threshold, unit scales, counters 0.
// Start of a mixed solve with residual replacement: fresh inner scalars (done when the outer
// ones are), the inner scale fixed at the initial residual norm.
__global__ void kb_ru_start(const BVec* __restrict__ vecs, int nprob) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nprob) return;
  const BVec& v = vecs[q];
  KScalars k{};
  k.eps = 1e-300;
  k.scale = 1.0;
  k.inv_scale = 1.0;
  k.done = (v.so->done | v.so->breakdown);
  *v.si = k;
  v.so->snorm = v.so->metric > 0.0 ? v.so->metric : 1.0;
  v.so->steps = 0;
  v.so->pad0 = 0;
}

__global__ void kb_init_scalars(const BVec* __restrict__ vecs, int nprob, double eps) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nprob) return;
  KScalars k{};
  k.eps = eps;
  k.scale = 1.0;
  k.inv_scale = 1.0;
  *vecs[q].so = k;
  *vecs[q].si = k;
}

__global__ void kb_active(const BVec* __restrict__ vecs, int nprob, int inner, int* active) {
  __shared__ double lds[1024 / kWave];
  double a = 0.0;
  for (int q = threadIdx.x; q < nprob; q += blockDim.x) {
    const KScalars* sc = inner ? vecs[q].si : vecs[q].so;
    // a BiCGStab problem broken down with a restart pending is still active
    a += (sc->done | (sc->breakdown & !sc->cycle_done)) ? 0.0 : 1.0;
  }
  const double s = block_sum(a, lds);
  if (threadIdx.x == 0) *active = static_cast<int>(s);
}

// A problem stopped from the host (a deadline): both scalar blocks broken down, no restart.
__global__ void kb_cancel(const BVec* __restrict__ vecs, int p) {
  if (threadIdx.x != 0) return;
  KScalars* scs[2] = {vecs[p].so, vecs[p].si};
  for (KScalars* sc : scs) {
    if (!sc) continue;
    sc->breakdown = 1;
    sc->cycle_done = 0;
  }
}

// ------------------------------------------------------------------------ batched BiCGStab
// Vector kernels over row-block items (p, rows [a, b)): one wavefront walks the rows, lanes the
// columns, so the Jacobi diagonal of every node follows from (i, j) without a division.  The
// arithmetic per node is the single-problem path's (k_krylov.hip): FP64 combinations of the
// stored values, rounded once to the vector precision, division by the folded diagonal.
__device__ __forceinline__ double kk_diag(const DevOp& op, int i, int j) {
  W7<double> w = const_w<double>(op);
  fold_center<2, double>(op, i, j, 0, w);
  return w.C;
}

template <typename S, int OP>
__global__ void __launch_bounds__(kLanes * kWavesPB)
kk_vec(const BKVec* __restrict__ vecs, const BItem* __restrict__ items, int n, double* partials) {
  const int w = wave_item();
  if (w >= n) return;
  const BItem it = items[w];
  const BKVec& v = vecs[it.p];
  const int lane = threadIdx.x & (kLanes - 1);
  double acc = 0.0;
  const KScalars* sc = v.sc;
  // a restart pending after a breakdown copies the residual into rh although the problem is
  // marked broken down until the scalars are reset
  // KK_X0 resets only the problems a mixed batch's completion restarts from x0 (steps == -1)
  const bool run = OP == KK_RESTART ? ldS(&sc->cycle_done, 0) != 0 && !ldS(&sc->done, 0)
                 : (OP == KK_X0 ? !done(sc) && ldS(&sc->steps, 0) == -1 : !done(sc));
  if (run) {
    const int nx = v.nx, ny = v.ny;
    S* r = static_cast<S*>(v.r);
    S* rh = static_cast<S*>(v.rh);
    S* p = static_cast<S*>(v.p);
    S* vv = static_cast<S*>(v.v);
    S* sv = static_cast<S*>(v.s);
    S* ph = static_cast<S*>(v.ph);
    S* sh = static_cast<S*>(v.sh);
    const double alpha = ldS(&sc->alpha, 0), beta = ldS(&sc->beta, 0), omega = ldS(&sc->omega, 0);
    const bool first = ldS(&sc->pad0, 0) == 0;   // first iteration since the (re)start
    const bool fs = ldS(&sc->flag_s, 0) != 0;
    const bool jac = v.jacobi != 0;
    for (int j = it.a; j < it.b && (OP != KK_XS || fs); ++j) {
      const long long row = (long long)j * nx;
      const bool bj = j == 0 || j == ny - 1;
      for (int i = lane; i < nx; i += kLanes) {
        const long long id = row + i;
        const bool bnd = bj || i == 0 || i == nx - 1;
        if constexpr (OP == KK_RH || OP == KK_RESTART) {
          stS(rh, id, ldS(r, id));
        } else if constexpr (OP == KK_X0) {
          stS(v.x, id, ldS(v.x0, id));
        } else if constexpr (OP == KK_RHO) {
          acc += double(ldS(rh, id)) * double(ldS(r, id));
        } else if constexpr (OP == KK_P) {
          const S pn = static_cast<S>(first ? double(ldS(r, id))
                                            : double(ldS(r, id)) + beta * (double(ldS(p, id)) - omega * double(ldS(vv, id))));
          stS(p, id, pn);
          if (jac) stS(ph, id, bnd ? pn : static_cast<S>(double(pn) / kk_diag(v.op64, i, j)));
        } else if constexpr (OP == KK_S) {
          const S si = static_cast<S>(double(ldS(r, id)) - alpha * double(ldS(vv, id)));
          stS(sv, id, si);
          acc += double(si) * double(si);
          if (jac) stS(sh, id, bnd ? si : static_cast<S>(double(si) / kk_diag(v.op64, i, j)));
        } else if constexpr (OP == KK_XS) {
          stS(v.x, id, ldS(v.x, id) + alpha * double(ldS(ph, id)));
        } else {
          stS(v.x, id, ldS(v.x, id) + (alpha * double(ldS(ph, id)) + omega * double(ldS(sh, id))));
        }
      }
    }
  }
  if constexpr (OP == KK_RHO || OP == KK_S) {
    acc = wave_sum(acc);
    if (lane == 0) partials[w] = acc;
  }
}

// Applies over level-0 strip items.  MODE 0: v = A ph (+ rh.v); 1: t = A sh (+ t.t, t.s);
// 2: r = b - A x in FP64 (+ |r|^2).  FP64 vectors use the sum form with the folded centre
// (k_stencil's FP64 rows); FP32 vectors the difference form (its FP32 rows).
template <typename S, int MODE>
__global__ void __launch_bounds__(kLanes * kWavesPB)
kk_apply(const BKVec* __restrict__ vecs, const BItem* __restrict__ items, int n, double* p1,
         double* p2) {
  using T = std::conditional_t<MODE == 2, double, S>;
  const int w = wave_item();
  if (w >= n) return;
  const BItem it = items[w];
  const BKVec& v = vecs[it.p];
  const int lane = threadIdx.x & (kLanes - 1);
  double a1 = 0.0, a2 = 0.0;
  const KScalars* sc = v.sc;
  bool run = !done(sc);
  if (MODE == 1 && run) run = ldS(&sc->flag_s, 0) == 0;
  if (run) {
    const DevOp& op = MODE == 2 ? v.op64 : v.op;
    const W7<T> w0 = const_w<T>(op);
    const T e = T(op.e);
    auto ax_of = [&](int i, int j, T xc, T xw, T xe, T xs, T xn) -> T {
      if (i == 0 || j == 0 || i == op.nx - 1 || j == op.ny - 1) return xc;
      if constexpr (std::is_same<T, double>::value) {
        W7<T> wf = w0;
        fold_center<2, T>(op, i, j, 0, wf);
        return wf.C * xc + wf.E * xe + wf.W * xw + wf.N * xn + wf.S * xs;
      } else {
        return apply_diff<2, T>(op, w0, e, i, j, 0, xc, xw, xe, xs, xn, T(0), T(0));
      }
    };
    if constexpr (MODE == 0) {
      S* out = static_cast<S*>(v.v);
      const S* rh = static_cast<const S*>(v.rh);
      strip_march<T>(static_cast<const S*>(v.ph), static_cast<const void*>(nullptr), op.nx, op.ny, it,
                     [&](int i, int j, long long id, T xc, T xw, T xe, T xs, T xn, T) {
        const S o = static_cast<S>(ax_of(i, j, xc, xw, xe, xs, xn));
        stS(out, id, o);
        a1 += double(ldS(rh, id)) * double(o);
      });
    } else if constexpr (MODE == 1) {
      S* out = static_cast<S*>(v.t);
      const S* sv = static_cast<const S*>(v.s);
      strip_march<T>(static_cast<const S*>(v.sh), static_cast<const void*>(nullptr), op.nx, op.ny, it,
                     [&](int i, int j, long long id, T xc, T xw, T xe, T xs, T xn, T) {
        const S o = static_cast<S>(ax_of(i, j, xc, xw, xe, xs, xn));
        stS(out, id, o);
        a1 += double(o) * double(o);
        a2 += double(o) * double(ldS(sv, id));
      });
    } else {
      S* out = static_cast<S*>(v.r);
      strip_march<double>(v.x, static_cast<const double*>(v.b), op.nx, op.ny, it,
                          [&](int i, int j, long long id, double xc, double xw, double xe, double xs,
                              double xn, double bj) {
        const double rv = bj - ax_of(i, j, xc, xw, xe, xs, xn);
        stS(out, id, static_cast<S>(rv));
        a1 += rv * rv;
      });
    }
  }
  a1 = wave_sum(a1);
  if (lane == 0) p1[w] = a1;
  if constexpr (MODE == 1) {
    a2 = wave_sum(a2);
    if (lane == 0) p2[w] = a2;
  }
}

// A vanishing rho, r^.v, t.t or omega stops the single-problem BiCGStab (reference semantics).
// The batch instead restarts the problem from its true residual (r^ = r, p = r), at most
// kKkRestarts times: the rest of the iteration skips it (broken down), and kk_restart clears
// the flag before the next one.  The residual is b - A x of the current x at every such point.
// A non-finite dot product before the solution update (an overflowing FP32 recurrence) is
// restarted the same way; a non-finite true residual ends the problem.
constexpr int kKkRestarts = 8;
__device__ __forceinline__ void kk_breakdown(KScalars* sc) {
  sc->breakdown = 1;
  if (sc->pad1 < kKkRestarts) sc->cycle_done = 1;
}

// Restart of every problem with a pending restart (after its rh = r copy): scalars of a fresh
// start, one more restart counted.
__global__ void kk_restart(const BKVec* __restrict__ vecs, int nprob) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nprob) return;
  KScalars* sc = vecs[q].sc;
  if (!sc->cycle_done || sc->done) return;
  sc->cycle_done = 0;
  sc->breakdown = 0;
  sc->pad0 = 0;
  sc->pad1 += 1;
}

// Number of problems that did not converge (not done, or broken down), one block.
__global__ void __launch_bounds__(1024) kk_failed(const BKVec* __restrict__ vecs, int nprob,
                                                 int* out) {
  __shared__ double lds[1024 / kWave];
  double a = 0.0;
  for (int q = threadIdx.x; q < nprob; q += blockDim.x) {
    const KScalars* sc = vecs[q].sc;
    a += (sc->done && !sc->breakdown) ? 0.0 : 1.0;
  }
  const double t = block_sum(a, lds);
  if (threadIdx.x == 0) *out = static_cast<int>(t);
}

__global__ void kk_reset_failed(const BKVec* __restrict__ vecs, int nprob, double eps) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nprob) return;
  KScalars* sc = vecs[q].sc;
  if (sc->done && !sc->breakdown) return;
  // continue from the current solution when the first pass cut the true residual by 1e3
  // (steps = -2), else start again from x0 (steps = -1)
  const bool keep = isfinite(sc->metric) && sc->metric < 1e-3 * sc->r0;
  KScalars k{};
  k.eps = eps;
  k.scale = 1.0;
  k.inv_scale = 1.0;
  k.steps = keep ? -2 : -1;
  k.r0 = sc->r0;           // the first pass's initial residual and target carry over
  k.target = sc->target;
  *sc = k;
}

// Segmented finalize of the batched BiCGStab: problem blockIdx.x sums its items' partials in a
// fixed order and applies the single-problem path's scalar step (k_krylov.hip k_fin_k).
__global__ void __launch_bounds__(256)
kk_finalize(const BKVec* __restrict__ vecs, const int* __restrict__ first,
            const int* __restrict__ cnt, int nprob, const double* __restrict__ p1,
            const double* __restrict__ p2, int op, double tol, int tol_mode) {
  __shared__ double lds[256 / kWave];
  const int q = blockIdx.x;
  if (q >= nprob) return;
  KScalars* sc = vecs[q].sc;
  if (sc->done | sc->breakdown) return;
  const int f = first[q], c = cnt[q];
  double a1 = 0.0, a2 = 0.0;
  for (int i = threadIdx.x; i < c; i += blockDim.x) {
    a1 += p1[f + i];
    if (p2) a2 += p2[f + i];
  }
  const double s1 = block_sum(a1, lds);
  const double s2 = p2 ? block_sum(a2, lds) : 0.0;
  if (threadIdx.x != 0) return;
  switch (op) {
    case KK_R0: {
      const double r0 = sqrt(s1);
      if (nonfinite(s1)) { sc->breakdown = 1; break; }
      sc->r0 = r0;
      sc->metric = r0;
      sc->target = tol_mode == PDEHIP_TOL_REL ? tol * r0 : tol;
      if (r0 < sc->target || r0 == 0.0) sc->done = 1;
      break;
    }
    case KK_FRHO:
      sc->rho = s1;
      if (fabs(s1) < sc->eps || nonfinite(s1)) { kk_breakdown(sc); break; }
      if (sc->pad0 > 0) sc->beta = (s1 / sc->rho_old) * (sc->alpha / sc->omega);
      break;
    case KK_FALPHA:
      if (fabs(s1) < sc->eps || nonfinite(s1)) { kk_breakdown(sc); break; }
      sc->alpha = sc->rho / s1;
      break;
    case KK_FS:
      sc->snorm = sqrt(s1);
      if (nonfinite(s1)) { kk_breakdown(sc); break; }
      if (sc->snorm < sc->target) sc->flag_s = 1;
      break;
    case KK_FOMEGA:
      if (sc->flag_s) {   // |s| met the target: x += alpha ph already done, stop here
        sc->metric = sc->snorm;
        sc->iter += 1;
        sc->done = 1;
        break;
      }
      sc->tt = s1;
      sc->ts = s2;
      if (fabs(s1) < sc->eps || nonfinite(s1) || nonfinite(s2)) { kk_breakdown(sc); break; }
      sc->omega = s2 / s1;
      break;
    case KK_FR:
      sc->metric = sqrt(s1);
      sc->iter += 1;
      sc->pad0 += 1;
      if (nonfinite(s1)) { sc->breakdown = 1; break; }
      if (sc->metric < sc->target) { sc->done = 1; break; }
      // stall window (tol > 0, the FP32 pass of a mixed batch): a true residual that has not
      // halved for `tol` iterations ends the pass for this problem (no restart); the FP64
      // completion continues from its current solution
      if (tol > 0.0) {
        if (sc->best_iter == 0 || sc->metric < 0.5 * sc->best) {
          sc->best = sc->metric;
          sc->best_iter = sc->iter;
        } else if (sc->iter - sc->best_iter >= static_cast<int>(tol)) {
          sc->stalled = 1;
          sc->breakdown = 1;
          sc->cycle_done = 0;
          break;
        }
      }
      if (fabs(sc->omega) < sc->eps) { kk_breakdown(sc); break; }
      sc->rho_old = sc->rho;
      break;
    case KK_R0K:
      // restart of the FP64 completion from the current solution: the initial residual and the
      // target of the first pass are kept
      if (nonfinite(s1)) { sc->breakdown = 1; break; }
      sc->metric = sqrt(s1);
      if (sc->metric < sc->target) sc->done = 1;
      break;
    default:
      break;
  }
}

// ------------------------------------------------------------------------ batched GMRES(m)
// The single-problem path's GMRES (runtime.cpp solve_gmres, k_krylov.hip) with every launch
// covering every problem of the batch: vector kernels over row-block items, applies over
// level-0 strips, per-problem segmented reductions, and the per-problem scalar work (Givens
// steps, back substitution) in one block per problem.  Problems leave a cycle when their
// residual estimate meets the target (cycle_done) and the batch when the true residual at the
// cycle's end does (done).

// Right Jacobi preconditioner at node (i, j): division by the folded diagonal on interior rows,
// identity on boundary rows (k_krylov.hip k_jacobi_pc).
__device__ __forceinline__ double kg_pc(const BGVec& v, int i, int j, double x) {
  if (!v.jacobi || i == 0 || j == 0 || i == v.nx - 1 || j == v.ny - 1) return x;
  return x / kk_diag(v.op, i, j);
}

constexpr int kGmPer = 8;                  // nodes per lane in a GMRES vector item
static_assert(kGmPer * kLanes == kGmresItemNodes, "GMRES vector items: 8 nodes per lane");

__device__ __forceinline__ bool kg_out_of_cycle(const KScalars* sc) {
  return done(sc) || ldS(&sc->cycle_done, 0) != 0;
}

// Vector kernels over element items (p, [a, b)) of at most kGmresItemNodes: lane L holds nodes
// a + 64 k + L (k < kGmPer) in registers, so each basis vector's loads for all of them are in
// flight together (a per-node loop over the basis would wait on one load pair at a time).
template <int OP>
__global__ void __launch_bounds__(kLanes * kWavesPB)
kg_vec(const BGVec* __restrict__ vecs, const BItem* __restrict__ items, int n, int col, int mm,
       double* partials, double* partials2) {
  const int w = wave_item();
  if (w >= n) return;
  const BItem it = items[w];
  const BGVec& v = vecs[it.p];
  const int lane = threadIdx.x & (kLanes - 1);
  const KScalars* sc = v.sc;
  const long long nn = v.n;
  long long ids[kGmPer];
  bool in[kGmPer];
#pragma unroll
  for (int k = 0; k < kGmPer; ++k) {
    ids[k] = (long long)it.a + k * kLanes + lane;
    in[k] = ids[k] < it.b;
    if (!in[k]) ids[k] = it.a;   // loads stay in the item; results are not stored
  }
  if constexpr (OP == KG_DOT) {
    // V_i . w for the group i = 8 blockIdx.y .. +7 (i <= col)
    const int i0 = static_cast<int>(blockIdx.y) * 8;
    const int cnt = min(8, col + 1 - i0);
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (!kg_out_of_cycle(sc)) {
      double wv[kGmPer];
#pragma unroll
      for (int k = 0; k < kGmPer; ++k) wv[k] = in[k] ? ldS(v.w, ids[k]) : 0.0;
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        if (t < cnt) {
          const double* Vi = v.V + (long long)(i0 + t) * nn;
#pragma unroll
          for (int k = 0; k < kGmPer; ++k) acc[t] += ldS(Vi, ids[k]) * wv[k];
        }
      }
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      if (t >= cnt) break;
      const double a = wave_sum(acc[t]);
      if (lane == 0) partials[(long long)(i0 + t) * n + w] = a;
    }
  } else {
    const bool in_cycle = OP == KG_START || OP == KG_UPDATE;
    const bool run = in_cycle ? !done(sc) : !kg_out_of_cycle(sc);
    double ww = 0.0;
    if (run) {
      if constexpr (OP == KG_AXPY1 || OP == KG_AXPY2 || OP == KG_UPDATE) {
        // z = sum_q c_q V_q over the nodes in registers, the basis vector by vector
        const double* c = OP == KG_UPDATE ? v.y : v.hc + (OP == KG_AXPY2 ? mm + 1 : 0);
        const int cnt = OP == KG_UPDATE ? ldS(&sc->steps, 0) : col + 1;
        double acc[kGmPer];
#pragma unroll
        for (int k = 0; k < kGmPer; ++k) acc[k] = 0.0;
        // four basis vectors' loads in flight at once; each node still sums q in order
        int q = 0;
        for (; q + 4 <= cnt; q += 4) {
          double c4[4], x4[4][kGmPer];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            c4[u] = ldS(c, q + u);
            const double* Vq = v.V + (long long)(q + u) * nn;
#pragma unroll
            for (int k = 0; k < kGmPer; ++k) x4[u][k] = ldS(Vq, ids[k]);
          }
#pragma unroll
          for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int k = 0; k < kGmPer; ++k) acc[k] += c4[u] * x4[u][k];
        }
        for (; q < cnt; ++q) {
          const double cq = ldS(c, q);
          const double* Vq = v.V + (long long)q * nn;
#pragma unroll
          for (int k = 0; k < kGmPer; ++k) acc[k] += cq * ldS(Vq, ids[k]);
        }
        const double sign = -1.0;
#pragma unroll
        for (int k = 0; k < kGmPer; ++k) {
          if (!in[k]) continue;
          if constexpr (OP == KG_UPDATE) {   // x += M^-1 (0 + z)
            const int j = static_cast<int>(ids[k] / v.nx), i = static_cast<int>(ids[k] - (long long)j * v.nx);
            const double z = 0.0 + 1.0 * acc[k];
            stS(v.x, ids[k], ldS(v.x, ids[k]) + 1.0 * kg_pc(v, i, j, z));
          } else {                          // w -= z
            const double wn = ldS(v.w, ids[k]) + sign * acc[k];
            stS(v.w, ids[k], wn);
            if (OP == KG_AXPY2) ww += wn * wn;
          }
        }
      } else {
        // |w| of this step, kept at hc[2 mm + 1] (the Givens step has already zeroed H[col+1][col])
        const double hn = OP == KG_NORMALIZE ? ldS(v.hc, 2 * mm + 1) : ldS(&sc->metric, 0);
        if (OP != KG_NORMALIZE || hn > 0.0) {
          const double* src = OP == KG_NORMALIZE ? v.w : v.r;
          double* dst = v.V + (OP == KG_NORMALIZE ? (long long)(col + 1) * nn : 0);
#pragma unroll
          for (int k = 0; k < kGmPer; ++k)
            if (in[k]) stS(dst, ids[k], ldS(src, ids[k]) / hn);
        }
      }
    }
    if constexpr (OP == KG_AXPY2) {
      ww = wave_sum(ww);
      if (lane == 0) partials2[w] = ww;
    }
  }
}

// MODE 0: w = A M^-1 V_col (the preconditioner applied to each stencil value on the fly);
// MODE 1: r = b - A x with partials |r|^2.  FP64 rows in sum form with the folded centre, as
// the single-problem FP64 apply.
template <int MODE>
__global__ void __launch_bounds__(kLanes * kWavesPB)
kg_apply(const BGVec* __restrict__ vecs, const BItem* __restrict__ items, int n, int col,
         double* partials) {
  const int w = wave_item();
  if (w >= n) return;
  const BItem it = items[w];
  const BGVec& v = vecs[it.p];
  const int lane = threadIdx.x & (kLanes - 1);
  const KScalars* sc = v.sc;
  double a1 = 0.0;
  const bool run = MODE == 0 ? !kg_out_of_cycle(sc) : !done(sc);
  if (run) {
    const DevOp& op = v.op;
    const W7<double> w0 = const_w<double>(op);
    auto ax_of = [&](int i, int j, double xc, double xw, double xe, double xs, double xn) -> double {
      if (i == 0 || j == 0 || i == op.nx - 1 || j == op.ny - 1) return xc;
      W7<double> wf = w0;
      fold_center<2, double>(op, i, j, 0, wf);
      return wf.C * xc + wf.E * xe + wf.W * xw + wf.N * xn + wf.S * xs;
    };
    if constexpr (MODE == 0) {
      const double* src = v.V + (long long)col * v.n;
      strip_march<double>(src, static_cast<const void*>(nullptr), op.nx, op.ny, it,
                          [&](int i, int j, long long id, double xc, double xw, double xe,
                              double xs, double xn, double) {
        double o;
        if (v.jacobi)
          o = ax_of(i, j, kg_pc(v, i, j, xc), kg_pc(v, i - 1, j, xw), kg_pc(v, i + 1, j, xe),
                    kg_pc(v, i, j - 1, xs), kg_pc(v, i, j + 1, xn));
        else
          o = ax_of(i, j, xc, xw, xe, xs, xn);
        stS(v.w, id, o);
      });
    } else {
      strip_march<double>(v.x, static_cast<const double*>(v.b), op.nx, op.ny, it,
                          [&](int i, int j, long long id, double xc, double xw, double xe,
                              double xs, double xn, double bj) {
        const double rv = bj - ax_of(i, j, xc, xw, xe, xs, xn);
        stS(v.r, id, rv);
        a1 += rv * rv;
      });
    }
  }
  if constexpr (MODE == 1) {
    a1 = wave_sum(a1);
    if (lane == 0) partials[w] = a1;
  }
}

// Givens step col of one problem (k_krylov.hip k_givens: the reference's rotation convention,
// krylov.cpp:262-295).  A breakdown ends the problem (no restart pending).
__device__ void kg_givens(const BGVec& v, KScalars* sc, int col, int mm) {
  double* H = v.H;
  double* cs = v.cs;
  double* sn = v.sn;
  double* g = v.g;
  for (int i = 0; i < col; ++i) {
    const double v1 = H[i * mm + col], v2 = H[(i + 1) * mm + col];
    H[i * mm + col] = cs[i] * v1 - sn[i] * v2;
    H[(i + 1) * mm + col] = sn[i] * v1 + cs[i] * v2;
  }
  const double h1 = H[(col + 1) * mm + col], h0 = H[col * mm + col];
  const double denom = sqrt(h0 * h0 + h1 * h1);
  if (denom < 1e-30) { sc->breakdown = 1; sc->cycle_done = 0; return; }
  cs[col] = h0 / denom;
  sn[col] = -h1 / denom;
  H[col * mm + col] = cs[col] * h0 - sn[col] * h1;
  H[(col + 1) * mm + col] = 0.0;
  const double g0 = g[col], g1 = g[col + 1];
  g[col] = cs[col] * g0 - sn[col] * g1;
  g[col + 1] = sn[col] * g0 + cs[col] * g1;
  sc->steps = col + 1;
  sc->iter += 1;
  sc->metric = fabs(g[col + 1]);
  if (fabs(g[col + 1]) < sc->target) sc->cycle_done = 1;
}

// Segmented finalize, one 512-thread block per problem.  The coefficient sums (GF_H1/GF_H2)
// take one wavefront per basis vector, with eight partial loads in flight per lane; every sum
// runs over the problem's items in a fixed order.
constexpr int kGfThreads = 512;
__global__ void __launch_bounds__(kGfThreads)
kg_finalize(const BGVec* __restrict__ vecs, const int* __restrict__ first,
            const int* __restrict__ cnt, int nitems, int nprob, const double* __restrict__ partials,
            int op, int col, int mm, double tol, int tol_mode) {
  __shared__ double lds[kGfThreads / kWave];
  const int q = blockIdx.x;
  if (q >= nprob) return;
  const BGVec& v = vecs[q];
  KScalars* sc = v.sc;
  if (sc->done | sc->breakdown) return;
  if ((op == GF_H1 || op == GF_H2 || op == GF_NORM) && sc->cycle_done) return;
  const int f = first[q], c = cnt[q];
  if (op == GF_H1 || op == GF_H2) {
    const int lane = threadIdx.x & (kLanes - 1);
    const int wv = static_cast<int>(threadIdx.x / kLanes);
    for (int i = wv; i <= col; i += static_cast<int>(blockDim.x / kLanes)) {
      double a = 0.0;
      const double* pi = partials + (long long)i * nitems + f;
#pragma unroll 8
      for (int t = lane; t < c; t += kLanes) a += pi[t];
      a = wave_sum(a);
      if (lane == 0) {
        if (op == GF_H1) {
          v.hc[i] = a;
        } else {
          v.hc[mm + 1 + i] = a;
          v.H[(long long)i * mm + col] = v.hc[i] + a;
        }
      }
    }
    return;
  }
  double a = 0.0;
  for (int t = threadIdx.x; t < c; t += blockDim.x) a += partials[f + t];
  const double s1 = block_sum(a, lds);
  if (threadIdx.x != 0) return;
  switch (op) {
    case GF_NORM:
      v.H[(long long)(col + 1) * mm + col] = sqrt(s1);
      v.hc[2 * mm + 1] = sqrt(s1);   // for the normalisation, which runs after the rotation
      kg_givens(v, sc, col, mm);
      break;
    case GF_R0: {
      const double r0 = sqrt(s1);
      if (nonfinite(s1)) { sc->breakdown = 1; break; }
      sc->r0 = r0;
      sc->metric = r0;
      sc->target = tol_mode == PDEHIP_TOL_REL ? tol * r0 : tol;
      if (r0 < sc->target || r0 == 0.0) sc->done = 1;
      break;
    }
    case GF_R:
      sc->metric = sqrt(s1);
      if (nonfinite(s1)) { sc->breakdown = 1; sc->cycle_done = 0; break; }
      if (sc->metric < sc->target) sc->done = 1;
      break;
    default:
      break;
  }
}

// Per-problem cycle bookkeeping, one 64-thread block per problem.
__global__ void kg_cycle(const BGVec* __restrict__ vecs, int nprob, int op, int mm) {
  const int q = blockIdx.x;
  if (q >= nprob) return;
  const BGVec& v = vecs[q];
  KScalars* sc = v.sc;
  if (sc->done | sc->breakdown) return;
  if (op == GC_START) {
    const double beta = sc->metric;
    for (int t = threadIdx.x; t < (mm + 1) * mm; t += blockDim.x) v.H[t] = 0.0;
    for (int t = threadIdx.x; t <= mm; t += blockDim.x) v.g[t] = t == 0 ? beta : 0.0;
    if (threadIdx.x == 0) { sc->cycle_done = 0; sc->steps = 0; }
    return;
  }
  if (threadIdx.x != 0) return;
  // back substitution over the cycle's steps (k_krylov.hip k_backsub)
  const double* H = v.H;
  const double* g = v.g;
  double* y = v.y;
  const int dim = sc->steps;
  for (int i = dim - 1; i >= 0; --i) {
    double sum = g[i];
    for (int t = i + 1; t < dim; ++t) sum -= H[i * mm + t] * y[t];
    if (fabs(H[i * mm + i]) < 1e-30) { sc->breakdown = 1; sc->cycle_done = 0; return; }
    y[i] = sum / H[i * mm + i];
  }
  for (int i = dim; i < mm; ++i) y[i] = 0.0;
}

// ------------------------------------------------------------------------------- finish
// Boundary closures of problem blockIdx.x's FP64 solution, one face pair per launch (phase).
__global__ void __launch_bounds__(256)
kb_closure(const BVec* __restrict__ vecs, const BFace* __restrict__ faces, int nprob, int phase) {
  const int q0 = blockIdx.x;
  if (q0 >= nprob) return;
  const BVec& v = vecs[q0];
  const FaceData& fd = faces[q0].fd;
  const int axis = phase & 1;
  const bool dir = phase >= 2;
  const int nx = v.nx, ny = v.ny;
  const int nq = axis == 0 ? ny : nx;
  double* u = v.x64;
  for (int q = threadIdx.x; q < nq; q += blockDim.x) {
    for (int side = 0; side < 2; ++side) {
      const int f = 2 * axis + side;
      const bool is_d = fd.kind[f] == PDEHIP_BC_DIRICHLET;
      if (is_d != dir) continue;
      long long node, inner;
      if (axis == 0) {
        node = (long long)q * nx + (side ? nx - 1 : 0);
        inner = (long long)q * nx + (side ? nx - 2 : 1);
      } else {
        node = (long long)(side ? ny - 1 : 0) * nx + q;
        inner = (long long)(side ? ny - 2 : 1) * nx + q;
      }
      if (is_d) {
        u[node] = fd.value[f][q];
      } else {
        double val;
        if (closure_nr_value(fd, f, side, q, u[inner], &val)) u[node] = val;
      }
    }
  }
}

// Reference residual norms of the FP64 solution over element items (one workgroup each).
__global__ void __launch_bounds__(256)
kb_resnorm(const BVec* __restrict__ vecs, const BFace* __restrict__ faces,
           const BItem* __restrict__ items, int n, double* sums, double* maxs) {
  __shared__ double lds[256 / kWave];
  const int w = blockIdx.x;
  if (w >= n) return;
  const BItem it = items[w];
  const BVec& v = vecs[it.p];
  const BFace& bf = faces[it.p];
  const DevOp& op = v.op64;
  const double* __restrict__ u = v.x64;
  const W7<double> wt = const_w<double>(op);
  double acc = 0.0, mx = 0.0;
  for (long long id = it.a + threadIdx.x; id < it.b; id += blockDim.x) {
    const IJK c = coords(op, id);
    const int i = c.i, j = c.j;
    double r;
    if (is_boundary2(op, i, j)) {
      if (!resnorm_boundary2(op, bf.fd, u, id, i, j, &r)) continue;
    } else {
      const double sn = wt.E * u[id + 1] + wt.W * u[id - 1] + wt.N * u[id + op.nx] + wt.S * u[id - op.nx];
      r = wt.C * u[id] + sn + (bf.f ? bf.f[id] : bf.f_const);
    }
    acc += r * r;
    mx = nmax(mx, fabs(r));
  }
  const double s = block_sum(acc, lds);
  const double m = block_max(mx, lds);
  if (threadIdx.x == 0) { sums[w] = s; maxs[w] = m; }
}

// Per-problem gather: sums of the solved-system residual partials (row items) and of the
// reference-norm partials (element items) in a fixed order, and the outer scalars.
__global__ void __launch_bounds__(64)
kb_finish_reduce(const BVec* __restrict__ vecs, const int* __restrict__ rfirst,
                 const int* __restrict__ rcnt, const double* __restrict__ rr_part,
                 const int* __restrict__ efirst, const int* __restrict__ ecnt,
                 const double* __restrict__ sums, const double* __restrict__ maxs, int nprob,
                 BReport* __restrict__ out) {
  const int q = blockIdx.x;
  if (q >= nprob) return;
  const int lane = threadIdx.x;
  double rr = 0.0, l2 = 0.0, mx = 0.0;
  for (int t = lane; t < rcnt[q]; t += 64) rr += rr_part[rfirst[q] + t];
  for (int t = lane; t < ecnt[q]; t += 64) {
    l2 += sums[efirst[q] + t];
    mx = nmax(mx, maxs[efirst[q] + t]);
  }
  rr = wave_sum(rr);
  l2 = wave_sum(l2);
  mx = wave_max(mx);
  if (lane != 0) return;
  const KScalars* so = vecs[q].so;
  const KScalars* si = vecs[q].si;
  BReport r{};
  r.rr = rr;
  r.l2sq = l2;
  r.linf = mx;
  r.metric = so->metric;
  r.r0 = so->r0;
  r.done = so->done;
  r.breakdown = so->breakdown | si->breakdown;
  r.iter = so->iter;
  r.steps = so->steps;
  r.inner = so->pad0;
  out[q] = r;
}

// Workgroups for n wavefront items, rounded up to a multiple of 8 (XCD-contiguous numbering).
unsigned waves_grid(int n) {
  const int wg = (n + kWavesPB - 1) / kWavesPB;
  return static_cast<unsigned>((wg + 7) / 8 * 8);
}

}  // namespace

void sweep(hipStream_t s, Prec p, const BEnt* tab, const BItem* items, int n, int c1, int dir,
           bool zero) {
  if (n <= 0) return;
  const dim3 g(waves_grid(n)), b(kLanes * kWavesPB);
#define SW_(PP, ZZ) hipLaunchKernelGGL((kb_sweep<PP, ZZ>), g, b, 0, s, tab, items, n, c1, dir)
  switch (p) {
    case P64: if (zero) SW_(P64, true); else SW_(P64, false); break;
    case P32: if (zero) SW_(P32, true); else SW_(P32, false); break;
    default: if (zero) SW_(P16, true); else SW_(P16, false); break;
  }
#undef SW_
}

void residual(hipStream_t s, Prec p, const BEnt* tab, const BItem* items, int n) {
  if (n <= 0) return;
  const dim3 g(waves_grid(n)), b(kLanes * kWavesPB);
  switch (p) {
    case P64: hipLaunchKernelGGL((kb_residual<P64>), g, b, 0, s, tab, items, n); break;
    case P32: hipLaunchKernelGGL((kb_residual<P32>), g, b, 0, s, tab, items, n); break;
    default: hipLaunchKernelGGL((kb_residual<P16>), g, b, 0, s, tab, items, n); break;
  }
}

void restrict_(hipStream_t s, Prec pf, Prec pc, const BEnt* fine, const BEnt* coarse,
               const BItem* items, int n, bool scale0) {
  if (n <= 0) return;
  const dim3 g(waves_grid(n)), b(kLanes * kWavesPB);
  const int sc = scale0 ? 1 : 0;
#define RB_(A, B) hipLaunchKernelGGL((kb_restrict<A, B>), g, b, 0, s, fine, coarse, items, n, sc)
  if (pf == P64 && pc == P64) RB_(P64, P64);
  else if (pf == P32 && pc == P16) RB_(P32, P16);
  else if (pf == P32 && pc == P32) RB_(P32, P32);
  else if (pf == P16 && pc == P16) RB_(P16, P16);
  else throw Error(PDEHIP_ERR_UNSUPPORTED, "batched restriction: unsupported precision pair");
#undef RB_
}

void prolong(hipStream_t s, Prec pc, Prec pf, const BEnt* coarse, const BEnt* fine,
             const BItem* items, int n, bool scale0) {
  if (n <= 0) return;
  const dim3 g(waves_grid(n)), b(kLanes * kWavesPB);
  const int sc = scale0 ? 1 : 0;
#define PB_(A, B) hipLaunchKernelGGL((kb_prolong<A, B>), g, b, 0, s, coarse, fine, items, n, sc)
  if (pc == P64 && pf == P64) PB_(P64, P64);
  else if (pc == P16 && pf == P32) PB_(P16, P32);
  else if (pc == P32 && pf == P32) PB_(P32, P32);
  else if (pc == P16 && pf == P16) PB_(P16, P16);
  else throw Error(PDEHIP_ERR_UNSUPPORTED, "batched prolongation: unsupported precision pair");
#undef PB_
}

void coarse_dense(hipStream_t s, Prec p, const BEnt* tab, const BDense* dense, const BItem* items,
                  int n) {
  if (n <= 0) return;
  const dim3 g(waves_grid(n)), b(kLanes * kWavesPB);
  switch (p) {
    case P64: hipLaunchKernelGGL((kb_dense<P64>), g, b, 0, s, tab, dense, items, n); break;
    case P32: hipLaunchKernelGGL((kb_dense<P32>), g, b, 0, s, tab, dense, items, n); break;
    default: hipLaunchKernelGGL((kb_dense<P16>), g, b, 0, s, tab, dense, items, n); break;
  }
}

void vec(hipStream_t s, Prec p, int op, const BVec* vecs, const BItem* items, int n,
         double* partials, double* maxpart) {
  if (n <= 0) return;
  const int inner = (p == P32) ? 1 : 0;
  if (p == P64)
    hipLaunchKernelGGL((kb_vec<P64>), dim3(n), dim3(256), 0, s, vecs, items, n, op, inner, partials, maxpart);
  else
    hipLaunchKernelGGL((kb_vec<P32>), dim3(n), dim3(256), 0, s, vecs, items, n, op, inner, partials, maxpart);
}

void apply_dot_tab(hipStream_t s, Prec p, const BVec* vecs, const BEnt* tab0, const BItem* rows,
                   int n, double* partials, bool outer, bool all) {
  if (n <= 0) return;
  const dim3 g(waves_grid(n)), b(kLanes * kWavesPB);
  if (outer && all)
    hipLaunchKernelGGL((kb_apply_dot<P64, true, true>), g, b, 0, s, vecs, tab0, rows, n, partials);
  else if (outer)
    hipLaunchKernelGGL((kb_apply_dot<P64, true>), g, b, 0, s, vecs, tab0, rows, n, partials);
  else if (p == P64)
    hipLaunchKernelGGL((kb_apply_dot<P64, false>), g, b, 0, s, vecs, tab0, rows, n, partials);
  else
    hipLaunchKernelGGL((kb_apply_dot<P32, false>), g, b, 0, s, vecs, tab0, rows, n, partials);
}

void finalize(hipStream_t s, const BVec* vecs, const int* first, const int* cnt, int nprob,
              const double* partials, const double* maxpart, int op, int inner, double tol,
              int tol_mode) {
  hipLaunchKernelGGL(kb_finalize, dim3(nprob), dim3(256), 0, s, vecs, first, cnt, nprob, partials,
                     maxpart, op, inner, tol, tol_mode);
}

void reset_inner(hipStream_t s, const BVec* vecs, int nprob) {
  hipLaunchKernelGGL(kb_reset_inner, dim3((nprob + 255) / 256), dim3(256), 0, s, vecs, nprob, 1);
}

void replace_residual(hipStream_t s, const BVec* vecs, const BItem* rows, int n, double* partials) {
  if (n <= 0) return;
  hipLaunchKernelGGL(kb_replace, dim3(waves_grid(n)), dim3(kLanes * kWavesPB), 0, s, vecs, rows, n,
                     partials);
}

void ru_start(hipStream_t s, const BVec* vecs, int nprob) {
  hipLaunchKernelGGL(kb_ru_start, dim3((nprob + 255) / 256), dim3(256), 0, s, vecs, nprob);
}

void init_scalars(hipStream_t s, const BVec* vecs, int nprob, double eps) {
  hipLaunchKernelGGL(kb_init_scalars, dim3((nprob + 255) / 256), dim3(256), 0, s, vecs, nprob, eps);
}

void closures(hipStream_t s, const BVec* vecs, const BFace* faces, int nprob, int phase) {
  hipLaunchKernelGGL(kb_closure, dim3(nprob), dim3(256), 0, s, vecs, faces, nprob, phase);
}

void resnorm(hipStream_t s, const BVec* vecs, const BFace* faces, const BItem* items, int n,
             double* sums, double* maxs) {
  if (n <= 0) return;
  hipLaunchKernelGGL(kb_resnorm, dim3(n), dim3(256), 0, s, vecs, faces, items, n, sums, maxs);
}

void finish_reduce(hipStream_t s, const BVec* vecs, const int* rfirst, const int* rcnt,
                   const double* rr_part, const int* efirst, const int* ecnt, const double* sums,
                   const double* maxs, int nprob, BReport* out) {
  hipLaunchKernelGGL(kb_finish_reduce, dim3(nprob), dim3(64), 0, s, vecs, rfirst, rcnt, rr_part,
                     efirst, ecnt, sums, maxs, nprob, out);
}

void active_count(hipStream_t s, const BVec* vecs, int nprob, int inner, int* active) {
  hipLaunchKernelGGL(kb_active, dim3(1), dim3(1024), 0, s, vecs, nprob, inner, active);
}

void cancel(hipStream_t s, const BVec* vecs, int p) {
  hipLaunchKernelGGL(kb_cancel, dim3(1), dim3(64), 0, s, vecs, p);
}


void kvec(hipStream_t s, Prec p, int op, const BKVec* vecs, const BItem* items, int n,
          double* partials) {
  if (n <= 0) return;
  const dim3 g(waves_grid(n)), b(kLanes * kWavesPB);
#define KV_(S, OP) hipLaunchKernelGGL((kk_vec<S, OP>), g, b, 0, s, vecs, items, n, partials)
#define KV_ALL(S)                                \
  switch (op) {                                  \
    case KK_RH: KV_(S, KK_RH); break;            \
    case KK_RHO: KV_(S, KK_RHO); break;          \
    case KK_P: KV_(S, KK_P); break;              \
    case KK_S: KV_(S, KK_S); break;              \
    case KK_XS: KV_(S, KK_XS); break;            \
    case KK_RESTART: KV_(S, KK_RESTART); break;  \
    case KK_X0: KV_(S, KK_X0); break;            \
    default: KV_(S, KK_X); break;                \
  }
  if (p == P64) { KV_ALL(double) } else { KV_ALL(float) }
#undef KV_ALL
#undef KV_
}

void kapply(hipStream_t s, Prec p, int mode, const BKVec* vecs, const BItem* items, int n,
            double* p1, double* p2) {
  if (n <= 0) return;
  const dim3 g(waves_grid(n)), b(kLanes * kWavesPB);
#define KA_(S, M) hipLaunchKernelGGL((kk_apply<S, M>), g, b, 0, s, vecs, items, n, p1, p2)
  if (p == P64) {
    if (mode == 0) KA_(double, 0); else if (mode == 1) KA_(double, 1); else KA_(double, 2);
  } else {
    if (mode == 0) KA_(float, 0); else if (mode == 1) KA_(float, 1); else KA_(float, 2);
  }
#undef KA_
}

void kfailed(hipStream_t s, const BKVec* vecs, int nprob, int* out) {
  hipLaunchKernelGGL(kk_failed, dim3(1), dim3(1024), 0, s, vecs, nprob, out);
}

void kreset_failed(hipStream_t s, const BKVec* vecs, int nprob, double eps) {
  hipLaunchKernelGGL(kk_reset_failed, dim3((nprob + 255) / 256), dim3(256), 0, s, vecs, nprob, eps);
}

void krestart(hipStream_t s, const BKVec* vecs, int nprob) {
  hipLaunchKernelGGL(kk_restart, dim3((nprob + 255) / 256), dim3(256), 0, s, vecs, nprob);
}

void kfinalize(hipStream_t s, const BKVec* vecs, const int* first, const int* cnt, int nprob,
               const double* p1, const double* p2, int op, double tol, int tol_mode) {
  hipLaunchKernelGGL(kk_finalize, dim3(nprob), dim3(256), 0, s, vecs, first, cnt, nprob, p1, p2,
                     op, tol, tol_mode);
}

void gvec(hipStream_t s, int op, const BGVec* vecs, const BItem* items, int n, int col, int mm,
          double* partials, double* partials2) {
  if (n <= 0) return;
  const dim3 b(kLanes * kWavesPB);
  const unsigned gx = waves_grid(n);
#define KG_(OP, GY) hipLaunchKernelGGL((kg_vec<OP>), dim3(gx, GY), b, 0, s, vecs, items, n, col, mm, partials, partials2)
  switch (op) {
    case KG_DOT: KG_(KG_DOT, static_cast<unsigned>((col + 1 + 7) / 8)); break;
    case KG_AXPY1: KG_(KG_AXPY1, 1u); break;
    case KG_AXPY2: KG_(KG_AXPY2, 1u); break;
    case KG_NORMALIZE: KG_(KG_NORMALIZE, 1u); break;
    case KG_START: KG_(KG_START, 1u); break;
    default: KG_(KG_UPDATE, 1u); break;
  }
#undef KG_
}

void gapply(hipStream_t s, int mode, const BGVec* vecs, const BItem* items, int n, int col,
            double* partials) {
  if (n <= 0) return;
  const dim3 g(waves_grid(n)), b(kLanes * kWavesPB);
  if (mode == 0) hipLaunchKernelGGL((kg_apply<0>), g, b, 0, s, vecs, items, n, col, partials);
  else hipLaunchKernelGGL((kg_apply<1>), g, b, 0, s, vecs, items, n, col, partials);
}

void gfinalize(hipStream_t s, const BGVec* vecs, const int* first, const int* cnt, int nitems,
               int nprob, const double* partials, int op, int col, int mm, double tol,
               int tol_mode) {
  hipLaunchKernelGGL(kg_finalize, dim3(nprob), dim3(kGfThreads), 0, s, vecs, first, cnt, nitems,
                     nprob, partials, op, col, mm, tol, tol_mode);
}

void gcycle(hipStream_t s, const BGVec* vecs, int nprob, int op, int mm) {
  hipLaunchKernelGGL(kg_cycle, dim3(nprob), dim3(64), 0, s, vecs, nprob, op, mm);
}

}  // namespace bk
}  // namespace pdehip
