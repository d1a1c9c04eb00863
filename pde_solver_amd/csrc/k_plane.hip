// k_plane.hip — 3D operator apply and residual as a plane march for gfx950.
//
// A workgroup (8 waves) owns a 62 x (8R-2) (i,j) tile (extended by a 1-node ring to 64 lanes
// x 8R rows) and walks a chunk of z planes.  Per plane p:
//   * x(p) and kappa(p+1) go to LDS (two slots each);
//   * FLUX: every node of the extended tile forms its east and north faces of plane p+1 once
//     (two LDS slots) and its front face in registers;
//   * the tile's plane p is produced from x neighbours in LDS / registers and the faces of
//     plane p formed one step earlier: y = A x (partials of x.y) or r = b - A x (partials of
//     r.r).  Boundary rows are identity.
// Double-buffered slots make one barrier per plane enough: a slot is rewritten two planes
// after it was read, and a thread can only be there once every thread has passed the barrier
// in between.  The FP64 operator divides exactly (face_coef); the FP32/FP16 operators keep
// rho = 1/kappa and form each face with one hardware reciprocal, within their own rounding.
#include <algorithm>
#include <type_traits>

#include "bufio.hpp"
#include "kernels.hpp"
#include "kernels_ext.hpp"

namespace pdehip {
namespace {

constexpr int kLanes = 64;
constexpr int kWaves = 8;
constexpr int kThreads = kLanes * kWaves;

template <Prec P> struct PT;
template <> struct PT<P64> { using T = double; using S = double; };
template <> struct PT<P32> { using T = float; using S = float; };
template <> struct PT<P16> { using T = float; using S = __half; };

template <typename S> __device__ __forceinline__ auto ldS(const S* p, long long i);
template <> __device__ __forceinline__ auto ldS<double>(const double* p, long long i) { return p[i]; }
template <> __device__ __forceinline__ auto ldS<float>(const float* p, long long i) { return p[i]; }
template <> __device__ __forceinline__ auto ldS<__half>(const __half* p, long long i) { return __half2float(p[i]); }

__device__ __forceinline__ bool skip(const KScalars* sc) { return sc && (sc->done | sc->breakdown); }

template <Prec P, typename T>
__device__ __forceinline__ T face_op(T kc, T kn, T f) {
  return face_coef<T>(kc, kn, f);
}

// A launch with fewer workgroups than the npart partials its finalize sums (4-byte levels run
// taller tiles than FP64 ones, so fewer of them) zeroes the ones past its grid.
__device__ __forceinline__ void pad_partials(double* partials, int npart) {
  if (partials && blockIdx.x == 0)
    for (int z = static_cast<int>(gridDim.x) + static_cast<int>(threadIdx.x); z < npart;
         z += static_cast<int>(blockDim.x))
      partials[z] = 0.0;
}

// Residual replacement of a mixed solve (MODE 2): the two FP64 solution buffers, the FP32
// correction, the scaled FP32 residual out and the outer scalars (xsel, snorm).  Reference
// residual norms (MODE 4): the source (field or constant), the max |r| partials and the
// degenerate-row flag.
struct ReplArgs {
  double* x[2];
  const float* e;
  float* r32;
  const KScalars* so;
  const double* f;
  double f_const;
  double* maxs;
  int* degenerate;
};

// MODE 0: y = A x, partials of x.y.  MODE 1: r = b - A x, partials of r.r.  MODE 2 (FP64 only):
// MODE 1 on x = x[sel] + s e, which also goes to x[1 - sel], and r / s to r32.  MODE 4 (FP64
// only): the reference residual L u + f of the interior rows with the unfolded weights
// (k_resnorm3's rows: the same weights, the neighbour sum in its order), nothing stored,
// partials of r.r and of max |r| (ra.maxs); a row whose |C| < 1e-12 sets *ra.degenerate.
template <int KIND, Prec P, int MODE, int R>
__global__ void __launch_bounds__(kThreads)
k_plane3(DevOp op, const void* __restrict__ bv, const void* __restrict__ xv, void* __restrict__ yv,
         double* partials, int npart, int tiles_x, int tiles_y, int ntiles, int kchunk,
         const KScalars* sc, ReplArgs ra) {
  static_assert((MODE != 2 && MODE != 4) || P == P64, "replacement / norms: FP64 operator only");
  using T = typename PT<P>::T;
  using S = typename PT<P>::S;
  constexpr int TYE = R * kWaves;
  constexpr int TY = TYE - 2;
  constexpr int TX = kLanes - 2;
  constexpr bool HK = KIND == PDEHIP_OP_FLUX;
  constexpr int KR = HK ? TYE : 1;
  __shared__ T lx[2][TYE][kLanes];
  __shared__ T lk[HK ? 2 : 1][KR][kLanes];
  __shared__ T lfe[HK ? 2 : 1][KR][kLanes];
  __shared__ T lfn[HK ? 2 : 1][KR][kLanes];
  __shared__ double lred[kWaves];
  if (skip(sc)) return;
  pad_partials(partials, npart);
  if constexpr (MODE == 4) pad_partials(ra.maxs, npart);
  const int per = static_cast<int>(gridDim.x >> 3);
  const int t = static_cast<int>(blockIdx.x & 7) * per + static_cast<int>(blockIdx.x >> 3);
  if (t >= ntiles) {
    if (partials && threadIdx.x == 0) partials[blockIdx.x] = 0.0;
    if constexpr (MODE == 4)
      if (threadIdx.x == 0) ra.maxs[blockIdx.x] = 0.0;
    return;
  }
  const S* __restrict__ x = static_cast<const S*>(xv);
  const S* __restrict__ b = static_cast<const S*>(bv);
  const T* __restrict__ kap = static_cast<const T*>(op.kappa);
  S* __restrict__ y = static_cast<S*>(yv);
  double* __restrict__ xo = nullptr;
  double rs = 0.0, rinv = 1.0;
  if constexpr (MODE == 2) {
    const int sel = ra.so->xsel;
    x = ra.x[sel];
    xo = ra.x[sel ^ 1];
    rs = ra.so->snorm;
    rinv = rs > 0.0 ? 1.0 / rs : 1.0;
  }
  const int tx = t % tiles_x;
  const int rest = t / tiles_x;
  const int ty = rest % tiles_y;
  const int cz = rest / tiles_y;
  const int lane = threadIdx.x & (kLanes - 1);
  const int wv = threadIdx.x / kLanes;
  const int i = tx * TX - 1 + lane;
  const int k0 = cz * kchunk;
  const int k1 = min(k0 + kchunk, op.nz);
  const int nz = op.nz;
  const long long sxy = op.sxy;
  const T fx = T(op.fx), fy = T(op.fy), fz = T(op.fz);
  const bool in_x = i >= 0 && i < op.nx;
  const bool bnd_x = i == 0 || i == op.nx - 1;
  const bool tile_x = lane >= 1 && lane <= kLanes - 2;

  // Lanes and rows outside the grid load a clamped in-grid node: no output reads their values.
  unsigned m_out = 0, m_bxy = 0, m_fxy = 0;
  long long col[R];
  const int ci = min(max(i, 0), op.nx - 1);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int lr = wv * R + r;
    const int j = ty * TY - 1 + lr;
    const bool in_y = j >= 0 && j < op.ny;
    col[r] = (long long)min(max(j, 0), op.ny - 1) * op.nx + ci;
    if (in_x && in_y && tile_x && lr >= 1 && lr <= TYE - 2) m_out |= 1u << r;
    if (bnd_x || j == 0 || j == op.ny - 1) m_bxy |= 1u << r;
    if (i == 1 || i == op.nx - 2 || j == 1 || j == op.ny - 2) m_fxy |= 1u << r;
  }
  const int jrow0 = ty * TY - 1 + wv * R;

  auto kcl = [&](int k) -> long long { return (long long)min(max(k, 0), nz - 1) * sxy; };
  auto ld_x = [&](int r, int k) -> T {
    const long long id = kcl(k) + col[r];
    if constexpr (MODE == 2) return fma(rs, static_cast<double>(ra.e[id]), x[id]);
    else return T(ldS(x, id));
  };
  auto ld_k = [&](int r, int k) -> T {
    if constexpr (HK) return kap[kcl(k) + col[r]];
    else return T(0);
  };
  auto ld_b = [&](int r, int k) -> T {
    if constexpr (MODE == 4) return ra.f ? ra.f[kcl(k) + col[r]] : ra.f_const;
    else if constexpr (MODE >= 1) return T(ldS(b, kcl(k) + col[r]));
    else return T(0);
  };

  // queues after the shift of step p: X = x(p-1..p+1), XN = x(p+2) (prefetch),
  // K = kappa(p+1..p+2), ZF = front faces zf(p-1..p+1) with zf(k) = face(kappa(k), kappa(k+1)),
  // B = b(p..p+1).  FLUX starts two planes early so zf(k0-1) and the faces of plane k0 exist.
  const int pstart = HK ? k0 - 2 : k0;
  T X[R][3], XN[R], K[R][2], ZF[R][3], B[R][2];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    X[r][0] = T(0);
    X[r][1] = ld_x(r, pstart - 1);
    X[r][2] = ld_x(r, pstart);
    XN[r] = ld_x(r, pstart + 1);
    K[r][0] = T(0);
    K[r][1] = ld_k(r, pstart + 1);
    ZF[r][0] = ZF[r][1] = ZF[r][2] = T(0);
    B[r][0] = T(0);
    B[r][1] = ld_b(r, pstart);
  }
  double acc = 0.0, mxv = 0.0;
  for (int p = pstart; p < k1; ++p) {
    const int sl = p & 1;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int lr = wv * R + r;
      X[r][0] = X[r][1]; X[r][1] = X[r][2]; X[r][2] = XN[r]; XN[r] = ld_x(r, p + 2);
      B[r][0] = B[r][1]; B[r][1] = ld_b(r, p + 1);
      if constexpr (HK) {
        K[r][0] = K[r][1]; K[r][1] = ld_k(r, p + 2);
        ZF[r][0] = ZF[r][1]; ZF[r][1] = ZF[r][2];
        lk[sl][lr][lane] = K[r][0];
      }
      lx[sl][lr][lane] = X[r][1];
    }
    __syncthreads();
    if constexpr (HK) {
      // faces of plane p+1 into slot (p+1) & 1
      const int fs = sl ^ 1;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int lr = wv * R + r;
        const T kc = K[r][0];
        ZF[r][2] = face_op<P, T>(kc, K[r][1], fz);
        lfe[fs][lr][lane] = lane < kLanes - 1 ? face_op<P, T>(kc, lk[sl][lr][lane + 1], fx) : T(0);
        lfn[fs][lr][lane] = lr < TYE - 1 ? face_op<P, T>(kc, lk[sl][lr + 1][lane], fy) : T(0);
      }
    }
    if (p >= k0) {
      // plane p of the tile; its faces were formed one step earlier (slot p & 1)
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (!((m_out >> r) & 1u)) continue;
        const int lr = wv * R + r;
        const int j = jrow0 + r;
        const T xc = X[r][1];
        T out;
        if (((m_bxy >> r) & 1u) || p == 0 || p == nz - 1) {
          if constexpr (MODE == 4) continue;   // interior rows only
          out = MODE == 0 ? xc : T(B[r][0] - xc);
        } else {
          W7<T> w;
          if constexpr (KIND == PDEHIP_OP_CONST) {
            w.C = T(op.w[0]); w.W = T(op.w[1]); w.E = T(op.w[2]); w.S = T(op.w[3]);
            w.N = T(op.w[4]); w.B = T(op.w[5]); w.F = T(op.w[6]);
          } else if constexpr (KIND == PDEHIP_OP_POINTWISE) {
            const long long id = p * sxy + col[r];
            const double* pw = op.pw;
            const long long n = op.n;
            w.C = T(pw[id]); w.W = T(pw[n + id]); w.E = T(pw[2 * n + id]); w.S = T(pw[3 * n + id]);
            w.N = T(pw[4 * n + id]); w.B = T(pw[5 * n + id]); w.F = T(pw[6 * n + id]);
          } else {
            w.W = lfe[sl][lr][lane - 1]; w.E = lfe[sl][lr][lane];
            w.S = lfn[sl][lr - 1][lane]; w.N = lfn[sl][lr][lane];
            T sum = w.W + w.E + w.S + w.N;
            w.B = ZF[r][0]; w.F = ZF[r][1];
            sum += w.B + w.F;
            w.C = T(op.e) - sum;
          }
          if constexpr (MODE == 4) {
            if (fabs(w.C) < 1e-12) { *ra.degenerate = 1; continue; }
            double sn = w.E * lx[sl][lr][lane + 1] + w.W * lx[sl][lr][lane - 1] +
                        w.N * lx[sl][lr + 1][lane] + w.S * lx[sl][lr - 1][lane];
            sn = sn + w.F * X[r][2] + w.B * X[r][0];
            const double rv = w.C * xc + sn + B[r][0];
            acc += rv * rv;
            mxv = nmax(mxv, fabs(rv));
            continue;
          }
          if (((m_fxy >> r) & 1u) || p == 1 || p == nz - 2) fold_center<3, T>(op, i, j, p, w);
          const T ax = apply_row<T>(w, xc, lx[sl][lr][lane + 1], lx[sl][lr][lane - 1],
                                    lx[sl][lr + 1][lane], lx[sl][lr - 1][lane], X[r][2], X[r][0]);
          out = MODE == 0 ? ax : T(B[r][0] - ax);
        }
        y[p * sxy + col[r]] = store_as<S>(static_cast<double>(out));
        if constexpr (MODE == 2) {
          xo[p * sxy + col[r]] = xc;
          ra.r32[p * sxy + col[r]] = static_cast<float>(out * rinv);
        }
        if (partials)
          acc += MODE == 0 ? static_cast<double>(xc) * static_cast<double>(out)
                           : static_cast<double>(out) * static_cast<double>(out);
      }
    }
  }
  if (partials) {
    const double s = block_sum(acc, lred);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
  }
  if constexpr (MODE == 4) {
    __syncthreads();   // lred reused
    const double m = block_max(mxv, lred);
    if (threadIdx.x == 0) ra.maxs[blockIdx.x] = m;
  }
}

// The same plane march for 4-byte arithmetic (FP32/FP16 storage, CONST and FLUX), with the
// bookkeeping taken out of the plane step: the loop is unrolled kRingP = 6 times and every
// register queue is a ring indexed by the step's phase at compile time (x: 6 planes, loads
// three steps ahead; b, kappa, front faces: 3), LDS slots are compile-time, LDS planes are
// padded by one row / column on every side (neighbour reads at immediate offsets, no clamps),
// global traffic goes through per-plane buffer descriptors, and stores of lanes outside the
// output tile are discarded by the descriptor's range check.  Same tile, same arithmetic in
// the same order as k_plane3, so the outputs and the partials are the same bits.  Steps past
// the chunk compute and discard.
//
// MODE 3 (FP32): the CG direction update folded into the apply: with bv = z and xv = p, every
// loaded node forms p' = z + beta p (beta = sc->beta, the FMA of the separate update), the
// march writes p' to pov and y = A p' with the partials of p'.y.  p' of a plane is formed just
// before its first use, so the z and p loads keep their three-plane lead.  pov must not be xv:
// neighbouring tiles read xv's halo.
constexpr int kRingP = 6;

template <int KIND, Prec P, int MODE, int R>
__global__ void __launch_bounds__(kThreads)
k_plane3r(DevOp op, const void* __restrict__ bv, const void* __restrict__ xv, void* __restrict__ yv,
          double* partials, int npart, int tiles_x, int tiles_y, int ntiles, int kchunk,
          const KScalars* sc, void* __restrict__ pov) {
  using T = typename PT<P>::T;
  using S = typename PT<P>::S;
  static_assert(sizeof(T) == 4, "ring plane march: 4-byte arithmetic types only");
  static_assert(MODE != 3 || P == P32, "fused direction update: FP32 vectors only");
  constexpr bool XP = MODE == 3;
  constexpr int TYE = R * kWaves;
  constexpr int TY = TYE - 2;
  constexpr int TX = kLanes - 2;
  constexpr bool HK = KIND == PDEHIP_OP_FLUX;
  constexpr int KR = HK ? TYE : 1;
  constexpr int PW = kLanes + 2;
  __shared__ T lx[2][TYE + 2][PW];
  __shared__ T lk[HK ? 2 : 1][KR + 2][PW];
  __shared__ T lfe[HK ? 2 : 1][KR + 2][PW];
  __shared__ T lfn[HK ? 2 : 1][KR + 2][PW];
  __shared__ double lred[kWaves];
  if (skip(sc)) return;
  pad_partials(partials, npart);
  const int per = static_cast<int>(gridDim.x >> 3);
  const int t = static_cast<int>(blockIdx.x & 7) * per + static_cast<int>(blockIdx.x >> 3);
  if (t >= ntiles) {
    if (partials && threadIdx.x == 0) partials[blockIdx.x] = 0.0;
    return;
  }
  const int tx = t % tiles_x;
  const int rest = t / tiles_x;
  const int ty = rest % tiles_y;
  const int cz = rest / tiles_y;
  const int lane = threadIdx.x & (kLanes - 1);
  const int wv = threadIdx.x / kLanes;
  const int i = tx * TX - 1 + lane;
  const int k0 = cz * kchunk;
  const int k1 = min(k0 + kchunk, op.nz);
  const int nz = op.nz;
  const long long sxy = op.sxy;
  const int pbytes = static_cast<int>(sxy * static_cast<long long>(sizeof(S)));
  const int kbytes = static_cast<int>(sxy * static_cast<long long>(sizeof(T)));
  const T fx = T(op.fx), fy = T(op.fy), fz = T(op.fz);
  const bool in_x = i >= 0 && i < op.nx;
  const bool bnd_x = i == 0 || i == op.nx - 1;
  const bool tile_x = lane >= 1 && lane <= kLanes - 2;

  unsigned m_out = 0, m_bxy = 0, m_fxy = 0;
  unsigned off_s[R], off_k[R], off_st[R];
  const int ci = min(max(i, 0), op.nx - 1);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int lr = wv * R + r;
    const int j = ty * TY - 1 + lr;
    const bool in_y = j >= 0 && j < op.ny;
    const unsigned col = static_cast<unsigned>(min(max(j, 0), op.ny - 1) * op.nx + ci);
    off_s[r] = col * static_cast<unsigned>(sizeof(S));
    off_k[r] = col * static_cast<unsigned>(sizeof(T));
    const bool out = in_x && in_y && tile_x && lr >= 1 && lr <= TYE - 2;
    off_st[r] = out ? off_s[r] : kDropOffset;
    if (out) m_out |= 1u << r;
    if (bnd_x || j == 0 || j == op.ny - 1) m_bxy |= 1u << r;
    if (i == 1 || i == op.nx - 2 || j == 1 || j == op.ny - 2) m_fxy |= 1u << r;
  }
  const int jrow0 = ty * TY - 1 + wv * R;

  auto kcl = [&](int k) -> long long { return static_cast<long long>(min(max(k, 0), nz - 1)) * sxy; };
  auto ld_x = [&](int r, int k) -> T { return BufIO<S>::ld(plane_rsrc(xv, kcl(k), sizeof(S), pbytes), off_s[r]); };
  auto ld_k = [&](int r, int k) -> T {
    if constexpr (HK) return BufIO<T>::ld(plane_rsrc(op.kappa, kcl(k), sizeof(T), kbytes), off_k[r]);
    else return T(0);
  };
  auto ld_b = [&](int r, int k) -> T {
    if constexpr (MODE == 1) return BufIO<S>::ld(plane_rsrc(bv, kcl(k), sizeof(S), pbytes), off_s[r]);
    else return T(0);
  };
  // MODE 3: z into X, p into XQ; beta
  auto ld_z = [&](int r, int k) -> T {
    if constexpr (XP) return BufIO<S>::ld(plane_rsrc(bv, kcl(k), sizeof(S), pbytes), off_s[r]);
    else return T(0);
  };
  const T beta = XP ? T(sc->beta) : T(0);

  // Rings, plane pstart + s at slot s mod L: X (6) x(p-1 .. p+4), B (3) b(p .. p+2),
  // K (3) kappa(p+1 .. p+3), ZF (3) zf(p-1 .. p+1) with zf(k) = face(kappa(k), kappa(k+1)).
  // MODE 3: X holds z until the step that first uses the plane turns it into p' with XQ's p.
  const int pstart = HK ? k0 - 2 : k0;
  T X[R][6], B[R][3], K[R][3], ZF[R][3], XQ[R][6];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if constexpr (XP) {
#pragma unroll
      for (int s = 0; s < 6; ++s) XQ[r][s] = T(0);
      X[r][5] = fma(beta, ld_x(r, pstart - 1), ld_z(r, pstart - 1));
      X[r][0] = fma(beta, ld_x(r, pstart), ld_z(r, pstart));
      X[r][1] = ld_z(r, pstart + 1);
      X[r][2] = ld_z(r, pstart + 2);
      X[r][3] = ld_z(r, pstart + 3);
      XQ[r][1] = ld_x(r, pstart + 1);
      XQ[r][2] = ld_x(r, pstart + 2);
      XQ[r][3] = ld_x(r, pstart + 3);
    } else {
      X[r][5] = ld_x(r, pstart - 1);
      X[r][0] = ld_x(r, pstart);
      X[r][1] = ld_x(r, pstart + 1);
      X[r][2] = ld_x(r, pstart + 2);
      X[r][3] = ld_x(r, pstart + 3);
    }
    X[r][4] = T(0);
    B[r][0] = ld_b(r, pstart);
    B[r][1] = ld_b(r, pstart + 1);
    B[r][2] = T(0);
    K[r][1] = ld_k(r, pstart + 1);
    K[r][2] = ld_k(r, pstart + 2);
    K[r][0] = T(0);
    ZF[r][0] = ZF[r][1] = ZF[r][2] = T(0);
  }
  double acc = 0.0;

  auto step = [&](auto phc, int p) {
    constexpr int PH = decltype(phc)::value;
    constexpr int x_m1 = (PH + 5) % 6, x_0 = PH % 6, x_1 = (PH + 1) % 6, x_4 = (PH + 4) % 6;
    constexpr int b_0 = PH % 3, b_2 = (PH + 2) % 3;
    constexpr int k_1 = (PH + 1) % 3, k_2 = (PH + 2) % 3, k_3 = PH % 3;
    constexpr int z_m1 = (PH + 2) % 3, z_0 = PH % 3, z_1 = (PH + 1) % 3;
    constexpr int sl = PH & 1, fs = (PH + 1) & 1;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int lr = wv * R + r;
      if constexpr (XP) {
        X[r][x_1] = fma(beta, XQ[r][x_1], X[r][x_1]);   // p' of plane p+1, first used below
        X[r][x_4] = ld_z(r, p + 4);
        XQ[r][x_4] = ld_x(r, p + 4);
      } else {
        X[r][x_4] = ld_x(r, p + 4);
      }
      B[r][b_2] = ld_b(r, p + 2);
      if constexpr (HK) {
        K[r][k_3] = ld_k(r, p + 3);
        lk[sl][lr + 1][lane + 1] = K[r][k_1];
      }
      lx[sl][lr + 1][lane + 1] = X[r][x_0];
    }
    __syncthreads();
    if constexpr (HK) {
      // faces of plane p+1 into slot (p+1) & 1
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int lr = wv * R + r;
        const T kc = K[r][k_1];
        ZF[r][z_1] = face_op<P, T>(kc, K[r][k_2], fz);
        const T fe = face_op<P, T>(kc, lk[sl][lr + 1][lane + 2], fx);
        lfe[fs][lr + 1][lane + 1] = lane < kLanes - 1 ? fe : T(0);
        lfn[fs][lr + 1][lane + 1] = lr < TYE - 1 ? face_op<P, T>(kc, lk[sl][lr + 2][lane + 1], fy) : T(0);
      }
    }
    if (p >= k0) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int lr = wv * R + r;
        const int j = jrow0 + r;
        const T xc = X[r][x_0];
        W7<T> w;
        if constexpr (KIND == PDEHIP_OP_CONST) {
          w.C = T(op.w[0]); w.W = T(op.w[1]); w.E = T(op.w[2]); w.S = T(op.w[3]);
          w.N = T(op.w[4]); w.B = T(op.w[5]); w.F = T(op.w[6]);
        } else {
          w.W = lfe[sl][lr + 1][lane]; w.E = lfe[sl][lr + 1][lane + 1];
          w.S = lfn[sl][lr][lane + 1]; w.N = lfn[sl][lr + 1][lane + 1];
          T sum = w.W + w.E + w.S + w.N;
          w.B = ZF[r][z_m1]; w.F = ZF[r][z_0];
          sum += w.B + w.F;
          w.C = T(op.e) - sum;
        }
        if (((m_fxy >> r) & 1u) || p == 1 || p == nz - 2) fold_center<3, T>(op, i, j, p, w);
        const T ax = apply_row<T>(w, xc, lx[sl][lr + 1][lane + 2], lx[sl][lr + 1][lane],
                                  lx[sl][lr + 2][lane + 1], lx[sl][lr][lane + 1], X[r][x_1],
                                  X[r][x_m1]);
        const bool bnd = ((m_bxy >> r) & 1u) || p == 0 || p == nz - 1;
        const T ain = bnd ? xc : ax;
        const T out = MODE == 1 ? T(B[r][b_0] - ain) : ain;
        const S so = store_as<S>(static_cast<double>(out));
        BufIO<S>::st(plane_rsrc(yv, static_cast<long long>(p) * sxy, sizeof(S), pbytes), off_st[r], so);
        if constexpr (XP)
          BufIO<S>::st(plane_rsrc(pov, static_cast<long long>(p) * sxy, sizeof(S), pbytes), off_st[r], xc);
        if (partials && ((m_out >> r) & 1u))
          acc += MODE != 1 ? static_cast<double>(xc) * static_cast<double>(out)
                           : static_cast<double>(out) * static_cast<double>(out);
      }
    }
  };

  const int steps = k1 - pstart;
  for (int s0 = 0; s0 < steps; s0 += kRingP) {
    const int p = pstart + s0;
    step(std::integral_constant<int, 0>{}, p);
    if (s0 + 1 < steps) step(std::integral_constant<int, 1>{}, p + 1);
    if (s0 + 2 < steps) step(std::integral_constant<int, 2>{}, p + 2);
    if (s0 + 3 < steps) step(std::integral_constant<int, 3>{}, p + 3);
    if (s0 + 4 < steps) step(std::integral_constant<int, 4>{}, p + 4);
    if (s0 + 5 < steps) step(std::integral_constant<int, 5>{}, p + 5);
  }
  if (partials) {
    const double s = block_sum(acc, lred);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
  }
}

struct Shape { int tiles_x, tiles_y, chunks, ntiles, kchunk, blocks; };

// Rows per wave of a plane march: 2 (tile 62 x 14), except the 4-byte residual: 4 (tile 62 x 30,
// less halo per output node and one barrier per 4 rows; measured faster in a mixed solve at
// 257^3, where the 4-row apply and fused direction update are slower than at 2 rows: at their
// register count two workgroups fit a CU against three at 2 rows).  The FP64 tile's LDS planes
// would not leave room for two workgroups per CU at 4 rows.  Ring and generic kernels of one
// precision and mode share the shape, so their partials are the same, and so do the apply and
// the fused direction update.
template <Prec P, int MODE> constexpr int plane_rows() { return P != P64 && MODE == 1 ? 4 : 2; }

// Launch shape of an R-row plane march.  Occupancy is that of the heaviest instantiation of
// the row count (FP64 FLUX residual at 2 rows, FP32 FLUX ring residual at 4), so every
// instantiation with the same R has the same shape.  The larger of the two shapes' workgroup
// counts is the partials count of every plane march on the grid (plane_blocks); the smaller
// launch zeroes the partials past its grid.
Shape shape_of(const DevOp& op, int R) {
  const int ty = R * kWaves - 2;
  Shape s;
  s.tiles_x = (op.nx + kLanes - 3) / (kLanes - 2);
  s.tiles_y = (op.ny + ty - 1) / ty;
  static const int occ2 = blocks_per_cu(k_plane3<PDEHIP_OP_FLUX, P64, 1, 2>, kThreads);
  static const int occ4 = blocks_per_cu(k_plane3r<PDEHIP_OP_FLUX, P32, 1, 4>, kThreads);
  const long long txy = (long long)s.tiles_x * s.tiles_y;
  s.kchunk = plane_chunk(txy, op.nz, device_cus() * (R == 2 ? occ2 : occ4), 2, 4);
  // at most kMaxPartials workgroups (the reduction partials' capacity)
  while (s.kchunk < op.nz && txy * ((op.nz + s.kchunk - 1) / s.kchunk) > kMaxPartials - 8) s.kchunk *= 2;
  s.chunks = (op.nz + s.kchunk - 1) / s.kchunk;
  s.ntiles = s.tiles_x * s.tiles_y * s.chunks;
  s.blocks = 8 * ((s.ntiles + 7) / 8);
  return s;
}

bool g_plane_ring = true;

template <int KIND, Prec P, int MODE>
void launch(const Launch& L, const DevOp& op, const void* b, const void* x, void* y,
            double* partials) {
  constexpr int R = plane_rows<P, MODE>();
  const Shape s = shape_of(op, R);
  const int npart = plane_blocks(op);
  // ring variant: 4-byte arithmetic, buffer offsets of a plane below 2^31 bytes
  if constexpr (sizeof(typename PT<P>::T) == 4 && KIND != PDEHIP_OP_POINTWISE) {
    if (g_plane_ring && op.sxy < (1LL << 29)) {
      hipLaunchKernelGGL((k_plane3r<KIND, P, MODE, R>), dim3(s.blocks), dim3(kThreads), 0,
                         L.stream, op, b, x, y, partials, npart, s.tiles_x, s.tiles_y, s.ntiles,
                         s.kchunk, L.sc, nullptr);
      return;
    }
  }
  hipLaunchKernelGGL((k_plane3<KIND, P, MODE, R>), dim3(s.blocks), dim3(kThreads), 0, L.stream,
                     op, b, x, y, partials, npart, s.tiles_x, s.tiles_y, s.ntiles, s.kchunk, L.sc,
                     ReplArgs{});
}

template <int MODE>
void launch_mode(const Launch& L, const DevOp& op, Prec p, const void* b, const void* x, void* y,
                 double* partials) {
  switch (op.kind) {
    case PDEHIP_OP_CONST:
      if (p == P64) launch<PDEHIP_OP_CONST, P64, MODE>(L, op, b, x, y, partials);
      else if (p == P32) launch<PDEHIP_OP_CONST, P32, MODE>(L, op, b, x, y, partials);
      else launch<PDEHIP_OP_CONST, P16, MODE>(L, op, b, x, y, partials);
      break;
    case PDEHIP_OP_POINTWISE:
      launch<PDEHIP_OP_POINTWISE, P64, MODE>(L, op, b, x, y, partials);
      break;
    default:
      if (p == P64) launch<PDEHIP_OP_FLUX, P64, MODE>(L, op, b, x, y, partials);
      else if (p == P32) launch<PDEHIP_OP_FLUX, P32, MODE>(L, op, b, x, y, partials);
      else launch<PDEHIP_OP_FLUX, P16, MODE>(L, op, b, x, y, partials);
  }
}

}  // namespace

int plane_blocks(const DevOp& op) { return std::max(shape_of(op, 2).blocks, shape_of(op, 4).blocks); }

void set_plane_ring(bool on) { g_plane_ring = on; }
bool plane_ring() { return g_plane_ring; }

void apply_plane(const Launch& L, const DevOp& op, Prec p, const void* x, void* y, double* partials) {
  launch_mode<0>(L, op, p, nullptr, x, y, partials);
}

void residual_plane(const Launch& L, const DevOp& op, Prec p, const void* b, const void* x, void* r,
                    double* partials) {
  launch_mode<1>(L, op, p, b, x, r, partials);
}

bool apply_xpay_plane(const Launch& L, const DevOp& op, Prec p, const void* z, const void* p_in,
                      void* p_out, void* y, double* partials) {
  if (op.nz <= 1 || !plane_stencils() || !g_plane_ring || p != P32 || op.sxy >= (1LL << 29) ||
      op.kind == PDEHIP_OP_POINTWISE || p_in == p_out)
    return false;
  constexpr int R = plane_rows<P32, 3>();
  const Shape s = shape_of(op, R);
  const int npart = plane_blocks(op);
  if (op.kind == PDEHIP_OP_CONST)
    hipLaunchKernelGGL((k_plane3r<PDEHIP_OP_CONST, P32, 3, R>), dim3(s.blocks), dim3(kThreads), 0,
                       L.stream, op, z, p_in, y, partials, npart, s.tiles_x, s.tiles_y, s.ntiles,
                       s.kchunk, L.sc, p_out);
  else
    hipLaunchKernelGGL((k_plane3r<PDEHIP_OP_FLUX, P32, 3, R>), dim3(s.blocks), dim3(kThreads), 0,
                       L.stream, op, z, p_in, y, partials, npart, s.tiles_x, s.tiles_y, s.ntiles,
                       s.kchunk, L.sc, p_out);
  return true;
}

bool resnorm_plane(hipStream_t s, const DevOp& op, const double* u, const double* f, double f_const,
                   double* sums, double* maxs, int* degenerate) {
  if (op.nz < 3 || op.nx < 3 || op.ny < 3 || !plane_stencils()) return false;
  const Shape sh = shape_of(op, plane_rows<P64, 4>());
  const int npart = resnorm_blocks(op);
  ReplArgs ra{};
  ra.f = f;
  ra.f_const = f_const;
  ra.maxs = maxs;
  ra.degenerate = degenerate;
#define PDEHIP_RN_(K)                                                                           \
  hipLaunchKernelGGL((k_plane3<K, P64, 4, plane_rows<P64, 4>()>), dim3(sh.blocks), dim3(kThreads), \
                     0, s, op, nullptr, u, nullptr, sums, npart, sh.tiles_x, sh.tiles_y,         \
                     sh.ntiles, sh.kchunk, nullptr, ra)
  switch (op.kind) {
    case PDEHIP_OP_CONST: PDEHIP_RN_(PDEHIP_OP_CONST); break;
    case PDEHIP_OP_POINTWISE: PDEHIP_RN_(PDEHIP_OP_POINTWISE); break;
    default: PDEHIP_RN_(PDEHIP_OP_FLUX); break;
  }
#undef PDEHIP_RN_
  return true;
}

bool replace_plane(const Launch& Lgate, const DevOp& op64, const double* b, double* const x[2],
                   const float* e, double* r64, float* r32, const KScalars* so, double* partials) {
  if (op64.nz <= 1 || !plane_stencils()) return false;
  const Shape s = shape_of(op64, plane_rows<P64, 2>());
  const int npart = plane_blocks(op64);
  ReplArgs ra{};
  ra.x[0] = x[0];
  ra.x[1] = x[1];
  ra.e = e;
  ra.r32 = r32;
  ra.so = so;
#define PDEHIP_REPL_(K)                                                                         \
  hipLaunchKernelGGL((k_plane3<K, P64, 2, plane_rows<P64, 2>()>), dim3(s.blocks), dim3(kThreads), \
                     0, Lgate.stream, op64, b, nullptr, r64, partials, npart, s.tiles_x,         \
                     s.tiles_y, s.ntiles, s.kchunk, Lgate.sc, ra)
  switch (op64.kind) {
    case PDEHIP_OP_CONST: PDEHIP_REPL_(PDEHIP_OP_CONST); break;
    case PDEHIP_OP_POINTWISE: PDEHIP_REPL_(PDEHIP_OP_POINTWISE); break;
    default: PDEHIP_REPL_(PDEHIP_OP_FLUX); break;
  }
#undef PDEHIP_REPL_
  return true;
}

}  // namespace pdehip
