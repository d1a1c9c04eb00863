// k_box.hip — multigrid smoothing of the coarse 3D levels on boxes for gfx950: one or two
// red-black Gauss-Seidel sweeps per launch, temporally blocked in LDS.
//
// The plane-march sweep (k_smooth.hip) is built for large grids: a workgroup walks a z chunk one
// plane at a time with one barrier per plane, so a launch costs its prologue and a few plane
// steps of latency whatever the grid's size.  On the coarse levels of a V-cycle (129^3 and
// below) the level sits in L2 / Infinity Cache and that latency is the whole cost.  Here a
// workgroup owns a BX x BY x BZ box of interior nodes (16 x 8, or 8 x 8 / 8 x 4 on the small
// levels, where more workgroups with less serial work each cut the latency) and loads the box
// plus a ring of H = 2 * sweeps nodes into LDS at once (every load of the box in flight together), then runs
// the colour stages of its sweeps in LDS with one barrier per stage:
//   stage s (0 .. 2*sweeps-1) updates colour c1 (s even) or the other colour (s odd) on the box
//   plus a ring of 2*sweeps-1-s nodes; the neighbours it reads sit one node further out and
//   were left by stage s-1, so the box's own nodes come out as sweeps of the whole grid leave
//   them, and no workgroup waits for another.
// Input modes: x_in from memory; x_in known to be zero (no x loads); or x_in plus the
// prolongation of the coarse-level correction, formed on load from a coarse box staged in LDS,
// so the V-cycle's prolongation and the post-smoothing sweeps after it take one launch.
// Every update uses the plane-march sweep's per-node arithmetic (weights_at, fold_center,
// gs_value, face_smooth; values rounded to the storage type after each update) and the
// prolongation uses the prolongation kernel's (transfer_dev.hpp), so x_out equals bit for bit
// what the prolongation launch and the plane-march sweeps leave.
#include <type_traits>

#include "kernels.hpp"
#include "smooth_dev.hpp"
#include "transfer_dev.hpp"

namespace pdehip {
namespace {

constexpr int kBoxThreads = 512;
constexpr int kBX = 16, kBY = 8;   // interior nodes of a box along x and y (z: BZ)

template <int N> using IC = std::integral_constant<int, N>;

template <typename S> __device__ __forceinline__ float back(S v);
template <> __device__ __forceinline__ float back<float>(float v) { return v; }
template <> __device__ __forceinline__ float back<__half>(__half v) { return __half2float(v); }

// IN: 0 x_in from memory, 1 x_in is zero, 2 x_in plus factor * prolong(uc) on interior nodes.
template <int KIND, Prec P, int NS, int BZ, int IN, int BX = kBX, int BY = kBY>
__global__ void __launch_bounds__(kBoxThreads)
k_box_rbgs(DevOp op, const void* __restrict__ bv, const void* __restrict__ xin_v,
           void* __restrict__ xout_v, int c1, int nbx, int nby, int nboxes, DevOp cop,
           const void* __restrict__ uc_v, double factor, const double* dscale,
           const KScalars* sc) {
  using T = typename PT<P>::T;
  using S = typename PT<P>::S;
  static_assert(sizeof(T) == 4, "box sweep: 4-byte arithmetic types only");
  constexpr bool HK = KIND == PDEHIP_OP_FLUX;
  constexpr int H = 2 * NS;
  constexpr int EX = BX + 2 * H, EY = BY + 2 * H, EZ = BZ + 2 * H;
  constexpr int EXY = EX * EY, EN = EXY * EZ;
  // coarse nodes under the extended box (prolongation on load)
  constexpr int CX = EX / 2 + 2, CY = EY / 2 + 2, CZ = EZ / 2 + 2, CN = CX * CY * CZ;
  // x and b in LDS in the storage type (every value the sweeps hold is representable in it),
  // rho = 1/kappa in the arithmetic type
  __shared__ S xs[EN];
  __shared__ S bs[EN];
  __shared__ T rs[HK ? EN : 1];
  __shared__ S cs[IN == 2 ? CN : 1];
  if (skip(sc)) return;
  const int t = xcd_tile();
  if (t >= nboxes) return;
  const int bx = t % nbx, rest = t / nbx, by = rest % nby, bz = rest / nby;
  const int ox = 1 + bx * BX, oy = 1 + by * BY, oz = 1 + bz * BZ;   // first interior node
  const int gx0 = ox - H, gy0 = oy - H, gz0 = oz - H;                  // grid node of LDS (0,0,0)
  const int nx = op.nx, ny = op.ny, nz = op.nz;
  const S* __restrict__ b = static_cast<const S*>(bv);
  const S* __restrict__ xin = static_cast<const S*>(xin_v);
  const T* __restrict__ kap = static_cast<const T*>(op.kappa);
  S* __restrict__ xout = static_cast<S*>(xout_v);
  const int tid = static_cast<int>(threadIdx.x);
  auto gid = [&](int i, int j, int k) -> long long { return ((long long)k * ny + j) * nx + i; };

  const int cx0 = gx0 >> 1, cy0 = gy0 >> 1, cz0 = gz0 >> 1;   // floor
  // the extended box (and the coarse box): every load issued before the first LDS store
  constexpr int NC = IN == 2 ? (CN + kBoxThreads - 1) / kBoxThreads : 1;
  S cv[NC];
  bool inc[NC];
  if constexpr (IN == 2) {
    const S* __restrict__ uc = static_cast<const S*>(uc_v);
#pragma unroll
    for (int q = 0; q < NC; ++q) {
      const int e = min(tid + q * kBoxThreads, CN - 1);
      const int ci = cx0 + e % CX, cj = cy0 + (e / CX) % CY, ck = cz0 + e / (CX * CY);
      inc[q] = ci >= 0 && ci < cop.nx && cj >= 0 && cj < cop.ny && ck >= 0 && ck < cop.nz;
      cv[q] = uc[inc[q] ? ((long long)ck * cop.ny + cj) * cop.nx + ci : 0];
    }
  }
  constexpr int NL = (EN + kBoxThreads - 1) / kBoxThreads;
  S xv[NL], bv_[NL];
  T rv[HK ? NL : 1];
  bool inb[NL];
#pragma unroll
  for (int q = 0; q < NL; ++q) {
    const int e = min(tid + q * kBoxThreads, EN - 1);
    const int gi = gx0 + e % EX, gj = gy0 + (e / EX) % EY, gk = gz0 + e / EXY;
    inb[q] = gi >= 0 && gi < nx && gj >= 0 && gj < ny && gk >= 0 && gk < nz;
    const long long id = inb[q] ? gid(gi, gj, gk) : 0;
    if constexpr (IN != 1) xv[q] = xin[id];
    bv_[q] = b[id];
    if constexpr (HK) rv[q] = kap[id];
  }
#pragma unroll
  for (int q = 0; q < NL; ++q) {
    const int e = tid + q * kBoxThreads;
    if (EN % kBoxThreads == 0 || e < EN) {
      if constexpr (IN != 1) xs[e] = inb[q] ? xv[q] : S(0.0f);
      else xs[e] = S(0.0f);
      bs[e] = inb[q] ? bv_[q] : S(0.0f);
      if constexpr (HK) rs[e] = inb[q] ? rv[q] : T(0);
    }
  }
  if constexpr (IN == 2) {
#pragma unroll
    for (int q = 0; q < NC; ++q) {
      const int e = tid + q * kBoxThreads;
      if (CN % kBoxThreads == 0 || e < CN) cs[e] = inc[q] ? cv[q] : S(0.0f);
    }
  }
  if constexpr (IN == 2) {
    __syncthreads();
    const double s = dscale ? factor * (*dscale) : factor;
    auto C = [&](int k, int j, int i) -> double {
      return static_cast<double>(back<S>(cs[((k - cz0) * CY + (j - cy0)) * CX + (i - cx0)]));
    };
    // values first, stores after (each node reads only the coarse box and itself)
    S pv[NL];
    bool upd[NL];
#pragma unroll
    for (int q = 0; q < NL; ++q) {
      const int e = min(tid + q * kBoxThreads, EN - 1);
      const int gi = gx0 + e % EX, gj = gy0 + (e / EX) % EY, gk = gz0 + e / EXY;
      upd[q] = (EN % kBoxThreads == 0 || tid + q * kBoxThreads < EN) && gi >= 1 && gi <= nx - 2 &&
               gj >= 1 && gj <= ny - 2 && gk >= 1 && gk <= nz - 2;
      if (!upd[q]) continue;
      const ProlongCol w = prolong_col(cop, gi, gj);
      const int kc = gk >> 1;
      const bool half = (gk & 1) != 0;
      double wz0, wz1 = 0.0;
      const int k0 = prolong_zplane(cop, kc, &wz0);
      const double p_lo = prolong_plane(w, C(k0, w.y0, w.x0), C(k0, w.y0, w.x1),
                                        C(k0, w.y1, w.x0), C(k0, w.y1, w.x1), wz0);
      double p_hi = 0.0;
      if (half) {
        const int k1 = prolong_zplane(cop, kc + 1, &wz1);
        p_hi = prolong_plane(w, C(k1, w.y0, w.x0), C(k1, w.y0, w.x1), C(k1, w.y1, w.x0),
                             C(k1, w.y1, w.x1), wz1);
      }
      const double fa = static_cast<double>(back<S>(xs[e]));
      pv[q] = store_as<S>(prolong_fine(fa, s, half, p_lo, p_hi));
    }
#pragma unroll
    for (int q = 0; q < NL; ++q)
      if (upd[q]) xs[tid + q * kBoxThreads] = pv[q];
  }
  __syncthreads();

  const T fx = T(op.fx), fy = T(op.fy), fz = T(op.fz);
  const int c2 = 1 - c1;
  auto stage = [&](auto stc) {
    constexpr int st = decltype(stc)::value;
    constexpr int r = 2 * NS - 1 - st;                 // ring of this stage
    constexpr int RX = BX + 2 * r, RY = BY + 2 * r, RZ = BZ + 2 * r, HX = RX / 2;
    constexpr int NI = HX * RY * RZ;
    constexpr int o = H - r;                           // LDS offset of the stage's region
    constexpr int NQ = (NI + kBoxThreads - 1) / kBoxThreads;
    const int colour = (st & 1) ? c2 : c1;
    // every update of the stage first (a stage reads only the other colour), stores after
    S nv[NQ];
    int ne[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int it = tid + q * kBoxThreads;
      ne[q] = -1;
      if (NI % kBoxThreads != 0 && it >= NI) continue;
      const int hx = it % HX, row = it / HX;
      const int ey = o + row % RY, ez = o + row / RY;
      const int gj = gy0 + ey, gk = gz0 + ez;
      // the region's nodes of this colour in row (ey, ez): every other x, starting at par
      const int par = (colour + gx0 + o + gj + gk) & 1;
      const int ex = o + 2 * hx + par;
      const int gi = gx0 + ex;
      if (gi < 1 || gi > nx - 2 || gj < 1 || gj > ny - 2 || gk < 1 || gk > nz - 2) continue;
      const int e = (ez * EY + ey) * EX + ex;
      T aw = T(0), ae = T(0), as = T(0), an = T(0), ab = T(0), af = T(0);
      if constexpr (HK) {
        const T rc = rs[e];
        aw = face_smooth<T>(rc, rs[e - 1], fx);
        ae = face_smooth<T>(rc, rs[e + 1], fx);
        as = face_smooth<T>(rc, rs[e - EX], fy);
        an = face_smooth<T>(rc, rs[e + EX], fy);
        ab = face_smooth<T>(rc, rs[e - EXY], fz);
        af = face_smooth<T>(rc, rs[e + EXY], fz);
      }
      W7<T> w = weights_at<3, KIND, T>(op, 0, aw, ae, as, an, ab, af);
      fold_center<3, T>(op, gi, gj, gk, w);
      const T v = gs_value<3, T>(w, back<S>(bs[e]), back<S>(xs[e + 1]), back<S>(xs[e - 1]),
                                 back<S>(xs[e + EX]), back<S>(xs[e - EX]), back<S>(xs[e + EXY]),
                                 back<S>(xs[e - EXY]));
      nv[q] = to_store<S, T>(v);
      ne[q] = e;
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q)
      if (ne[q] >= 0) xs[ne[q]] = nv[q];
    __syncthreads();
  };
  stage(IC<0>{});
  stage(IC<1>{});
  if constexpr (NS == 2) {
    stage(IC<2>{});
    stage(IC<3>{});
  }

  // out: the box's interior nodes, plus the boundary layer beside it on the grid faces it touches
  const int xl = ox == 1 ? 0 : ox, yl = oy == 1 ? 0 : oy, zl = oz == 1 ? 0 : oz;
  const int xh = ox + BX >= nx - 1 ? nx : ox + BX;
  const int yh = oy + BY >= ny - 1 ? ny : oy + BY;
  const int zh = oz + BZ >= nz - 1 ? nz : oz + BZ;
  constexpr int OX = BX + 2, OY = BY + 2, OZ = BZ + 2, ON = OX * OY * OZ;
  for (int it = tid; it < ON; it += kBoxThreads) {
    const int gi = ox - 1 + it % OX, gj = oy - 1 + (it / OX) % OY, gk = oz - 1 + it / (OX * OY);
    if (gi < xl || gi >= xh || gj < yl || gj >= yh || gk < zl || gk >= zh) continue;
    const int e = ((gk - gz0) * EY + (gj - gy0)) * EX + (gi - gx0);
    xout[gid(gi, gj, gk)] = xs[e];
  }
}

// ------------------------------------------------------------------ residual + restriction
// One workgroup per box of CX x CY x CZ coarse nodes (8 x 4 x 4; 4 x 4 x 4 on the small
// levels): the fine residual r = b - A x (the
// plane kernels' per-node arithmetic: faces by face_coef, apply_row, r rounded to the storage
// type) on every fine node the box's full weighting reads, the face nodes its folded closure
// terms read included, kept in LDS; then each coarse node's value (restrict_value,
// transfer_dev.hpp) from there.  The fine residual never goes to memory, and a level's residual
// and restriction take one launch.
constexpr int kCX = 8, kCY = 4, kCZ = 4;
constexpr int kRRThreads = 256;

template <int KIND, Prec PF, Prec PC, int CX = kCX, int CY = kCY, int CZ = kCZ>
__global__ void __launch_bounds__(kRRThreads)
k_box_rr(DevOp f, const void* __restrict__ bv, const void* __restrict__ xv, DevOp c,
         void* __restrict__ bcv, double scale, const double* dscale, int ncx, int ncy,
         int nboxes, const KScalars* sc) {
  using T = typename PT<PF>::T;
  using S = typename PT<PF>::S;
  using SC = typename PT<PC>::S;
  static_assert(sizeof(T) == 4, "box residual-restriction: 4-byte arithmetic types only");
  constexpr bool HK = KIND == PDEHIP_OP_FLUX;
  // fine residual region [2*I0 - 2, 2*(I0 + CX - 1) + 2] per axis, x region one node wider
  constexpr int RX = 2 * CX + 3, RY = 2 * CY + 3, RZ = 2 * CZ + 3, RN = RX * RY * RZ;
  constexpr int XX = RX + 2, XY = RY + 2, XZ = RZ + 2, XXY = XX * XY, XN = XXY * XZ;
  __shared__ S xs[XN];
  __shared__ T rs[HK ? XN : 1];
  __shared__ S bs[RN];
  __shared__ S res[RN];
  if (skip(sc)) return;
  const int t = xcd_tile();
  if (t >= nboxes) return;
  const int I0 = (t % ncx) * CX, rest = t / ncx, J0 = (rest % ncy) * CY, K0 = (rest / ncy) * CZ;
  const int rx0 = 2 * I0 - 2, ry0 = 2 * J0 - 2, rz0 = 2 * K0 - 2;   // residual region origin
  const int nx = f.nx, ny = f.ny, nz = f.nz;
  const S* __restrict__ b = static_cast<const S*>(bv);
  const S* __restrict__ x = static_cast<const S*>(xv);
  const T* __restrict__ kap = static_cast<const T*>(f.kappa);
  SC* __restrict__ bc = static_cast<SC*>(bcv);
  const int tid = static_cast<int>(threadIdx.x);
  auto gid = [&](int i, int j, int k) -> long long { return ((long long)k * ny + j) * nx + i; };
  auto in_grid = [&](int i, int j, int k) {
    return i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz;
  };

  // loads: x and rho on the x region, b on the residual region, every load before any store
  constexpr int NX = (XN + kRRThreads - 1) / kRRThreads;
  constexpr int NB = (RN + kRRThreads - 1) / kRRThreads;
  S xv_[NX], bv_[NB];
  T rv[HK ? NX : 1];
  bool inx[NX], inb[NB];
#pragma unroll
  for (int q = 0; q < NX; ++q) {
    const int e = min(tid + q * kRRThreads, XN - 1);
    const int gi = rx0 - 1 + e % XX, gj = ry0 - 1 + (e / XX) % XY, gk = rz0 - 1 + e / XXY;
    inx[q] = in_grid(gi, gj, gk);
    const long long id = inx[q] ? gid(gi, gj, gk) : 0;
    xv_[q] = x[id];
    if constexpr (HK) rv[q] = kap[id];
  }
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const int e = min(tid + q * kRRThreads, RN - 1);
    const int gi = rx0 + e % RX, gj = ry0 + (e / RX) % RY, gk = rz0 + e / (RX * RY);
    inb[q] = in_grid(gi, gj, gk);
    bv_[q] = b[inb[q] ? gid(gi, gj, gk) : 0];
  }
#pragma unroll
  for (int q = 0; q < NX; ++q) {
    const int e = tid + q * kRRThreads;
    if (XN % kRRThreads == 0 || e < XN) {
      xs[e] = inx[q] ? xv_[q] : S(0.0f);
      if constexpr (HK) rs[e] = inx[q] ? rv[q] : T(0);
    }
  }
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const int e = tid + q * kRRThreads;
    if (RN % kRRThreads == 0 || e < RN) bs[e] = inb[q] ? bv_[q] : S(0.0f);
  }
  __syncthreads();

  // fine residual on the residual region (identity rows on the boundary, zero off the grid)
  const T fx = T(f.fx), fy = T(f.fy), fz = T(f.fz);
  for (int e = tid; e < RN; e += kRRThreads) {
    const int lx = e % RX, ly = (e / RX) % RY, lz = e / (RX * RY);
    const int gi = rx0 + lx, gj = ry0 + ly, gk = rz0 + lz;
    T out = T(0);
    if (in_grid(gi, gj, gk)) {
      const int ex = ((lz + 1) * XY + (ly + 1)) * XX + (lx + 1);
      const T xc = back<S>(xs[ex]);
      const T bb = back<S>(bs[e]);
      if (gi == 0 || gj == 0 || gk == 0 || gi == nx - 1 || gj == ny - 1 || gk == nz - 1) {
        out = T(bb - xc);
      } else {
        T aw = T(0), ae = T(0), as = T(0), an = T(0), ab = T(0), af = T(0);
        if constexpr (HK) {
          const T rc = rs[ex];
          aw = face_coef<T>(rc, rs[ex - 1], fx);
          ae = face_coef<T>(rc, rs[ex + 1], fx);
          as = face_coef<T>(rc, rs[ex - XX], fy);
          an = face_coef<T>(rc, rs[ex + XX], fy);
          ab = face_coef<T>(rc, rs[ex - XXY], fz);
          af = face_coef<T>(rc, rs[ex + XXY], fz);
        }
        W7<T> w = weights_at<3, KIND, T>(f, 0, aw, ae, as, an, ab, af);
        fold_center<3, T>(f, gi, gj, gk, w);
        const T ax = apply_row<T>(w, xc, back<S>(xs[ex + 1]), back<S>(xs[ex - 1]),
                                  back<S>(xs[ex + XX]), back<S>(xs[ex - XX]),
                                  back<S>(xs[ex + XXY]), back<S>(xs[ex - XXY]));
        out = T(bb - ax);
      }
    }
    res[e] = store_as<S>(static_cast<double>(out));
  }
  __syncthreads();

  // the box's coarse nodes
  const double s = dscale ? scale * (*dscale) : scale;
  auto r = [&](int i, int j, int k) -> double {
    return static_cast<double>(back<S>(res[((k - rz0) * RY + (j - ry0)) * RX + (i - rx0)]));
  };
  for (int q = tid; q < CX * CY * CZ; q += kRRThreads) {
    const int ic = I0 + q % CX, jc = J0 + (q / CX) % CY, kc = K0 + q / (CX * CY);
    if (ic >= c.nx || jc >= c.ny || kc >= c.nz) continue;
    const long long out = ((long long)kc * c.ny + jc) * c.nx + ic;
    if (ic == 0 || jc == 0 || kc == 0 || ic == c.nx - 1 || jc == c.ny - 1 || kc == c.nz - 1) {
      bc[out] = store_as<SC>(0.0);
      continue;
    }
    bc[out] = store_as<SC>(restrict_value<true>(f, c, r, ic, jc, kc, s));
  }
}

template <int KIND, Prec PF, Prec PC, int CX, int CY, int CZ>
void launch_rr(const Launch& L, const DevOp& f, const void* b, const void* x, const DevOp& c,
               void* bc, double scale, const double* dscale) {
  const int ncx = (c.nx + CX - 1) / CX, ncy = (c.ny + CY - 1) / CY, ncz = (c.nz + CZ - 1) / CZ;
  const int nboxes = ncx * ncy * ncz;
  const dim3 g(static_cast<unsigned>(8 * ((nboxes + 7) / 8)));
  hipLaunchKernelGGL((k_box_rr<KIND, PF, PC, CX, CY, CZ>), g, dim3(kRRThreads), 0, L.stream, f, b,
                     x, c, bc, scale, dscale, ncx, ncy, nboxes, L.sc);
}

// small: the level is within the small-box bound (coarse boxes of 4 x 4 x 4 nodes)
template <int KIND, Prec PF, Prec PC>
void rr_by_size(const Launch& L, const DevOp& f, const void* b, const void* x, const DevOp& c,
                void* bc, double scale, const double* dscale, bool small) {
  if (small) launch_rr<KIND, PF, PC, 4, 4, 4>(L, f, b, x, c, bc, scale, dscale);
  else launch_rr<KIND, PF, PC, kCX, kCY, kCZ>(L, f, b, x, c, bc, scale, dscale);
}

template <int KIND>
bool rr_by_prec(const Launch& L, const DevOp& f, Prec pf, const void* b, const void* x,
                const DevOp& c, Prec pc, void* bc, double scale, const double* dscale, bool small) {
  if (pf == P32 && pc == P32) rr_by_size<KIND, P32, P32>(L, f, b, x, c, bc, scale, dscale, small);
  else if (pf == P32 && pc == P16) rr_by_size<KIND, P32, P16>(L, f, b, x, c, bc, scale, dscale, small);
  else if (pf == P16 && pc == P16) rr_by_size<KIND, P16, P16>(L, f, b, x, c, bc, scale, dscale, small);
  else return false;
  return true;
}

template <int KIND, Prec P, int NS, int BZ, int IN, int BX, int BY>
void launch(const Launch& L, const DevOp& op, const void* b, const void* xin, void* xout, int c1,
            const DevOp& cop, const void* uc, double factor, const double* dscale) {
  const int nbx = (op.nx - 2 + BX - 1) / BX;
  const int nby = (op.ny - 2 + BY - 1) / BY;
  const int nbz = (op.nz - 2 + BZ - 1) / BZ;
  const int nboxes = nbx * nby * nbz;
  const dim3 g(static_cast<unsigned>(8 * ((nboxes + 7) / 8)));
  hipLaunchKernelGGL((k_box_rbgs<KIND, P, NS, BZ, IN, BX, BY>), g, dim3(kBoxThreads), 0, L.stream,
                     op, b, xin, xout, c1, nbx, nby, nboxes, cop, uc, factor, dscale, L.sc);
}

template <int KIND, Prec P, int NS, int BZ, int BX = kBX, int BY = kBY>
void by_input(const Launch& L, const DevOp& op, const void* b, const void* xin, void* xout, int c1,
              int in_mode, const DevOp& cop, const void* uc, double factor, const double* dscale) {
  if (in_mode == 1) launch<KIND, P, NS, BZ, 1, BX, BY>(L, op, b, xin, xout, c1, cop, uc, factor, dscale);
  else if (in_mode == 2) launch<KIND, P, NS, BZ, 2, BX, BY>(L, op, b, xin, xout, c1, cop, uc, factor, dscale);
  else launch<KIND, P, NS, BZ, 0, BX, BY>(L, op, b, xin, xout, c1, cop, uc, factor, dscale);
}

// small: 0, or the x extent of the small boxes used on this level (8: 8 x 8 x 4, 4: 8 x 4 x 4;
// 4-plane boxes only)
template <int KIND, Prec P>
void by_shape(const Launch& L, const DevOp& op, const void* b, const void* xin, void* xout, int c1,
              int sweeps, int bz, int in_mode, const DevOp& cop, const void* uc, double factor,
              const double* dscale, int small) {
  if (bz == 4 && small == 8) {
    if (sweeps == 2) by_input<KIND, P, 2, 4, 8, 8>(L, op, b, xin, xout, c1, in_mode, cop, uc, factor, dscale);
    else by_input<KIND, P, 1, 4, 8, 8>(L, op, b, xin, xout, c1, in_mode, cop, uc, factor, dscale);
  } else if (bz == 4 && small == 4) {
    if (sweeps == 2) by_input<KIND, P, 2, 4, 8, 4>(L, op, b, xin, xout, c1, in_mode, cop, uc, factor, dscale);
    else by_input<KIND, P, 1, 4, 8, 4>(L, op, b, xin, xout, c1, in_mode, cop, uc, factor, dscale);
  } else if (sweeps == 2) {
    if (bz == 4) by_input<KIND, P, 2, 4>(L, op, b, xin, xout, c1, in_mode, cop, uc, factor, dscale);
    else by_input<KIND, P, 2, 8>(L, op, b, xin, xout, c1, in_mode, cop, uc, factor, dscale);
  } else {
    if (bz == 4) by_input<KIND, P, 1, 4>(L, op, b, xin, xout, c1, in_mode, cop, uc, factor, dscale);
    else by_input<KIND, P, 1, 8>(L, op, b, xin, xout, c1, in_mode, cop, uc, factor, dscale);
  }
}

// the level's sides are all within the small-box bound
bool small_level(const DevOp& op) {
  const int lim = smooth_tuning().box_small;
  return lim > 0 && op.nx <= lim && op.ny <= lim && op.nz <= lim;
}

}  // namespace

bool box_resrestrict(const Launch& L, const DevOp& fine, Prec pf, const void* b, const void* x,
                     const DevOp& coarse, Prec pc, void* bc, double scale, const double* dev_scale) {
  if (fine.nz < 3 || coarse.nz < 3 || fine.n >= (1LL << 31)) return false;
  if (2 * coarse.nx - 1 != fine.nx || 2 * coarse.ny - 1 != fine.ny || 2 * coarse.nz - 1 != fine.nz)
    return false;
  const bool small = small_level(fine);
  if (fine.kind == PDEHIP_OP_CONST)
    return rr_by_prec<PDEHIP_OP_CONST>(L, fine, pf, b, x, coarse, pc, bc, scale, dev_scale, small);
  if (fine.kind == PDEHIP_OP_FLUX)
    return rr_by_prec<PDEHIP_OP_FLUX>(L, fine, pf, b, x, coarse, pc, bc, scale, dev_scale, small);
  return false;
}

bool box_sweeps(const Launch& L, const DevOp& op, Prec p, const void* b, const void* x_in,
                void* x_out, int first_color, int sweeps, int in_mode, const DevOp* coarse,
                Prec pc, const void* uc, double factor, const double* dev_scale) {
  if (op.nz < 3 || op.nx < 3 || op.ny < 3 || sweeps < 1 || sweeps > 2 || p == P64) return false;
  if (op.kind != PDEHIP_OP_CONST && op.kind != PDEHIP_OP_FLUX) return false;
  if (op.n >= (1LL << 31)) return false;
  if (in_mode == 2 && (!coarse || pc != p || coarse->nz < 3)) return false;
  if (in_mode != 1 && x_in == x_out) return false;   // other boxes read x_in's ring
  const DevOp cop = coarse ? *coarse : op;
  const int bz = smooth_tuning().box_z == 4 ? 4 : 8;
  const int small = small_level(op) ? smooth_tuning().box_sxy : 0;
  if (op.kind == PDEHIP_OP_CONST) {
    if (p == P32) by_shape<PDEHIP_OP_CONST, P32>(L, op, b, x_in, x_out, first_color, sweeps, bz, in_mode, cop, uc, factor, dev_scale, small);
    else by_shape<PDEHIP_OP_CONST, P16>(L, op, b, x_in, x_out, first_color, sweeps, bz, in_mode, cop, uc, factor, dev_scale, small);
  } else {
    if (p == P32) by_shape<PDEHIP_OP_FLUX, P32>(L, op, b, x_in, x_out, first_color, sweeps, bz, in_mode, cop, uc, factor, dev_scale, small);
    else by_shape<PDEHIP_OP_FLUX, P16>(L, op, b, x_in, x_out, first_color, sweeps, bz, in_mode, cop, uc, factor, dev_scale, small);
  }
  return true;
}

}  // namespace pdehip
