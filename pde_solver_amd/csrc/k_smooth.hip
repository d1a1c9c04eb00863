// k_smooth.hip — fused red-black Gauss-Seidel sweep for gfx950: both colours of one sweep in
// a single pass over HBM, out of place (x_in -> x_out).
//
// The two-launch in-place sweep (k_rbgs) touches every cache line of x, b and kappa once per
// colour, so one sweep moves about twice the compulsory bytes.  Here one workgroup owns an
// (i,j) tile and marches it through a chunk of z planes:
//
//   * extended tile: 64 columns (one wavefront, lane = x) by R*8 rows (8 waves x R rows per
//     thread); the output tile is the extended tile minus a 2-node ring (60 x (8R-4));
//   * for plane p the first colour is updated on the tile plus a 1-node ring from old values
//     of the second colour (planes p-1, p, p+1), then the second colour of plane p-1 is
//     updated on the tile from the new first-colour values (planes p-2, p-1, p) — the ring
//     makes the tile self-contained, so no workgroup waits for another;
//   * z neighbours of a node live in the owning thread's registers (4-plane queue of x and of
//     kappa, 2 planes of b); in-plane neighbours come from two LDS plane slots;
//   * x_out receives every node of the output tile (boundary nodes keep x_in).
//
// Every node's update uses the same weights, the same FMA chain (gs_value) and the same rounding
// to the storage type as the in-place two-colour sweep, so x_out equals bit for bit what
// rbgs(c1); rbgs(1-c1) would leave in x.  HBM traffic per sweep is the compulsory
// (3 s + s_k) N; the ring and the two extra planes per z chunk are re-reads served by L2 /
// Infinity Cache.  Workgroups are numbered so that each XCD (round-robin dispatch) receives a
// contiguous run of tiles, which keeps ring re-reads inside one XCD's L2.
#include <type_traits>

#include "bufio.hpp"
#include "kernels.hpp"
#include "smooth_dev.hpp"

namespace pdehip {
namespace {

constexpr int kLanes = 64;
constexpr int kHalo = 2;
constexpr int kTileX = kLanes - 2 * kHalo;
constexpr int kWaves = 8;
constexpr int kThreads = kLanes * kWaves;

__device__ __forceinline__ int mod3(int p) { return ((p % 3) + 3) % 3; }

// ------------------------------------------------------------------------------- 3D
// Plane step p (p runs from the prologue start to k1):
//   1. queues advance; loads for x(p+2), b(p+2), kappa(p+4) are issued (consumed two steps
//      later, so their latency overlaps two steps of arithmetic);
//   2. x(p) and kappa(p+1) go to LDS; barrier;
//   3. FLUX: every node of the extended tile computes its east, north and front faces of plane
//      p+1 (LDS face ring; front faces in registers) — each face once;
//   4. first colour on plane p (tile + ring): faces of plane p, z faces from registers;
//   5. second colour on plane p-1 (tile): faces of plane p-1; plane p-1 of the tile is stored.
// ONEBAR (4-byte types): LDS slots x3 for x, x2 for kappa, x4 for faces, so every slot is
// rewritten only after a barrier that all of its readers have passed — one barrier per plane.
// Otherwise (FP64, where those slots would not fit next to enough workgroups): x2 / x1 / x3
// slots and a second barrier at the end of the step.
// Loads are issued unconditionally (out-of-grid lanes and planes read a clamped in-grid node),
// so the compiler counts outstanding loads instead of draining them at every branch.
__device__ __forceinline__ int modn(int p, int n) { return ((p % n) + n) % n; }

template <int KIND, Prec P, int R, bool ZERO>
__global__ void __launch_bounds__(kThreads)
k_rbgs_fused3(DevOp op, const void* __restrict__ bv, const void* __restrict__ xin_v,
              void* __restrict__ xout_v, int c1, int tiles_x, int tiles_y, int ntiles,
              int kchunk, const KScalars* sc) {
  using T = typename PT<P>::T;
  using S = typename PT<P>::S;
  constexpr int TYE = R * kWaves;
  constexpr int TY = TYE - 2 * kHalo;
  constexpr bool HK = KIND == PDEHIP_OP_FLUX;
  constexpr bool ONEBAR = sizeof(T) == 4;
  constexpr int NU = ONEBAR ? 3 : 2, NK = ONEBAR ? 2 : 1, NF = ONEBAR ? 4 : 3;
  constexpr int KR = HK ? TYE : 1;
  __shared__ T lu[NU][TYE][kLanes];
  __shared__ T lk[HK ? NK : 1][KR][kLanes];
  __shared__ T lfe[HK ? NF : 1][KR][kLanes];
  __shared__ T lfn[HK ? NF : 1][KR][kLanes];
  if (skip(sc)) return;
  const int t = xcd_tile();
  if (t >= ntiles) return;
  const S* __restrict__ b = static_cast<const S*>(bv);
  const S* __restrict__ xin = static_cast<const S*>(xin_v);
  const T* __restrict__ kap = static_cast<const T*>(op.kappa);
  S* __restrict__ xout = static_cast<S*>(xout_v);
  const int tx = t % tiles_x;
  const int rest = t / tiles_x;
  const int ty = rest % tiles_y;
  const int cz = rest / tiles_y;
  const int lane = threadIdx.x & (kLanes - 1);
  const int wv = threadIdx.x / kLanes;
  const int i = tx * kTileX - kHalo + lane;
  const int k0 = cz * kchunk;
  const int k1 = min(k0 + kchunk, op.nz);
  const int nz = op.nz;
  const int c2 = 1 - c1;
  const long long sxy = op.sxy;
  const T fx = T(op.fx), fy = T(op.fy), fz = T(op.fz);

  const bool in_x = i >= 0 && i < op.nx;
  const bool int_x = i >= 1 && i <= op.nx - 2;
  const bool ring_x = lane >= 1 && lane <= kLanes - 2;
  const bool tile_x = lane >= kHalo && lane < kLanes - kHalo;

  // per-row flags (bit r): first-colour candidate (ring & interior), second-colour candidate
  // (tile & interior), output (tile & in grid), next to a face (rows that may fold).  Lanes and
  // rows outside the grid load a clamped in-grid node: no update ever reads their values.
  unsigned m_r1 = 0, m_r2 = 0, m_out = 0, m_fxy = 0;
  long long col[R];
  const int ci = min(max(i, 0), op.nx - 1);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int lr = wv * R + r;
    const int j = ty * TY - kHalo + lr;
    const bool in_y = j >= 0 && j < op.ny;
    const bool int_y = j >= 1 && j <= op.ny - 2;
    const bool ring_y = lr >= 1 && lr <= TYE - 2;
    const bool tile_y = lr >= kHalo && lr < TYE - kHalo;
    col[r] = (long long)min(max(j, 0), op.ny - 1) * op.nx + ci;
    if (ring_x && ring_y && int_x && int_y) m_r1 |= 1u << r;
    if (tile_x && tile_y && int_x && int_y) m_r2 |= 1u << r;
    if (tile_x && tile_y && in_x && in_y) m_out |= 1u << r;
    if (i == 1 || i == op.nx - 2 || j == 1 || j == op.ny - 2) m_fxy |= 1u << r;
  }
  const int jrow0 = ty * TY - kHalo + wv * R;

  // plane index clamped into the grid (uniform across the workgroup)
  auto kcl = [&](int k) -> long long { return (long long)min(max(k, 0), nz - 1) * sxy; };
  auto ld_u = [&](int r, int k) -> T {
    if constexpr (ZERO) return T(0);
    else return T(ldS(xin, kcl(k) + col[r]));
  };
  auto ld_k = [&](int r, int k) -> T {
    if constexpr (HK) return kap[kcl(k) + col[r]];
    else return T(0);
  };
  auto ld_b = [&](int r, int k) -> T { return T(ldS(b, kcl(k) + col[r])); };

  // queues after the shift of step p: U = x(p-2..p+2), B = b(p-1..p+2), K = kappa(p+1..p+4),
  // ZF = front faces zf(p-2..p+1) with zf(k) = face(kappa(k), kappa(k+1))
  const int pstart = HK ? k0 - 3 : k0 - 1;
  T U[R][5], B[R][4], K[R][4], ZF[R][4];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    U[r][0] = T(0);
    U[r][1] = ld_u(r, pstart - 2);
    U[r][2] = ld_u(r, pstart - 1);
    U[r][3] = ld_u(r, pstart);
    U[r][4] = ld_u(r, pstart + 1);
    B[r][0] = T(0);
    B[r][1] = ld_b(r, pstart - 1);
    B[r][2] = ld_b(r, pstart);
    B[r][3] = ld_b(r, pstart + 1);
    K[r][0] = T(0);
    K[r][1] = ld_k(r, pstart + 1);
    K[r][2] = ld_k(r, pstart + 2);
    K[r][3] = ld_k(r, pstart + 3);
    ZF[r][0] = ZF[r][1] = ZF[r][2] = ZF[r][3] = T(0);
  }

  for (int p = pstart; p <= k1; ++p) {
    const int su = modn(p, NU);
    const int sk = modn(p, NK);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int lr = wv * R + r;
      U[r][0] = U[r][1]; U[r][1] = U[r][2]; U[r][2] = U[r][3]; U[r][3] = U[r][4];
      U[r][4] = ld_u(r, p + 2);
      B[r][0] = B[r][1]; B[r][1] = B[r][2]; B[r][2] = B[r][3]; B[r][3] = ld_b(r, p + 2);
      if constexpr (HK) {
        K[r][0] = K[r][1]; K[r][1] = K[r][2]; K[r][2] = K[r][3]; K[r][3] = ld_k(r, p + 4);
        ZF[r][0] = ZF[r][1]; ZF[r][1] = ZF[r][2]; ZF[r][2] = ZF[r][3];
        lk[sk][lr][lane] = K[r][0];
      }
      lu[su][lr][lane] = U[r][2];
    }
    __syncthreads();
    if constexpr (HK) {
      const int fs = modn(p + 1, NF);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int lr = wv * R + r;
        const T kc = K[r][0];
        ZF[r][3] = face_smooth<T>(kc, K[r][1], fz);
        lfe[fs][lr][lane] = lane < kLanes - 1 ? face_smooth<T>(kc, lk[sk][lr][lane + 1], fx) : T(0);
        lfn[fs][lr][lane] = lr < TYE - 1 ? face_smooth<T>(kc, lk[sk][lr + 1][lane], fy) : T(0);
      }
    }
    // first colour on plane p (tile + ring)
    if (p >= k0 - 1 && p >= 1 && p <= nz - 2) {
      const int f0 = HK ? modn(p, NF) : 0;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int j = jrow0 + r;
        if (((m_r1 >> r) & 1u) && ((i + j + p) & 1) == c1) {
          const int lr = wv * R + r;
          const long long id = p * sxy + col[r];
          T aw = T(0), ae = T(0), as = T(0), an = T(0);
          if constexpr (HK) {
            aw = lfe[f0][lr][lane - 1]; ae = lfe[f0][lr][lane];
            as = lfn[f0][lr - 1][lane]; an = lfn[f0][lr][lane];
          }
          W7<T> w = weights_at<3, KIND, T>(op, id, aw, ae, as, an, ZF[r][1], ZF[r][2]);
          if (((m_fxy >> r) & 1u) || p == 1 || p == nz - 2) fold_center<3, T>(op, i, j, p, w);
          const T v = rounded<S, T>(gs_value<3, T>(w, B[r][1], lu[su][lr][lane + 1],
                                                   lu[su][lr][lane - 1], lu[su][lr + 1][lane],
                                                   lu[su][lr - 1][lane], U[r][3], U[r][1]));
          U[r][2] = v;
          lu[su][lr][lane] = v;
        }
      }
    }
    // second colour on plane q = p - 1 (tile), then the tile's plane q goes out
    const int q = p - 1;
    if (q >= k0 && q < k1) {
      const int sq = modn(q, NU);
      if (q >= 1 && q <= nz - 2) {
        const int f1 = HK ? modn(q, NF) : 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int j = jrow0 + r;
          if (((m_r2 >> r) & 1u) && ((i + j + q) & 1) == c2) {
            const int lr = wv * R + r;
            const long long id = q * sxy + col[r];
            T aw = T(0), ae = T(0), as = T(0), an = T(0);
            if constexpr (HK) {
              aw = lfe[f1][lr][lane - 1]; ae = lfe[f1][lr][lane];
              as = lfn[f1][lr - 1][lane]; an = lfn[f1][lr][lane];
            }
            W7<T> w = weights_at<3, KIND, T>(op, id, aw, ae, as, an, ZF[r][0], ZF[r][1]);
            if (((m_fxy >> r) & 1u) || q == 1 || q == nz - 2) fold_center<3, T>(op, i, j, q, w);
            U[r][1] = rounded<S, T>(gs_value<3, T>(w, B[r][0], lu[sq][lr][lane + 1],
                                                   lu[sq][lr][lane - 1], lu[sq][lr + 1][lane],
                                                   lu[sq][lr - 1][lane], U[r][2], U[r][0]));
          }
        }
      }
#pragma unroll
      for (int r = 0; r < R; ++r)
        if ((m_out >> r) & 1u) xout[q * sxy + col[r]] = to_store<S, T>(U[r][1]);
    }
    if constexpr (!ONEBAR) __syncthreads();
  }
}

// ------------------------------------------------------------------------------- 3D, ring
// Same tile, plane pipeline, colour order and per-node arithmetic as k_rbgs_fused3 (so the
// same bits in x_out), restructured for 4-byte types so that the plane step carries no
// bookkeeping:
//   * the plane loop is unrolled kRing = 6 times and every register queue is a ring indexed
//     by the step's phase (compile-time), so queues never shift (no register moves);
//   * LDS slots (3 for x, 2 for kappa, 3 for the face planes) are compile-time too, and each
//     LDS plane is padded by one row / column on every side, so every neighbour read is a
//     fixed immediate offset from a per-thread base and no read needs a clamp;
//   * global traffic goes through buffer instructions on a per-plane descriptor: the plane
//     offset is scalar, the in-plane offset a per-thread constant (no 64-bit vector address
//     arithmetic).  Stores of lanes outside the output tile use an offset past the
//     descriptor's range, which the hardware discards (no exec-mask branches);
//   * colour updates are computed on every lane and selected, instead of branching.
// Loads run two planes ahead for x and one for b and kappa.  Steps past the chunk's last plane
// (the loop runs a multiple of kRing steps) compute and discard; the host picks z chunks whose
// step count is a multiple of kRing where it can.
constexpr int kRing = 6;
template <int N> using IC = std::integral_constant<int, N>;

// DOT: also the partial of b . x_out over the nodes the workgroup stores (one per block; the
// stored nodes of all blocks cover the grid once), for the CG dot product r . z when b is the
// residual and x_out the preconditioned one.
template <int KIND, Prec P, int R, bool ZERO, bool DOT = false>
__global__ void __launch_bounds__(kThreads)
k_rbgs_ring3(DevOp op, const void* __restrict__ bv, const void* __restrict__ xin_v,
             void* __restrict__ xout_v, int c1, int tiles_x, int tiles_y, int ntiles,
             int kchunk, const KScalars* sc, double* partials = nullptr) {
  using T = typename PT<P>::T;
  using S = typename PT<P>::S;
  static_assert(sizeof(T) == 4, "ring sweep: 4-byte arithmetic types only");
  constexpr int TYE = R * kWaves;
  constexpr int TY = TYE - 2 * kHalo;
  constexpr bool HK = KIND == PDEHIP_OP_FLUX;
  constexpr int KR = HK ? TYE : 1;
  constexpr int PW = kLanes + 2;   // padded LDS row
  __shared__ T lu[3][TYE + 2][PW];
  __shared__ T lk[HK ? 2 : 1][KR + 2][PW];
  __shared__ T lfe[HK ? 3 : 1][KR + 2][PW];
  __shared__ T lfn[HK ? 3 : 1][KR + 2][PW];
  if (skip(sc)) return;
  const int t = xcd_tile();
  if (t >= ntiles) {
    if constexpr (DOT) {
      if (threadIdx.x == 0) partials[blockIdx.x] = 0.0;
    }
    return;
  }
  const int tx = t % tiles_x;
  const int rest = t / tiles_x;
  const int ty = rest % tiles_y;
  const int cz = rest / tiles_y;
  const int lane = threadIdx.x & (kLanes - 1);
  const int wv = threadIdx.x / kLanes;
  const int i = tx * kTileX - kHalo + lane;
  double dacc = 0.0;   // DOT
  const int k0 = cz * kchunk;
  const int k1 = min(k0 + kchunk, op.nz);
  const int nz = op.nz;
  const int c2 = 1 - c1;
  const long long sxy = op.sxy;
  const int pbytes = static_cast<int>(sxy * static_cast<long long>(sizeof(S)));
  const int kbytes = static_cast<int>(sxy * static_cast<long long>(sizeof(T)));
  const T fx = T(op.fx), fy = T(op.fy), fz = T(op.fz);

  const bool int_x = i >= 1 && i <= op.nx - 2;
  const bool in_x = i >= 0 && i < op.nx;
  const bool ring_x = lane >= 1 && lane <= kLanes - 2;
  const bool tile_x = lane >= kHalo && lane < kLanes - kHalo;
  unsigned m_r1 = 0, m_r2 = 0, m_fxy = 0;
  unsigned off_s[R], off_k[R], off_st[R];   // byte offsets in a plane (loads clamped in-grid)
  const int ci = min(max(i, 0), op.nx - 1);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int lr = wv * R + r;
    const int j = ty * TY - kHalo + lr;
    const bool in_y = j >= 0 && j < op.ny;
    const bool int_y = j >= 1 && j <= op.ny - 2;
    const bool ring_y = lr >= 1 && lr <= TYE - 2;
    const bool tile_y = lr >= kHalo && lr < TYE - kHalo;
    const unsigned col = static_cast<unsigned>(min(max(j, 0), op.ny - 1) * op.nx + ci);
    off_s[r] = col * static_cast<unsigned>(sizeof(S));
    off_k[r] = col * static_cast<unsigned>(sizeof(T));
    off_st[r] = (tile_x && tile_y && in_x && in_y) ? off_s[r] : kDropOffset;
    if (ring_x && ring_y && int_x && int_y) m_r1 |= 1u << r;
    if (tile_x && tile_y && int_x && int_y) m_r2 |= 1u << r;
    if (i == 1 || i == op.nx - 2 || j == 1 || j == op.ny - 2) m_fxy |= 1u << r;
  }
  const int jrow0 = ty * TY - kHalo + wv * R;
  // parity of (i + j) for row 0 of this thread; row r adds r
  const int par0 = (i + jrow0) & 1;

  auto kcl = [&](int k) -> long long { return static_cast<long long>(min(max(k, 0), nz - 1)) * sxy; };
  auto ld_u = [&](int r, int k) -> T {
    if constexpr (ZERO) return T(0);
    else return BufIO<S>::ld(plane_rsrc(xin_v, kcl(k), sizeof(S), pbytes), off_s[r]);
  };
  auto ld_b = [&](int r, int k) -> T {
    return BufIO<S>::ld(plane_rsrc(bv, kcl(k), sizeof(S), pbytes), off_s[r]);
  };
  auto ld_k = [&](int r, int k) -> T {
    if constexpr (HK) return BufIO<T>::ld(plane_rsrc(op.kappa, kcl(k), sizeof(T), kbytes), off_k[r]);
    else return T(0);
  };

  // Rings, with plane pstart + s at slot s mod L: U (L 6) holds x(p-2 .. p+3), B (L 3)
  // b(p-1 .. p+1), K (L 3) kappa(p+1 .. p+3), ZF (L 6) front faces zf(p-2 .. p+1).
  const int pstart = HK ? k0 - 3 : k0 - 1;
  T U[R][6], B[R][3], K[R][3], ZF[R][6];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    U[r][4] = ld_u(r, pstart - 2);
    U[r][5] = ld_u(r, pstart - 1);
    U[r][0] = ld_u(r, pstart);
    U[r][1] = ld_u(r, pstart + 1);
    U[r][2] = ld_u(r, pstart + 2);
    U[r][3] = T(0);
    B[r][2] = ld_b(r, pstart - 1);
    B[r][0] = ld_b(r, pstart);
    B[r][1] = T(0);
    K[r][0] = T(0);
    K[r][1] = ld_k(r, pstart + 1);
    K[r][2] = ld_k(r, pstart + 2);
#pragma unroll
    for (int s = 0; s < 6; ++s) ZF[r][s] = T(0);
  }

  auto step = [&](auto phc, int p) {
    constexpr int PH = decltype(phc)::value;
    constexpr int u_m2 = (PH + 4) % 6, u_m1 = (PH + 5) % 6, u_0 = PH % 6, u_1 = (PH + 1) % 6,
                  u_3 = (PH + 3) % 6;
    constexpr int b_m1 = (PH + 2) % 3, b_0 = PH % 3, b_1 = (PH + 1) % 3;
    constexpr int k_1 = (PH + 1) % 3, k_2 = (PH + 2) % 3, k_3 = PH % 3;
    constexpr int z_m2 = (PH + 4) % 6, z_m1 = (PH + 5) % 6, z_0 = PH % 6, z_1 = (PH + 1) % 6;
    constexpr int su = PH % 3, sq = (PH + 2) % 3, sk = PH % 2;
    constexpr int f_p1 = (PH + 1) % 3, f_0 = PH % 3, f_m1 = (PH + 2) % 3;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int lr = wv * R + r;
      U[r][u_3] = ld_u(r, p + 3);
      B[r][b_1] = ld_b(r, p + 1);
      if constexpr (HK) {
        K[r][k_3] = ld_k(r, p + 3);
        lk[sk][lr + 1][lane + 1] = K[r][k_1];
      }
      lu[su][lr + 1][lane + 1] = U[r][u_0];
    }
    __syncthreads();
    if constexpr (HK) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int lr = wv * R + r;
        const T kc = K[r][k_1];
        ZF[r][z_1] = face_smooth<T>(kc, K[r][k_2], fz);
        const T fe = face_smooth<T>(kc, lk[sk][lr + 1][lane + 2], fx);
        lfe[f_p1][lr + 1][lane + 1] = lane < kLanes - 1 ? fe : T(0);
        lfn[f_p1][lr + 1][lane + 1] = lr < TYE - 1 ? face_smooth<T>(kc, lk[sk][lr + 2][lane + 1], fy) : T(0);
      }
    }
    // first colour on plane p (tile + ring)
    if (p >= k0 - 1 && p >= 1 && p <= nz - 2) {
      const int cp = (par0 + p) & 1;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int lr = wv * R + r;
        const int j = jrow0 + r;
        const bool upd = ((m_r1 >> r) & 1u) && ((cp + r) & 1) == c1;
        T aw = T(0), ae = T(0), as = T(0), an = T(0);
        if constexpr (HK) {
          aw = lfe[f_0][lr + 1][lane]; ae = lfe[f_0][lr + 1][lane + 1];
          as = lfn[f_0][lr][lane + 1]; an = lfn[f_0][lr + 1][lane + 1];
        }
        W7<T> w = weights_at<3, KIND, T>(op, 0, aw, ae, as, an, ZF[r][z_m1], ZF[r][z_0]);
        if (((m_fxy >> r) & 1u) || p == 1 || p == nz - 2) fold_center<3, T>(op, i, j, p, w);
        const T v = rounded<S, T>(gs_value<3, T>(w, B[r][b_0], lu[su][lr + 1][lane + 2],
                                                 lu[su][lr + 1][lane], lu[su][lr + 2][lane + 1],
                                                 lu[su][lr][lane + 1], U[r][u_1], U[r][u_m1]));
        U[r][u_0] = upd ? v : U[r][u_0];
        lu[su][lr + 1][lane + 1] = U[r][u_0];
      }
    }
    // second colour on plane q = p - 1 (tile), then the tile's plane q goes out
    const int q = p - 1;
    if (q >= k0 && q < k1) {
      if (q >= 1 && q <= nz - 2) {
        const int cq = (par0 + q) & 1;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int lr = wv * R + r;
          const int j = jrow0 + r;
          const bool upd = ((m_r2 >> r) & 1u) && ((cq + r) & 1) == c2;
          T aw = T(0), ae = T(0), as = T(0), an = T(0);
          if constexpr (HK) {
            aw = lfe[f_m1][lr + 1][lane]; ae = lfe[f_m1][lr + 1][lane + 1];
            as = lfn[f_m1][lr][lane + 1]; an = lfn[f_m1][lr + 1][lane + 1];
          }
          W7<T> w = weights_at<3, KIND, T>(op, 0, aw, ae, as, an, ZF[r][z_m2], ZF[r][z_m1]);
          if (((m_fxy >> r) & 1u) || q == 1 || q == nz - 2) fold_center<3, T>(op, i, j, q, w);
          const T v = rounded<S, T>(gs_value<3, T>(w, B[r][b_m1], lu[sq][lr + 1][lane + 2],
                                                   lu[sq][lr + 1][lane], lu[sq][lr + 2][lane + 1],
                                                   lu[sq][lr][lane + 1], U[r][u_0], U[r][u_m2]));
          U[r][u_m1] = upd ? v : U[r][u_m1];
        }
      }
      const __amdgpu_buffer_rsrc_t ro = plane_rsrc(xout_v, static_cast<long long>(q) * sxy,
                                                   sizeof(S), pbytes);
#pragma unroll
      for (int r = 0; r < R; ++r) BufIO<S>::st(ro, off_st[r], to_store<S, T>(U[r][u_m1]));
      if constexpr (DOT) {
        // U holds S-representable values (rounded updates or loaded x_in): the stored values
#pragma unroll
        for (int r = 0; r < R; ++r)
          if (off_st[r] != kDropOffset) dacc += double(B[r][b_m1]) * double(U[r][u_m1]);
      }
    }
  };

  const int steps = k1 - pstart + 1;
  for (int s0 = 0; s0 < steps; s0 += kRing) {
    const int p = pstart + s0;
    step(IC<0>{}, p);
    step(IC<1>{}, p + 1);
    step(IC<2>{}, p + 2);
    step(IC<3>{}, p + 3);
    step(IC<4>{}, p + 4);
    step(IC<5>{}, p + 5);
  }
  if constexpr (DOT) {
    __shared__ double lred[kWaves];
    const double tot = block_sum(dacc, lred);
    if (threadIdx.x == 0) partials[blockIdx.x] = tot;
  }
}

// ------------------------------------------------------------------------------- 3D, register rows
// k_rbgs_reg3: the ring sweep's plane pipeline, colour order and per-node arithmetic (so the same
// bits in x_out), with the in-plane neighbours held in registers.  A wave owns kRR consecutive
// rows of a 64 x (8 kRR) extended tile (lane = column; output tile 60 x (8 kRR - 4)):
//   * x neighbours come from the adjacent lanes by DPP wave shifts, y neighbours from the
//     thread's own rows; only a wave's first and last rows meet another wave's rows, through
//     LDS slots of two rows per wave (x) and one row per wave (rho, north faces);
//   * every node forms its east and north faces of a plane once; the west face is the west
//     lane's east face (DPP), the south face the row below's north face;
//   * register rings indexed by the step's phase of a 6-step unrolled plane loop, buffer loads
//     and stores on per-plane descriptors, colour updates computed on every lane and selected,
//     as in k_rbgs_ring3.
// One barrier per plane: x slots are rewritten three planes later, rho slots two, face slots
// three, each after a barrier that every reader of the old contents has passed.  A step forms
// the faces of the next plane after its colour stages, so the register rings hold two face
// planes and three front-face planes.
constexpr int kRR = 4;

// Wave shifts with zero fill (bound_ctrl), so the compiler can fold the shift into the VOP2
// instruction that consumes it.  Lanes 0 and 63 read 0: they are the tile's outer halo, whose
// faces, centres and updates are formed but never selected or stored.
__device__ __forceinline__ float dpp_west(float v) {   // lane l gets lane l-1 (lane 0: 0)
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x138,
                                                            0xf, 0xf, true));
}
__device__ __forceinline__ float dpp_east(float v) {   // lane l gets lane l+1 (lane 63: 0)
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x130,
                                                            0xf, 0xf, true));
}

template <int KIND, Prec P, bool ZERO, bool DOT = false>
__global__ void __launch_bounds__(kThreads, 4)   // four waves per SIMD (two workgroups per CU)
k_rbgs_reg3(DevOp op, const void* __restrict__ bv, const void* __restrict__ xin_v,
            void* __restrict__ xout_v, int c1, int tiles_x, int tiles_y, int ntiles,
            int kchunk, const KScalars* sc, double* partials = nullptr) {
  using T = typename PT<P>::T;
  using S = typename PT<P>::S;
  static_assert(sizeof(T) == 4, "register-row sweep: 4-byte arithmetic types only");
  constexpr int R = kRR;
  constexpr int TYE = R * kWaves;
  constexpr int TY = TYE - 2 * kHalo;
  constexpr bool HK = KIND == PDEHIP_OP_FLUX;
  __shared__ T lxb[3][kWaves][2][kLanes];          // x of each wave's first / last row
  __shared__ T lkb[HK ? 2 : 1][kWaves][kLanes];    // rho of each wave's first row
  __shared__ T lfn[HK ? 3 : 1][kWaves][kLanes];    // north faces of each wave's last row
  if (skip(sc)) return;
  const int t = xcd_tile();
  if (t >= ntiles) {
    if constexpr (DOT) {
      if (threadIdx.x == 0) partials[blockIdx.x] = 0.0;
    }
    return;
  }
  const int tx = t % tiles_x;
  const int rest = t / tiles_x;
  const int ty = rest % tiles_y;
  const int cz = rest / tiles_y;
  const int lane = threadIdx.x & (kLanes - 1);
  const int wv = threadIdx.x / kLanes;
  const int wdn = (wv + kWaves - 1) & (kWaves - 1);   // the wave below (wraps: its values unused)
  const int wup = (wv + 1) & (kWaves - 1);            // the wave above
  const int i = tx * kTileX - kHalo + lane;
  double dacc = 0.0;   // DOT
  const int k0 = cz * kchunk;
  const int k1 = min(k0 + kchunk, op.nz);
  const int nz = op.nz;
  const int c2 = 1 - c1;
  const long long sxy = op.sxy;
  const int pbytes = static_cast<int>(sxy * static_cast<long long>(sizeof(S)));
  const int kbytes = static_cast<int>(sxy * static_cast<long long>(sizeof(T)));
  const T fx = T(op.fx), fy = T(op.fy), fz = T(op.fz);

  const bool int_x = i >= 1 && i <= op.nx - 2;
  const bool in_x = i >= 0 && i < op.nx;
  const bool ring_x = lane >= 1 && lane <= kLanes - 2;
  const bool tile_x = lane >= kHalo && lane < kLanes - kHalo;
  unsigned m_r1 = 0, m_r2 = 0, m_fxy = 0;
  unsigned off_s[R], off_k[R], off_st[R];
  const int ci = min(max(i, 0), op.nx - 1);
  const int jrow0 = ty * TY - kHalo + wv * R;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int lr = wv * R + r;
    const int j = jrow0 + r;
    const bool in_y = j >= 0 && j < op.ny;
    const bool int_y = j >= 1 && j <= op.ny - 2;
    const bool ring_y = lr >= 1 && lr <= TYE - 2;
    const bool tile_y = lr >= kHalo && lr < TYE - kHalo;
    const unsigned col = static_cast<unsigned>(min(max(j, 0), op.ny - 1) * op.nx + ci);
    off_s[r] = col * static_cast<unsigned>(sizeof(S));
    off_k[r] = col * static_cast<unsigned>(sizeof(T));
    off_st[r] = (tile_x && tile_y && in_x && in_y) ? off_s[r] : kDropOffset;
    if (ring_x && ring_y && int_x && int_y) m_r1 |= 1u << r;
    if (tile_x && tile_y && int_x && int_y) m_r2 |= 1u << r;
    if (i == 1 || i == op.nx - 2 || j == 1 || j == op.ny - 2) m_fxy |= 1u << r;
  }
  const int par0 = (i + jrow0) & 1;
  const bool top_ok = wv * R + R - 1 < TYE - 1;   // the last row has a north neighbour in the tile

  auto kcl = [&](int k) -> long long { return static_cast<long long>(min(max(k, 0), nz - 1)) * sxy; };
  auto ld_u = [&](int r, int k) -> T {
    if constexpr (ZERO) return T(0);
    else return BufIO<S>::ld(plane_rsrc(xin_v, kcl(k), sizeof(S), pbytes), off_s[r]);
  };
  auto ld_b = [&](int r, int k) -> T {
    return BufIO<S>::ld(plane_rsrc(bv, kcl(k), sizeof(S), pbytes), off_s[r]);
  };
  auto ld_k = [&](int r, int k) -> T {
    if constexpr (HK) return BufIO<T>::ld(plane_rsrc(op.kappa, kcl(k), sizeof(T), kbytes), off_k[r]);
    else return T(0);
  };

  // Rings, with plane pstart + s at slot s mod L: U (6) x(p-2 .. p+3), B (3) b(p-1 .. p+1),
  // K (3) rho(p+1 .. p+3), ZF (3) zf(p-2 .. p) and FE / FN (2) faces of planes p-1 .. p during
  // the colour stages of step p; the faces of plane p+1 are formed after them, in the slots the
  // second stage has just finished with.  RC (2): recip of the folded centre, formed by a node's
  // first stage and reused by its second (the same value gs_value would recompute).  The first step is the first colour's ring plane
  // k0 - 1; the faces and front faces it needs are formed before the loop.
  const int pstart = k0 - 1;
  T U[R][6], B[R][3], K[R][3], ZF[R][3], FE[R][2], FN[R][2], RC[R][2];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    U[r][4] = ld_u(r, pstart - 2);
    U[r][5] = ld_u(r, pstart - 1);
    U[r][0] = ld_u(r, pstart);
    U[r][1] = ld_u(r, pstart + 1);
    U[r][2] = ld_u(r, pstart + 2);
    U[r][3] = T(0);
    B[r][2] = ld_b(r, pstart - 1);
    B[r][0] = ld_b(r, pstart);
    B[r][1] = T(0);
    K[r][0] = T(0);
    K[r][1] = ld_k(r, pstart + 1);
    K[r][2] = ld_k(r, pstart + 2);
#pragma unroll
    for (int s = 0; s < 3; ++s) ZF[r][s] = T(0);
#pragma unroll
    for (int s = 0; s < 2; ++s) { FE[r][s] = T(0); FN[r][s] = T(0); RC[r][s] = T(0); }
  }
  if constexpr (HK) {
    // faces of plane k0 - 1 (slot 0) and front faces zf(k0 - 2) (slot 2), zf(k0 - 1) (slot 0)
    T km2[R], km1[R];
#pragma unroll
    for (int r = 0; r < R; ++r) { km2[r] = ld_k(r, pstart - 1); km1[r] = ld_k(r, pstart); }
    lkb[1][wv][lane] = km1[0];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) {
      ZF[r][2] = face_smooth<T>(km2[r], km1[r], fz);
      ZF[r][0] = face_smooth<T>(km1[r], K[r][1], fz);
      FE[r][0] = face_smooth<T>(km1[r], dpp_east(km1[r]), fx);
      FN[r][0] = face_smooth<T>(km1[r], r < R - 1 ? km1[r + 1] : lkb[1][wup][lane], fy);
    }
    lfn[0][wv][lane] = top_ok ? FN[R - 1][0] : T(0);
  }

  auto step = [&](auto phc, int p) {
    constexpr int PH = decltype(phc)::value;
    constexpr int u_m2 = (PH + 4) % 6, u_m1 = (PH + 5) % 6, u_0 = PH % 6, u_1 = (PH + 1) % 6,
                  u_3 = (PH + 3) % 6;
    constexpr int b_m1 = (PH + 2) % 3, b_0 = PH % 3, b_1 = (PH + 1) % 3;
    constexpr int k_1 = (PH + 1) % 3, k_2 = (PH + 2) % 3, k_3 = PH % 3;
    constexpr int z_m2 = (PH + 1) % 3, z_m1 = (PH + 2) % 3, z_0 = PH % 3;
    constexpr int f_0 = PH % 2, f_m1 = (PH + 1) % 2;
    constexpr int g_p1 = (PH + 1) % 3, g_0 = PH % 3, g_m1 = (PH + 2) % 3;   // lfn slots
    constexpr int sx = PH % 3, sxq = (PH + 2) % 3, sk = PH % 2;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      U[r][u_3] = ld_u(r, p + 3);
      B[r][b_1] = ld_b(r, p + 1);
      if constexpr (HK) K[r][k_3] = ld_k(r, p + 3);
    }
    lxb[sx][wv][0][lane] = U[0][u_0];
    lxb[sx][wv][1][lane] = U[R - 1][u_0];
    if constexpr (HK) lkb[sk][wv][lane] = K[0][k_1];
    __syncthreads();
    // first colour on plane p (tile + ring)
    if (p >= k0 - 1 && p >= 1 && p <= nz - 2) {
      const int cp = (par0 + p) & 1;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int j = jrow0 + r;
        const bool upd = ((m_r1 >> r) & 1u) && ((cp + r) & 1) == c1;
        T aw = T(0), ae = T(0), as = T(0), an = T(0);
        if constexpr (HK) {
          ae = FE[r][f_0];
          aw = dpp_west(ae);
          an = FN[r][f_0];
          as = r > 0 ? FN[r - 1][f_0] : lfn[g_0][wdn][lane];
        }
        W7<T> w = weights_at<3, KIND, T>(op, 0, aw, ae, as, an, ZF[r][z_m1], ZF[r][z_0]);
        if (((m_fxy >> r) & 1u) || p == 1 || p == nz - 2) fold_center<3, T>(op, i, j, p, w);
        RC[r][f_0] = recip<T>(w.C);
        const T xc = U[r][u_0];
        const T xn = r < R - 1 ? U[r + 1][u_0] : lxb[sx][wup][0][lane];
        const T xs = r > 0 ? U[r - 1][u_0] : lxb[sx][wdn][1][lane];
        const T v = rounded<S, T>((B[r][b_0] - gs_sum<3, T>(w, dpp_east(xc), dpp_west(xc), xn, xs,
                                                            U[r][u_1], U[r][u_m1])) * RC[r][f_0]);
        U[r][u_0] = upd ? v : xc;
      }
      lxb[sx][wv][0][lane] = U[0][u_0];
      lxb[sx][wv][1][lane] = U[R - 1][u_0];
    }
    // second colour on plane q = p - 1 (tile), then the tile's plane q goes out
    const int q = p - 1;
    if (q >= k0 && q < k1) {
      if (q >= 1 && q <= nz - 2) {
        const int cq = (par0 + q) & 1;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int j = jrow0 + r;
          const bool upd = ((m_r2 >> r) & 1u) && ((cq + r) & 1) == c2;
          T aw = T(0), ae = T(0), as = T(0), an = T(0);
          if constexpr (HK) {
            ae = FE[r][f_m1];
            aw = dpp_west(ae);
            an = FN[r][f_m1];
            as = r > 0 ? FN[r - 1][f_m1] : lfn[g_m1][wdn][lane];
          }
          // neighbour weights only: the folded centre's reciprocal is the first stage's
          const W7<T> w = weights_at<3, KIND, T>(op, 0, aw, ae, as, an, ZF[r][z_m2], ZF[r][z_m1]);
          const T xc = U[r][u_m1];
          const T xn = r < R - 1 ? U[r + 1][u_m1] : lxb[sxq][wup][0][lane];
          const T xs = r > 0 ? U[r - 1][u_m1] : lxb[sxq][wdn][1][lane];
          const T v = rounded<S, T>((B[r][b_m1] - gs_sum<3, T>(w, dpp_east(xc), dpp_west(xc), xn,
                                                               xs, U[r][u_0], U[r][u_m2])) *
                                    RC[r][f_m1]);
          U[r][u_m1] = upd ? v : xc;
        }
      }
      const __amdgpu_buffer_rsrc_t ro = plane_rsrc(xout_v, static_cast<long long>(q) * sxy,
                                                   sizeof(S), pbytes);
#pragma unroll
      for (int r = 0; r < R; ++r) BufIO<S>::st(ro, off_st[r], to_store<S, T>(U[r][u_m1]));
      if constexpr (DOT) {
#pragma unroll
        for (int r = 0; r < R; ++r)
          if (off_st[r] != kDropOffset) dacc += double(B[r][b_m1]) * double(U[r][u_m1]);
      }
    }
    if constexpr (HK) {
      // faces of plane p+1: east and north of every node, front in the ring (into the slots of
      // plane p-1, whose last reader was the second colour above)
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const T kc = K[r][k_1];
        ZF[r][z_m2] = face_smooth<T>(kc, K[r][k_2], fz);
        FE[r][f_m1] = face_smooth<T>(kc, dpp_east(kc), fx);
        const T kn = r < R - 1 ? K[r + 1][k_1] : lkb[sk][wup][lane];
        FN[r][f_m1] = face_smooth<T>(kc, kn, fy);
      }
      lfn[g_p1][wv][lane] = top_ok ? FN[R - 1][f_m1] : T(0);
    }
  };

  // steps k0 - 1 .. k1 (the last one holds only the second colour of plane k1 - 1); the step
  // count is the same in every thread, so every barrier is reached by all of them
  const int steps = k1 - pstart + 1;
  for (int s0 = 0; s0 < steps; s0 += kRing) {
    const int p = pstart + s0;
    step(IC<0>{}, p);
    if (s0 + 1 < steps) step(IC<1>{}, p + 1);
    if (s0 + 2 < steps) step(IC<2>{}, p + 2);
    if (s0 + 3 < steps) step(IC<3>{}, p + 3);
    if (s0 + 4 < steps) step(IC<4>{}, p + 4);
    if (s0 + 5 < steps) step(IC<5>{}, p + 5);
  }
  if constexpr (DOT) {
    __shared__ double lred[kWaves];
    const double tot = block_sum(dacc, lred);
    if (threadIdx.x == 0) partials[blockIdx.x] = tot;
  }
}

// ------------------------------------------------------------------------------- 2D
// One wavefront per 60-column strip; rows are the marching axis, x neighbours come from the
// adjacent lanes (ds_bpermute), y neighbours from the thread's row queue.  No LDS, no barrier.
// Faces: east from the lane to the right, north from the next row (both by face_smooth).
template <int KIND, Prec P, bool ZERO>
__global__ void __launch_bounds__(kThreads)
k_rbgs_fused2(DevOp op, const void* __restrict__ bv, const void* __restrict__ xin_v,
              void* __restrict__ xout_v, int c1, int strips, int ntiles, int jchunk,
              const KScalars* sc) {
  using T = typename PT<P>::T;
  using S = typename PT<P>::S;
  constexpr bool HK = KIND == PDEHIP_OP_FLUX;
  if (skip(sc)) return;
  const int t = xcd_tile() * kWaves + static_cast<int>(threadIdx.x / kLanes);
  if (t >= ntiles) return;   // whole wavefronts leave together
  const S* __restrict__ b = static_cast<const S*>(bv);
  const S* __restrict__ xin = static_cast<const S*>(xin_v);
  const T* __restrict__ kap = static_cast<const T*>(op.kappa);
  S* __restrict__ xout = static_cast<S*>(xout_v);
  const int sx = t % strips;
  const int cy = t / strips;
  const int lane = threadIdx.x & (kLanes - 1);
  const int i = sx * kTileX - kHalo + lane;
  const int j0 = cy * jchunk;
  const int j1 = min(j0 + jchunk, op.ny);
  const int ny = op.ny;
  const int c2 = 1 - c1;
  const bool in_x = i >= 0 && i < op.nx;
  const bool r1 = lane >= 1 && lane <= kLanes - 2 && i >= 1 && i <= op.nx - 2;
  const bool r2 = lane >= kHalo && lane < kLanes - kHalo && i >= 1 && i <= op.nx - 2;
  const bool out = lane >= kHalo && lane < kLanes - kHalo && in_x;
  const long long nx = op.nx;
  const T fx = T(op.fx), fy = T(op.fy);

  auto ld_u = [&](int j) -> T {
    if constexpr (ZERO) {
      return T(0);
    } else {
      const bool ok = in_x && j >= 0 && j < ny;
      const T v = T(ldS(xin, ok ? j * nx + i : 0));
      return ok ? v : T(0);
    }
  };
  auto ld_k = [&](int j) -> T {
    if constexpr (HK) {
      const bool ok = in_x && j >= 0 && j < ny;
      const T v = kap[ok ? j * nx + i : 0];
      return ok ? v : T(0);
    } else {
      return T(0);
    }
  };
  auto ld_b = [&](int j) -> T {
    const bool ok = r1 && j >= 1 && j < ny - 1;
    const T v = T(ldS(b, ok ? j * nx + i : 0));
    return ok ? v : T(0);
  };
  // queues after the shift of step p: U = x(p-2..p+1), B = b(p-1..p+1), K = kappa(p-2..p+2)
  T U[4], B[3], K[5];
  U[0] = U[1] = T(0); U[2] = ld_u(j0 - 2); U[3] = ld_u(j0 - 1);
  B[0] = B[1] = T(0); B[2] = ld_b(j0 - 1);
  K[0] = T(0); K[1] = ld_k(j0 - 3); K[2] = ld_k(j0 - 2); K[3] = ld_k(j0 - 1); K[4] = ld_k(j0);
  for (int p = j0 - 1; p <= j1; ++p) {
    U[0] = U[1]; U[1] = U[2]; U[2] = U[3]; U[3] = ld_u(p + 1);
    B[0] = B[1]; B[1] = B[2]; B[2] = ld_b(p + 1);
    if constexpr (HK) { K[0] = K[1]; K[1] = K[2]; K[2] = K[3]; K[3] = K[4]; K[4] = ld_k(p + 2); }
    // first colour on row p
    {
      const T uw = __shfl_up(U[2], 1, kLanes), ue = __shfl_down(U[2], 1, kLanes);
      T kw = T(0), ke = T(0);
      if constexpr (HK) { kw = __shfl_up(K[2], 1, kLanes); ke = __shfl_down(K[2], 1, kLanes); }
      if (r1 && p >= 1 && p <= ny - 2 && ((i + p) & 1) == c1) {
        const long long id = p * nx + i;
        T aw = T(0), ae = T(0), as = T(0), an = T(0);
        if constexpr (HK) {
          aw = face_smooth<T>(K[2], kw, fx); ae = face_smooth<T>(K[2], ke, fx);
          as = face_smooth<T>(K[2], K[1], fy); an = face_smooth<T>(K[2], K[3], fy);
        }
        W7<T> w = weights_at<2, KIND, T>(op, id, aw, ae, as, an, T(0), T(0));
        fold_center<2, T>(op, i, p, 0, w);
        U[2] = rounded<S, T>(gs_value<2, T>(w, B[1], ue, uw, U[3], U[1], T(0), T(0)));
      }
    }
    // second colour on row q = p - 1
    const int q = p - 1;
    if (q >= j0 && q < j1) {
      const T uw = __shfl_up(U[1], 1, kLanes), ue = __shfl_down(U[1], 1, kLanes);
      T kw = T(0), ke = T(0);
      if constexpr (HK) { kw = __shfl_up(K[1], 1, kLanes); ke = __shfl_down(K[1], 1, kLanes); }
      if (r2 && q >= 1 && q <= ny - 2 && ((i + q) & 1) == c2) {
        const long long id = q * nx + i;
        T aw = T(0), ae = T(0), as = T(0), an = T(0);
        if constexpr (HK) {
          aw = face_smooth<T>(K[1], kw, fx); ae = face_smooth<T>(K[1], ke, fx);
          as = face_smooth<T>(K[1], K[0], fy); an = face_smooth<T>(K[1], K[2], fy);
        }
        W7<T> w = weights_at<2, KIND, T>(op, id, aw, ae, as, an, T(0), T(0));
        fold_center<2, T>(op, i, q, 0, w);
        U[1] = rounded<S, T>(gs_value<2, T>(w, B[0], ue, uw, U[2], U[0], T(0), T(0)));
      }
      if (out) xout[q * nx + i] = to_store<S, T>(U[1]);
    }
  }
}

// ------------------------------------------------------------------------------- 3D, paired
// Divergence-free variant: each lane owns a column pair (2l, 2l+1) of a 128-wide extended tile
// (output 124 wide).  In any row and plane the first colour sits at the same pair offset s in
// every lane (s depends on row, plane and tile origin only), and the second colour of plane
// p-1 at the same offset, so both colour phases keep every lane busy.  Same plane pipeline,
// same face ring and the same arithmetic per node as k_rbgs_fused3.
constexpr int kPairCols = 2 * kLanes;
constexpr int kPairTileX = kPairCols - 2 * kHalo;

template <int KIND, Prec P, int R, bool ZERO>
__global__ void __launch_bounds__(kThreads)
k_rbgs_pair3(DevOp op, const void* __restrict__ bv, const void* __restrict__ xin_v,
             void* __restrict__ xout_v, int c1, int tiles_x, int tiles_y, int ntiles, int kchunk,
             const KScalars* sc) {
  using T = typename PT<P>::T;
  using S = typename PT<P>::S;
  constexpr int TYE = R * kWaves;
  constexpr int TY = TYE - 2 * kHalo;
  constexpr bool HK = KIND == PDEHIP_OP_FLUX;
  constexpr int KR = HK ? TYE : 1;
  constexpr int W = kPairCols;
  __shared__ T lu[2][TYE][W];
  __shared__ T lk[KR][W];
  __shared__ T lfe[HK ? 3 : 1][KR][W];
  __shared__ T lfn[HK ? 3 : 1][KR][W];
  if (skip(sc)) return;
  const int t = xcd_tile();
  if (t >= ntiles) return;
  const S* __restrict__ b = static_cast<const S*>(bv);
  const S* __restrict__ xin = static_cast<const S*>(xin_v);
  const T* __restrict__ kap = static_cast<const T*>(op.kappa);
  S* __restrict__ xout = static_cast<S*>(xout_v);
  const int tx = t % tiles_x;
  const int rest = t / tiles_x;
  const int ty = rest % tiles_y;
  const int cz = rest / tiles_y;
  const int lane = threadIdx.x & (kLanes - 1);
  const int wv = threadIdx.x / kLanes;
  const int i0 = tx * kPairTileX - kHalo;      // global x of extended column 0 (even)
  const int k0 = cz * kchunk;
  const int k1 = min(k0 + kchunk, op.nz);
  const int nz = op.nz;
  const long long sxy = op.sxy;
  const T fx = T(op.fx), fy = T(op.fy), fz = T(op.fz);
  const int jrow0 = ty * TY - kHalo + wv * R;

  // masks, bit (2r + s): in grid, first-colour candidate, second-colour candidate, output
  unsigned m_in = 0, m_r1 = 0, m_r2 = 0, m_out = 0;
  long long col[R][2];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int lr = wv * R + r;
    const int j = jrow0 + r;
    const bool in_y = j >= 0 && j < op.ny;
    const bool int_y = j >= 1 && j <= op.ny - 2;
    const bool ring_y = lr >= 1 && lr <= TYE - 2;
    const bool tile_y = lr >= kHalo && lr < TYE - kHalo;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int c = 2 * lane + q;
      const int i = i0 + c;
      const bool in_x = i >= 0 && i < op.nx;
      const bool int_x = i >= 1 && i <= op.nx - 2;
      const bool ring_x = c >= 1 && c <= W - 2;
      const bool tile_x = c >= kHalo && c < W - kHalo;
      const unsigned bit = 1u << (2 * r + q);
      col[r][q] = in_x && in_y ? (long long)j * op.nx + i : 0;
      if (in_x && in_y) m_in |= bit;
      if (ring_x && ring_y && int_x && int_y) m_r1 |= bit;
      if (tile_x && tile_y && int_x && int_y) m_r2 |= bit;
      if (tile_x && tile_y && in_x && in_y) m_out |= bit;
    }
  }
  auto ld_u = [&](int r, int q, int k) -> T {
    if constexpr (ZERO) {
      return T(0);
    } else {
      const bool ok = ((m_in >> (2 * r + q)) & 1u) && k >= 0 && k < nz;
      const T v = T(ldS(xin, ok ? k * sxy + col[r][q] : 0));
      return ok ? v : T(0);
    }
  };
  auto ld_k = [&](int r, int q, int k) -> T {
    if constexpr (HK) {
      const bool ok = ((m_in >> (2 * r + q)) & 1u) && k >= 0 && k < nz;
      const T v = kap[ok ? k * sxy + col[r][q] : 0];
      return ok ? v : T(0);
    } else {
      return T(0);
    }
  };
  auto ld_b = [&](int r, int q, int k) -> T {
    const bool ok = ((m_r1 >> (2 * r + q)) & 1u) && k >= 1 && k < nz - 1;
    const T v = T(ldS(b, ok ? k * sxy + col[r][q] : 0));
    return ok ? v : T(0);
  };

  const int pstart = HK ? k0 - 3 : k0 - 1;
  T U[R][2][4], B[R][2][3], K[R][2][3], ZF[R][2][4];
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      U[r][q][0] = U[r][q][1] = T(0);
      U[r][q][2] = ld_u(r, q, pstart - 1);
      U[r][q][3] = ld_u(r, q, pstart);
      B[r][q][0] = B[r][q][1] = T(0);
      B[r][q][2] = ld_b(r, q, pstart);
      K[r][q][0] = T(0);
      K[r][q][1] = ld_k(r, q, pstart + 1);
      K[r][q][2] = ld_k(r, q, pstart + 2);
      ZF[r][q][0] = ZF[r][q][1] = ZF[r][q][2] = ZF[r][q][3] = T(0);
    }

  for (int p = pstart; p <= k1; ++p) {
    const int sl = p & 1;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int lr = wv * R + r;
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        U[r][q][0] = U[r][q][1]; U[r][q][1] = U[r][q][2]; U[r][q][2] = U[r][q][3];
        U[r][q][3] = ld_u(r, q, p + 1);
        B[r][q][0] = B[r][q][1]; B[r][q][1] = B[r][q][2]; B[r][q][2] = ld_b(r, q, p + 1);
        if constexpr (HK) {
          K[r][q][0] = K[r][q][1]; K[r][q][1] = K[r][q][2]; K[r][q][2] = ld_k(r, q, p + 3);
          ZF[r][q][0] = ZF[r][q][1]; ZF[r][q][1] = ZF[r][q][2]; ZF[r][q][2] = ZF[r][q][3];
          lk[lr][2 * lane + q] = K[r][q][0];
        }
        lu[sl][lr][2 * lane + q] = U[r][q][2];
      }
    }
    __syncthreads();
    if constexpr (HK) {
      const int fs = mod3(p + 1);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int lr = wv * R + r;
        const T ke = lane < kLanes - 1 ? lk[lr][2 * lane + 2] : T(0);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int c = 2 * lane + q;
          const T kc = K[r][q][0];
          ZF[r][q][3] = face_smooth<T>(kc, K[r][q][1], fz);
          lfe[fs][lr][c] = face_smooth<T>(kc, q == 0 ? K[r][1][0] : ke, fx);
          lfn[fs][lr][c] = lr < TYE - 1 ? face_smooth<T>(kc, lk[lr + 1][c], fy) : T(0);
        }
      }
    }
    // first colour on plane p (tile + ring).  The pair offset s of a row is wave-uniform: branch
    // on it into two compile-time bodies (a runtime index into the register arrays would put
    // them in scratch memory).
    if (p >= k0 - 1 && p >= 1 && p <= nz - 2) {
      const int f0 = HK ? mod3(p) : 0;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int j = jrow0 + r;
        const int lr = wv * R + r;
        auto body = [&](auto qc) {
          constexpr int q = decltype(qc)::value;
          if (!((m_r1 >> (2 * r + q)) & 1u)) return;
          const int c = 2 * lane + q;
          const long long id = p * sxy + col[r][q];
          T aw = T(0), ae = T(0), as = T(0), an = T(0);
          if constexpr (HK) {
            aw = lfe[f0][lr][c - 1]; ae = lfe[f0][lr][c];
            as = lfn[f0][lr - 1][c]; an = lfn[f0][lr][c];
          }
          W7<T> w = weights_at<3, KIND, T>(op, id, aw, ae, as, an, ZF[r][q][1], ZF[r][q][2]);
          fold_center<3, T>(op, i0 + c, j, p, w);
          const T v = rounded<S, T>(gs_value<3, T>(w, B[r][q][1], lu[sl][lr][c + 1],
                                                   lu[sl][lr][c - 1], lu[sl][lr + 1][c],
                                                   lu[sl][lr - 1][c], U[r][q][3], U[r][q][1]));
          U[r][q][2] = v;
          lu[sl][lr][c] = v;
        };
        if ((c1 + i0 + j + p) & 1) body(std::integral_constant<int, 1>{});
        else body(std::integral_constant<int, 0>{});
      }
    }
    // second colour on plane q = p - 1 (tile): the same pair offset as plane p's first colour
    const int qp = p - 1;
    if (qp >= k0 && qp < k1) {
      const int sq = sl ^ 1;
      if (qp >= 1 && qp <= nz - 2) {
        const int f1 = HK ? mod3(qp) : 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int j = jrow0 + r;
          const int lr = wv * R + r;
          auto body = [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            if (!((m_r2 >> (2 * r + q)) & 1u)) return;
            const int c = 2 * lane + q;
            const long long id = qp * sxy + col[r][q];
            T aw = T(0), ae = T(0), as = T(0), an = T(0);
            if constexpr (HK) {
              aw = lfe[f1][lr][c - 1]; ae = lfe[f1][lr][c];
              as = lfn[f1][lr - 1][c]; an = lfn[f1][lr][c];
            }
            W7<T> w = weights_at<3, KIND, T>(op, id, aw, ae, as, an, ZF[r][q][0], ZF[r][q][1]);
            fold_center<3, T>(op, i0 + c, j, qp, w);
            U[r][q][1] = rounded<S, T>(gs_value<3, T>(w, B[r][q][0], lu[sq][lr][c + 1],
                                                      lu[sq][lr][c - 1], lu[sq][lr + 1][c],
                                                      lu[sq][lr - 1][c], U[r][q][2], U[r][q][0]));
          };
          if ((c1 + i0 + j + p) & 1) body(std::integral_constant<int, 1>{});
          else body(std::integral_constant<int, 0>{});
        }
      }
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int q = 0; q < 2; ++q)
          if ((m_out >> (2 * r + q)) & 1u) xout[qp * sxy + col[r][q]] = to_store<S, T>(U[r][q][1]);
    }
    __syncthreads();
  }
}

SmoothTuning g_tune{};

template <int KIND, Prec P>
void launch3(const Launch& L, const DevOp& op, const void* b, const void* xin, void* xout, int c1,
             bool zero_in) {
  const int rows = g_tune.rows == 4 ? 4 : 2;
  const int ty = rows * kWaves - 2 * kHalo;
  const int tiles_x = (op.nx + kTileX - 1) / kTileX;
  const int tiles_y = (op.ny + ty - 1) / ty;
  // z chunk: configured, or chosen so the launch fills the resident workgroup slots in whole
  // rounds (occupancy of the exact kernel instantiation)
  static const int occ4f = blocks_per_cu(k_rbgs_fused3<KIND, P, 4, false>, kThreads);
  static const int occ2f = blocks_per_cu(k_rbgs_fused3<KIND, P, 2, false>, kThreads);
  const int capacity = device_cus() * (rows == 4 ? occ4f : occ2f);
  const int prologue = KIND == PDEHIP_OP_FLUX ? 3 : 1;
  const int kchunk = g_tune.zchunk > 0 ? g_tune.zchunk
                                       : plane_chunk((long long)tiles_x * tiles_y, op.nz, capacity, prologue, 4);
  const int chunks = (op.nz + kchunk - 1) / kchunk;
  const int ntiles = tiles_x * tiles_y * chunks;
  const dim3 g(static_cast<unsigned>(8 * ((ntiles + 7) / 8)));
#define PDEHIP_F3(RR, ZZ)                                                                        \
  hipLaunchKernelGGL((k_rbgs_fused3<KIND, P, RR, ZZ>), g, dim3(kThreads), 0, L.stream, op, b,     \
                     xin, xout, c1, tiles_x, tiles_y, ntiles, kchunk, L.sc)
  if (rows == 4) {
    if (zero_in) PDEHIP_F3(4, true); else PDEHIP_F3(4, false);
  } else {
    if (zero_in) PDEHIP_F3(2, true); else PDEHIP_F3(2, false);
  }
#undef PDEHIP_F3
}

// z chunk for the ring sweep: the step count of a chunk (kchunk + prologue + 1) is a multiple
// of kRing, chosen like plane_chunk (whole rounds of resident workgroups) with the padded step
// count as the cost of a chunk.
int ring_chunk(long long tiles, int nz, int capacity, int prologue) {
  int best_kc = nz;
  long long best = -1;
  for (int kc = 2; kc <= nz + kRing; ++kc) {
    if ((kc + prologue + 1) % kRing != 0) continue;
    const int chunks = (nz + kc - 1) / kc;
    const long long n = tiles * chunks;
    const long long rounds = (n + capacity - 1) / capacity;
    const long long cost = rounds * (kc + prologue + 1);
    if (best < 0 || cost < best) { best = cost; best_kc = kc; }
    if (chunks == 1) break;
  }
  return best_kc;
}

template <int KIND, Prec P>
int launch3_ring(const Launch& L, const DevOp& op, const void* b, const void* xin, void* xout,
                 int c1, bool zero_in, double* partials = nullptr, int cap = 0) {
  constexpr int rows = 2;
  const int ty = rows * kWaves - 2 * kHalo;
  const int tiles_x = (op.nx + kTileX - 1) / kTileX;
  const int tiles_y = (op.ny + ty - 1) / ty;
  static const int occ = blocks_per_cu(k_rbgs_ring3<KIND, P, rows, false>, kThreads);
  const int prologue = KIND == PDEHIP_OP_FLUX ? 3 : 1;
  const int kchunk = g_tune.zchunk > 0 ? g_tune.zchunk
                                       : ring_chunk((long long)tiles_x * tiles_y, op.nz,
                                                    device_cus() * occ, prologue);
  const int chunks = (op.nz + kchunk - 1) / kchunk;
  const int ntiles = tiles_x * tiles_y * chunks;
  const dim3 g(static_cast<unsigned>(8 * ((ntiles + 7) / 8)));
  if (partials) {
    // the dot variant: one partial per block, when they fit the caller's buffer
    if (zero_in || static_cast<int>(g.x) > cap) return 0;
    hipLaunchKernelGGL((k_rbgs_ring3<KIND, P, rows, false, true>), g, dim3(kThreads), 0, L.stream,
                       op, b, xin, xout, c1, tiles_x, tiles_y, ntiles, kchunk, L.sc, partials);
    return static_cast<int>(g.x);
  }
  if (zero_in)
    hipLaunchKernelGGL((k_rbgs_ring3<KIND, P, rows, true>), g, dim3(kThreads), 0, L.stream, op, b,
                       xin, xout, c1, tiles_x, tiles_y, ntiles, kchunk, L.sc, nullptr);
  else
    hipLaunchKernelGGL((k_rbgs_ring3<KIND, P, rows, false>), g, dim3(kThreads), 0, L.stream, op, b,
                       xin, xout, c1, tiles_x, tiles_y, ntiles, kchunk, L.sc, nullptr);
  return 0;
}

template <int KIND, Prec P>
int launch3_reg(const Launch& L, const DevOp& op, const void* b, const void* xin, void* xout,
                int c1, bool zero_in, double* partials = nullptr, int cap = 0) {
  const int ty = kRR * kWaves - 2 * kHalo;
  const int tiles_x = (op.nx + kTileX - 1) / kTileX;
  const int tiles_y = (op.ny + ty - 1) / ty;
  static const int occ = blocks_per_cu(k_rbgs_reg3<KIND, P, false>, kThreads);
  // a chunk of kchunk planes takes kchunk + 2 plane steps (any count: no dead steps)
  const int kchunk = g_tune.zchunk > 0 ? g_tune.zchunk
                                       : plane_chunk((long long)tiles_x * tiles_y, op.nz,
                                                     device_cus() * occ, 2, 2);
  const int chunks = (op.nz + kchunk - 1) / kchunk;
  const int ntiles = tiles_x * tiles_y * chunks;
  const dim3 g(static_cast<unsigned>(8 * ((ntiles + 7) / 8)));
  if (partials) {
    // the fused dot runs on FP32 levels only (an FP16 level's dot sweep is the ring kernel's)
    if constexpr (P == P32) {
      if (zero_in || static_cast<int>(g.x) > cap) return 0;
      hipLaunchKernelGGL((k_rbgs_reg3<KIND, P, false, true>), g, dim3(kThreads), 0, L.stream, op,
                         b, xin, xout, c1, tiles_x, tiles_y, ntiles, kchunk, L.sc, partials);
      return static_cast<int>(g.x);
    } else {
      return 0;
    }
  }
  if (zero_in)
    hipLaunchKernelGGL((k_rbgs_reg3<KIND, P, true>), g, dim3(kThreads), 0, L.stream, op, b, xin,
                       xout, c1, tiles_x, tiles_y, ntiles, kchunk, L.sc, nullptr);
  else
    hipLaunchKernelGGL((k_rbgs_reg3<KIND, P, false>), g, dim3(kThreads), 0, L.stream, op, b, xin,
                       xout, c1, tiles_x, tiles_y, ntiles, kchunk, L.sc, nullptr);
  return 0;
}

template <int KIND, Prec P>
void launch3_pair(const Launch& L, const DevOp& op, const void* b, const void* xin, void* xout,
                  int c1, bool zero_in) {
  constexpr int rows = 2;
  const int ty = rows * kWaves - 2 * kHalo;
  const int tiles_x = (op.nx + kPairTileX - 1) / kPairTileX;
  const int tiles_y = (op.ny + ty - 1) / ty;
  static const int occ = blocks_per_cu(k_rbgs_pair3<KIND, P, rows, false>, kThreads);
  const int prologue = KIND == PDEHIP_OP_FLUX ? 3 : 1;
  const int kchunk = g_tune.zchunk > 0 ? g_tune.zchunk
                                       : plane_chunk((long long)tiles_x * tiles_y, op.nz,
                                                     device_cus() * occ, prologue, 4);
  const int chunks = (op.nz + kchunk - 1) / kchunk;
  const int ntiles = tiles_x * tiles_y * chunks;
  const dim3 g(static_cast<unsigned>(8 * ((ntiles + 7) / 8)));
  if (zero_in)
    hipLaunchKernelGGL((k_rbgs_pair3<KIND, P, rows, true>), g, dim3(kThreads), 0, L.stream, op, b,
                       xin, xout, c1, tiles_x, tiles_y, ntiles, kchunk, L.sc);
  else
    hipLaunchKernelGGL((k_rbgs_pair3<KIND, P, rows, false>), g, dim3(kThreads), 0, L.stream, op, b,
                       xin, xout, c1, tiles_x, tiles_y, ntiles, kchunk, L.sc);
}

template <int KIND, Prec P>
void launch2(const Launch& L, const DevOp& op, const void* b, const void* xin, void* xout, int c1,
             bool zero_in) {
  const int strips = (op.nx + kTileX - 1) / kTileX;
  // rows per wavefront: enough wavefronts to fill the chip, at least 16 rows each
  int jchunk = g_tune.ychunk > 0 ? g_tune.ychunk : 64;
  while (jchunk > 16 && (long long)strips * ((op.ny + jchunk - 1) / jchunk) < 4096) jchunk /= 2;
  const int chunks = (op.ny + jchunk - 1) / jchunk;
  const int ntiles = strips * chunks;
  const int groups = (ntiles + kWaves - 1) / kWaves;
  const dim3 g(static_cast<unsigned>(8 * ((groups + 7) / 8)));
  if (zero_in)
    hipLaunchKernelGGL((k_rbgs_fused2<KIND, P, true>), g, dim3(kThreads), 0, L.stream, op, b, xin,
                       xout, c1, strips, ntiles, jchunk, L.sc);
  else
    hipLaunchKernelGGL((k_rbgs_fused2<KIND, P, false>), g, dim3(kThreads), 0, L.stream, op, b, xin,
                       xout, c1, strips, ntiles, jchunk, L.sc);
}

template <int KIND>
void launch_kind(const Launch& L, const DevOp& op, Prec p, const void* b, const void* xin,
                 void* xout, int c1, bool zero_in) {
  if (op.nz > 1) {
    const bool pair = g_tune.pair != 0;
    // ring sweep: buffer offsets of a plane must stay below 2^31 bytes
    const bool ring = g_tune.ring != 0 && g_tune.rows == 2 && op.sxy < (1LL << 29);
    switch (p) {
      case P64: launch3<KIND, P64>(L, op, b, xin, xout, c1, zero_in); break;
      case P32:
        if (pair) launch3_pair<KIND, P32>(L, op, b, xin, xout, c1, zero_in);
        else if (ring && g_tune.reg) launch3_reg<KIND, P32>(L, op, b, xin, xout, c1, zero_in);
        else if (ring) launch3_ring<KIND, P32>(L, op, b, xin, xout, c1, zero_in);
        else launch3<KIND, P32>(L, op, b, xin, xout, c1, zero_in);
        break;
      default:
        if (pair) launch3_pair<KIND, P16>(L, op, b, xin, xout, c1, zero_in);
        else if (ring && g_tune.reg) launch3_reg<KIND, P16>(L, op, b, xin, xout, c1, zero_in);
        else if (ring) launch3_ring<KIND, P16>(L, op, b, xin, xout, c1, zero_in);
        else launch3<KIND, P16>(L, op, b, xin, xout, c1, zero_in);
        break;
    }
  } else {
    switch (p) {
      case P64: launch2<KIND, P64>(L, op, b, xin, xout, c1, zero_in); break;
      case P32: launch2<KIND, P32>(L, op, b, xin, xout, c1, zero_in); break;
      default: launch2<KIND, P16>(L, op, b, xin, xout, c1, zero_in); break;
    }
  }
}

}  // namespace

void set_smooth_tuning(const SmoothTuning& t) { g_tune = t; }
SmoothTuning smooth_tuning() { return g_tune; }

int rbgs_sweep_dot(const Launch& L, const DevOp& op, Prec p, const void* b, const void* xin,
                   void* xout, int first_color, double* partials, int cap) {
  const bool ring = g_tune.ring != 0 && g_tune.rows == 2 && g_tune.pair == 0 && g_tune.dot != 0 &&
                    op.sxy < (1LL << 29) && op.nz > 1 && (p == P32 || p == P16);
  if (!ring) return 0;
  // register rows for FP32 levels; an FP16 level (not a mixed plan's level 0) takes the ring
  // sweep, whose FP16 dot variant fits the register budget without spilling
  if (g_tune.reg && p == P32) {
    if (op.kind == PDEHIP_OP_CONST)
      return launch3_reg<PDEHIP_OP_CONST, P32>(L, op, b, xin, xout, first_color, false, partials, cap);
    if (op.kind == PDEHIP_OP_FLUX)
      return launch3_reg<PDEHIP_OP_FLUX, P32>(L, op, b, xin, xout, first_color, false, partials, cap);
    return 0;
  }
  if (op.kind == PDEHIP_OP_CONST)
    return p == P32 ? launch3_ring<PDEHIP_OP_CONST, P32>(L, op, b, xin, xout, first_color, false, partials, cap)
                    : launch3_ring<PDEHIP_OP_CONST, P16>(L, op, b, xin, xout, first_color, false, partials, cap);
  if (op.kind == PDEHIP_OP_FLUX)
    return p == P32 ? launch3_ring<PDEHIP_OP_FLUX, P32>(L, op, b, xin, xout, first_color, false, partials, cap)
                    : launch3_ring<PDEHIP_OP_FLUX, P16>(L, op, b, xin, xout, first_color, false, partials, cap);
  return 0;
}

void rbgs_sweep(const Launch& L, const DevOp& op, Prec p, const void* b, const void* xin,
                void* xout, int first_color, bool zero_in) {
  switch (op.kind) {
    case PDEHIP_OP_CONST: launch_kind<PDEHIP_OP_CONST>(L, op, p, b, xin, xout, first_color, zero_in); break;
    case PDEHIP_OP_POINTWISE:
      if (op.nz > 1) launch3<PDEHIP_OP_POINTWISE, P64>(L, op, b, xin, xout, first_color, zero_in);
      else launch2<PDEHIP_OP_POINTWISE, P64>(L, op, b, xin, xout, first_color, zero_in);
      break;
    default: launch_kind<PDEHIP_OP_FLUX>(L, op, p, b, xin, xout, first_color, zero_in); break;
  }
}

}  // namespace pdehip
