// kernels.hpp — host-side launchers of the device kernels (defined in k_*.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace pdehip {

// Storage/compute precision of one level's vectors.
//  P64 : double vectors, double kappa, double arithmetic
//  P32 : float vectors, float kappa, float arithmetic
//  P16 : __half vectors (scaled), float kappa, float arithmetic
enum Prec { P64 = 0, P32 = 1, P16 = 2 };

inline size_t prec_bytes(Prec p) { return p == P64 ? 8 : (p == P32 ? 4 : 2); }
inline size_t kappa_bytes(Prec p) { return p == P64 ? 8 : 4; }

// Compute units of the current device and resident workgroups per CU of a kernel (cached by
// the callers per kernel instantiation).
int device_cus();
template <typename K>
int blocks_per_cu(K kernel, int threads) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, threads, 0) != hipSuccess || n < 1) n = 1;
  return n;
}

struct Launch {
  hipStream_t stream;
  const KScalars* sc;   // device scalars; kernels return early once sc->done (may be null)
};

// ---- stencils (k_stencil.hip)
// y = A x (identity boundary rows).  If dot_partials != null, also block partials of x.y.
void apply(const Launch& L, const DevOp& op, Prec p, const void* x, void* y, double* dot_partials);
// r = b - A x on every row (boundary rows: b - x).  If partials != null, block partials of r.r.
void residual(const Launch& L, const DevOp& op, Prec p, const void* b, const void* x, void* r,
              double* partials);
// One colour of red-black Gauss-Seidel for A x = b (interior rows of colour (i+j+k)%2 == color).
void rbgs(const Launch& L, const DevOp& op, Prec p, const void* b, void* x, int color);
// One full red-black sweep (colour first_color, then the other) out of place: x_out = the x an
// in-place rbgs(first_color); rbgs(1 - first_color) would leave, read from x_in.  zero_in: x_in
// is known to be zero everywhere (its loads are skipped; x_in may be null).  (k_smooth.hip)
void rbgs_sweep(const Launch& L, const DevOp& op, Prec p, const void* b, const void* x_in,
                void* x_out, int first_color, bool zero_in);
// rbgs_sweep (x_in not zero) that also leaves the partials of b . x_out, one per block, when
// the 3D FP32/FP16 ring sweep serves this level and its grid fits `cap` partials.  Returns the
// number of partials written, or 0 after launching nothing (the caller sweeps and dots apart).
int rbgs_sweep_dot(const Launch& L, const DevOp& op, Prec p, const void* b, const void* x_in,
                   void* x_out, int first_color, double* partials, int cap);
// Launch shape of the fused sweep (process-wide).
struct SmoothTuning {
  int fused = 1;     // 0: multigrid smooths with the two-launch in-place sweep
  int rows = 2;      // 3D: rows per thread (tile 60 x (8*rows - 4)), 2 or 4
  int zchunk = 0;    // 3D: z planes per workgroup (0: chosen from occupancy and grid size)
  int ychunk = 64;   // 2D: rows per wavefront (upper bound)
  int pair = 0;      // 3D FP32/FP16: column-pair (divergence-free) tiles
  int ring = 1;      // 3D FP32/FP16: ring-buffered, 6-step unrolled sweep (k_rbgs_ring3)
  int reg = 1;       // ... with in-plane neighbours in registers, 4 rows per wave (k_rbgs_reg3)
  int dot = 1;       // the last level-0 sweep of a mixed V-cycle also leaves the r.z partials
  int box = 65;      // 3D FP32/FP16 coarse levels of mixed V-cycles with every side <= box
                     // smooth on boxes (k_box.hip); 0: plane-march sweeps on every level
  int box2 = 33;     // ... two sweeps per box launch where every side is <= box2, else one
  int box_z = 4;     // interior planes of a box (4 or 8); 4 with box2 = 33 measured 1.2%
                     // faster than 8 with 65 on a mixed 257^3 solve (more, smaller boxes on
                     // the latency-bound levels)
  int box_rr = 65;   // 3D FP32/FP16 levels of mixed V-cycles with every side <= box_rr form
                     // the residual and its restriction in one launch (k_box.hip); 0: off
  int box_small = 33;  // box levels with every side <= box_small use small boxes: sweeps on
                       // 8 x box_sxy x 4 boxes (with box_z = 4), residual-restriction on
                       // 4 x 4 x 4 coarse boxes (more workgroups, less serial work each, on
                       // the latency-bound levels); 0: off
  int box_sxy = 4;     // y extent of the small sweep boxes (4 or 8)
};
void set_smooth_tuning(const SmoothTuning& t);
SmoothTuning smooth_tuning();
// `sweeps` (1 or 2) red-black sweeps (colour first_color first) of a 3D FP32/FP16 CONST or FLUX
// level on boxes, temporally blocked in LDS (k_box.hip), out of place.  in_mode 0: from x_in;
// 1: from zero (x_in not read); 2: from x_in + factor (* *dev_scale) * the prolongation of the
// coarse correction uc (level `coarse`, precision pc == p), formed on load.  x_out equals what
// prolong_add and rbgs_sweep launches would leave.  Returns false (nothing launched) when the
// level is not one the box kernels take.
bool box_sweeps(const Launch& L, const DevOp& op, Prec p, const void* b, const void* x_in,
                void* x_out, int first_color, int sweeps, int in_mode, const DevOp* coarse,
                Prec pc, const void* uc, double factor, const double* dev_scale);
// restrict_fw of the residual b - A x of a 3D FP32/FP16 CONST or FLUX level in one launch
// (k_box.hip): the fine residual is formed in LDS and never stored.  The coarse values equal
// what residual() then restrict_fw() leave.  Returns false (nothing launched) when the level
// pair is not one the kernel takes.
bool box_resrestrict(const Launch& L, const DevOp& fine, Prec pf, const void* b, const void* x,
                     const DevOp& coarse, Prec pc, void* bc, double scale, const double* dev_scale);
// Number of blocks `partials`-producing kernels use for this op (<= kMaxPartials).
int partial_blocks(const DevOp& op);
// 3D apply / residual as plane marches (k_plane.hip); apply()/residual() route 3D levels there
// unless switched off (A/B timing).
void apply_plane(const Launch& L, const DevOp& op, Prec p, const void* x, void* y, double* partials);
void residual_plane(const Launch& L, const DevOp& op, Prec p, const void* b, const void* x, void* r,
                    double* partials);
int plane_blocks(const DevOp& op);
// CG direction update and apply in one 3D plane march (FP32 CONST / FLUX): p_out = z + beta p_in
// (beta from L.sc, the same FMA as cg_xpay), y = A p_out, partials of p_out.y as apply().  p_out
// must differ from p_in.  Returns false (nothing launched) where the march does not apply.
bool apply_xpay_plane(const Launch& L, const DevOp& op, Prec p, const void* z, const void* p_in,
                      void* p_out, void* y, double* partials);
// Residual replacement of a mixed solve in one 3D plane march (FP64 operator op64): with
// sel = so->xsel, s = so->snorm, x = x[sel] + s e (x[sel] and e read only), the march writes
// x[1 - sel] = x, r64 = b - A x, r32 = r64 / s and the partials of r64.r64; a set gate (done or
// breakdown) skips it.  Every output node is written by one thread and no input is written, so
// neighbouring tiles never see a half-updated solution.  Returns false (nothing launched) when
// the operator is not a 3D plane-march one.
bool replace_plane(const Launch& Lgate, const DevOp& op64, const double* b, double* const x[2],
                   const float* e, double* r64, float* r32, const KScalars* so, double* partials);
// Reference residual norms of a 3D FP64 level as a plane march (interior rows, L u + f with the
// operator's unfolded weights): plane_blocks(op) partials of r.r into sums and of max |r| into
// maxs, *degenerate set by a row whose |C| < 1e-12.  Returns false (nothing launched) when the
// grid is not a 3D plane-march one.
bool resnorm_plane(hipStream_t s, const DevOp& op, const double* u, const double* f, double f_const,
                   double* sums, double* maxs, int* degenerate);
// Partials capacity that covers either routing of this op.
int partial_blocks_max(const DevOp& op);
void set_plane_stencils(bool on);
bool plane_stencils();
// 3D FP32/FP16 plane march: the ring-buffered variant (k_plane3r) or k_plane3.
void set_plane_ring(bool on);
// 3D restrictions of standard coarsenings as a plane march (1, default) or one coarse node per
// thread (0); the same bits either way.
void set_restrict_march(bool on);
bool restrict_march();
bool plane_ring();

// Reference-semantics relaxation (FP64): interior update (-(f + N(u)) - sum w u_nb)/C.
struct NonlinearTerms {
  int n;
  pdehip_nonlinear t[PDEHIP_MAX_NONLINEAR];
};
void jacobi_ref(const Launch& L, const DevOp& op, const double* f, double f_const,
                const NonlinearTerms& nl, const double* x, double* next, double* maxdelta_partials,
                int* degenerate);
void sor_rb_ref(const Launch& L, const DevOp& op, const double* f, double f_const,
                const NonlinearTerms& nl, double* x, double omega, int color,
                double* maxdelta_partials, int* degenerate);
int relax_blocks(const DevOp& op);

// ---- boundary closures (k_aux.hip)
struct FaceData {
  int kind[6];
  const double* value[6];   // per face node
  const double* alpha[6];
  const double* beta[6];
  const double* gamma[6];
  double h[6];
  int convention;
};
// Reference closure order: axis by axis (x faces, y faces, z faces).
// (double: solutions; float: the FP32 state of a time series)
template <typename S> void closures_nr(const Launch& L, const DevOp& op, const FaceData& fd, S* u);
template <typename S> void closures_dirichlet(const Launch& L, const DevOp& op, const FaceData& fd, S* u);
// ---- multigrid transfers
// Coarse right-hand side: full weighting of the fine residual on every coarse node, then the
// transpose of the homogeneous closure extension (coarse.theta) folded in, times scale
// (times *dev_scale when given).  Coarse boundary nodes come out zero.  One pass.
void restrict_fw(const Launch& L, const DevOp& fine, Prec pf, const void* rf, const DevOp& coarse,
                 Prec pc, void* bc, double scale, const double* dev_scale);
// fine interior += factor * trilinear interpolation of the coarse correction, whose boundary
// values are taken as the homogeneous closure extension (theta times the inward neighbour).
void prolong_add(const Launch& L, const DevOp& coarse, Prec pc, const void* uc, const DevOp& fine,
                 Prec pf, void* uf, double factor, const double* dev_scale);
// Coarsest level: x[map] = Ainv * b[map] in FP64.
void coarse_dense(const Launch& L, Prec p, const double* ainv, const int* map, int m,
                  const void* b, void* x);
// kappa restriction (FLUX hierarchy set-up), FP64 in/out.
void restrict_kappa(hipStream_t s, const DevOp& fine, const double* kf, const DevOp& coarse,
                    double* kc);

// ---- vectors and scalars
void fill(hipStream_t s, Prec p, void* x, long long n, double v);
void convert(hipStream_t s, Prec pin, const void* in, Prec pout, void* out, long long n, double scale);
// out = 1/in in FP32 (FP32/FP16 operators keep FLUX conductivities as rho = 1/kappa).
void reciprocal_f32(hipStream_t s, const double* in, float* out, long long n);
// partials of a.b (or a.a when b == a)
void dot(const Launch& L, Prec p, const void* a, const void* b, long long n, double* partials);
int vec_blocks(long long n);

// Finalize step codes: sum the partials and update the scalars (single block).
enum FinOp {
  FIN_CG_ALPHA = 0,    // pq = sum; alpha = rz/pq
  FIN_CG_RR = 1,       // rr = sum; metric = sqrt(rr); done = metric < target
  FIN_CG_BETA = 2,     // rz_new = sum; beta = rz_new/rz; rz = rz_new
  FIN_CG_INIT = 3,     // rz = sum (first r.z)
  FIN_R0 = 4,          // r0 = sqrt(sum); target from tol; metric = r0; done if r0 < target
  FIN_STORE = 5,       // out[slot] = sum
  FIN_MAX = 6,         // out[slot] = max
  FIN_RELAX = 7,       // metric = max; iter++; done = metric < tol (relaxation stop rule)
  // mixed-precision CG with residual replacement (inner scalars, outer scalars)
  FIN_RU_CHECK = 8,    // as FIN_CG_RR without a stop; flag_s = replacement due
  FIN_RU_APPLY = 9,    // after a replacement: outer metric = sqrt(sum), stop on the true residual
  FIN_RU_START = 10,   // outer snorm = outer metric (the inner scale); inner done = outer done
  FIN_RU_APPLY_PP = 11,  // FIN_RU_APPLY after a fused replacement: a replacement that ran also
                         // switches the current FP64 solution buffer (xsel) and voids the FP32
                         // correction (fresh) until the next CG update
};
void finalize(const Launch& L, KScalars* sc, const double* partials, int nparts, int op,
              double tol, int tol_mode, double* out_slot);

// CG updates: x += alpha p, r -= alpha q, partials of r.r (and max|r| into maxpart if non-null).
// With `fresh` non-null and *fresh set, x is taken as zero (x = alpha p): the FP32 correction
// of a mixed solve after a fused replacement.
void cg_update(const Launch& L, Prec p, void* x, void* r, const void* pv, const void* q,
               long long n, double* partials, double* maxpart, const int* fresh = nullptr);
// p = z + beta p
void cg_xpay(const Launch& L, Prec p, const void* z, void* pv, long long n);
// y += a x with a from device scalar (mode) or constant
void axpy_dev(const Launch& L, Prec p, void* y, const void* x, long long n, const double* a, double sign);

}  // namespace pdehip
