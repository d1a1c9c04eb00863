// batch.hpp — batched multigrid-preconditioned CG for many independent 2D problems.
//
// Each problem keeps its own resident Plan (operators, hierarchy, boundary data, dense coarse
// inverse, work vectors); the batch engine drives all of them in lock step with batched
// launches: every operation of a PCG iteration (operator apply, vector updates, each level's
// smoother sweeps, residual, restriction, prolongation, coarse solve, reductions) is ONE
// launch whose work items cover every unfinished problem.  Work items are wavefront-sized
// (a strip or a block of rows of one problem), so small and large grids share a launch and
// the chip fills even when every single problem is small.  Each problem has its own device
// scalars; kernels skip the items of problems that have converged, and the reductions are
// segmented per problem (items of a problem are contiguous) and deterministic.
#pragma once

#include <functional>
#include <memory>
#include <vector>

#include "runtime.hpp"

namespace pdehip {

// One (problem, level) entry of the device table.
struct BEnt {
  DevOp op;          // the level's operator in the path's precision
  void* u;           // level correction (level 0: the preconditioner output z)
  void* b;           // level right-hand side (level 0: the preconditioner input r)
  void* r;           // level residual
  void* alt;         // second smoother buffer
  KScalars* sc;      // the problem's scalars of the current Krylov loop (inner or outer)
  double pfac;       // prolongation factor into this level: sigma(coarse) / sigma(this)
  int pad;
};

// Level-0 vectors of one problem.
struct BVec {
  void *x, *r, *p, *q, *z;       // Krylov vectors in the inner precision
  double *x64, *r64, *b64, *x0;  // FP64 outer vectors (mixed) / solution (FP64)
  KScalars* so;                  // outer scalars (FP64 refinement; FP64 path: the CG scalars)
  KScalars* si;                  // inner scalars (mixed)
  long long n;
  int nx, ny;
  DevOp op64;                    // level-0 FP64 operator (outer residual)
};

// Level-0 vectors of one problem of a batched BiCGStab (reference BiCGStabSolve semantics:
// the residual is recomputed from the FP64 solution every iteration).  Krylov vectors are FP64,
// or FP32 in mixed precision; the solution, right-hand side and residual arithmetic stay FP64.
struct BKVec {
  void *r, *rh, *p, *v, *s, *t, *ph, *sh;   // ph / sh alias p / s without a preconditioner
  double *x, *b, *x0;
  KScalars* sc;
  DevOp op;       // the applies' operator (FP32 difference form in mixed precision)
  DevOp op64;     // FP64 operator: residual and Jacobi diagonal
  int nx, ny, jacobi, pad;
};

// Level-0 data of one problem of a batched GMRES(m) (reference GMRESSolve, krylov.cpp:187-333,
// in the single-problem path's form: classical Gram-Schmidt applied twice, right Jacobi
// preconditioning optional).  FP64 throughout: the Arnoldi basis V ((m+1) vectors of n), the
// Hessenberg matrix H ((m+1) x m, leading dimension m), the Givens rotations, the least-squares
// right-hand side g, the solution y, and the two CGS passes' coefficients hc (m+1 apart).
struct BGVec {
  double *V, *H, *cs, *sn, *g, *y, *hc;
  double *x, *r, *b, *x0, *w;
  KScalars* sc;
  DevOp op;       // FP64 operator (folded rows)
  long long n;
  int nx, ny, jacobi, pad;
};

// Finish data of one problem: boundary data, source.
struct BFace { FaceData fd; const double* f; double f_const; };

// Per-problem results of a batched solve, gathered on the device and copied once.
struct BReport {
  double rr;          // |b - A x|^2 of the solved (folded) system
  double l2sq, linf;  // reference residual norms (squared L2, max)
  double metric, r0;  // outer scalars
  int done, breakdown, iter, steps, inner, pad;
};

// Work item: problem p and a range (meaning depends on the list).
struct BItem { int p, a, b, c; };

// Work-item shape of one level: strip height in rows of the smoother and the residual (and of
// the level-0 operator applies), nodes per prolongation row block, coarse nodes per restriction
// block (whole rows, at least one).
struct LevelShape { int ychunk = 16; int row_nodes = 256; int crow_nodes = 128; };

struct BatchPlan;
struct BatchDeleter { void operator()(BatchPlan* b) const; };
using BatchPtr = std::unique_ptr<BatchPlan, BatchDeleter>;
// plans (optional): resident plans of the problems, created on ctx with opt (moved into the
// batch where non-null; the rest are created).  batch_release_plans gives them back.
BatchPtr batch_create(Ctx* ctx, int count, const pdehip_problem* probs, const pdehip_options& opt,
                      const pdehip_batch_config* cfg,
                      std::vector<std::unique_ptr<Plan>>* plans = nullptr);
// The batch's plans, in problem order, after its work has finished (the batch keeps none).
std::vector<std::unique_ptr<Plan>> batch_release_plans(BatchPlan& B);
// Called on the host after each polling chunk of a solve has been enqueued (for example to
// cancel problems whose deadline passed); the stream order puts anything it enqueues between
// that chunk and the next.
void batch_set_chunk_hook(BatchPlan& B, std::function<void()> hook);
// Stop problem i from the next enqueued launch on: its scalars are marked broken down, so every
// later kernel skips it and its report says PDEHIP_ERR_BREAKDOWN.
void batch_cancel(BatchPlan& B, int i);
// Returns PDEHIP_OK, or the status of the first problem that did not converge.  MG-PCG batches
// run the lock-step MG-PCG; BiCGStab batches the lock-step BiCGStab.
int batch_solve(BatchPlan& B, pdehip_report* reports);
void batch_download(BatchPlan& B, int i, double* u_out);
// Mean duration (ms) of one launch of a batched level-0 kernel over every problem (HIP events
// on the batch stream): 0 smoother sweep pair, 1 residual, 2 Krylov apply + dot, 3 FP64
// residual.  algorithmic_bytes: compulsory bytes of one launch.
double batch_time_kernel(BatchPlan& B, int kernel, int reps, double* algorithmic_bytes);
size_t batch_bytes(const BatchPlan& B);
void batch_destroy(BatchPtr& B);
// True when the batched engine handles this set: 2D constant-coefficient linear operators, and
// MG-PCG without convection, BiCGStab (no preconditioner or Jacobi), or FP64 GMRES(m) (no
// preconditioner or Jacobi).
bool batch_eligible(int count, const pdehip_problem* probs, const pdehip_options& opt,
                    std::string* why);

// True when a plan built for this batch's options can join it (MG-PCG plans need a dense
// coarsest level).
bool batch_plan_ok(const Plan& P, const pdehip_options& opt);

// ---- batched kernels (k_batch.hip)
namespace bk {
// Smoother: one red-black sweep of every listed (problem, strip, row range), out of place
// between u and alt (dir 0: u -> alt, 1: alt -> u).  zero: u is zero (skip its loads).
void sweep(hipStream_t s, Prec p, const BEnt* tab, const BItem* items, int n, int c1, int dir,
           bool zero);
// r = b - A u over listed (problem, strip, row range) items (the smoother's items).
void residual(hipStream_t s, Prec p, const BEnt* tab, const BItem* items, int n);
// Coarse right-hand side (FW + folded closure transpose) for listed coarse (problem, row range).
void restrict_(hipStream_t s, Prec pf, Prec pc, const BEnt* fine, const BEnt* coarse,
               const BItem* items, int n, bool scale0);
// Fine += pfac * interpolate(closure-extended coarse) for listed fine (problem, row range).
void prolong(hipStream_t s, Prec pc, Prec pf, const BEnt* coarse, const BEnt* fine,
             const BItem* items, int n, bool scale0);
// Coarsest-level dense solve: x[map] = Ainv b[map], one wavefront per (problem, row) item.
struct BDense { const double* ainv; const int* map; int m; int pad; };
void coarse_dense(hipStream_t s, Prec p, const BEnt* tab, const BDense* dense, const BItem* items,
                  int n);
// Level-0 Krylov vector kernels over (problem, element range) items; partials per item.
// V_CG_UPDATE of an inner (mixed) solve adds |r64| alpha p to the FP64 solution directly.
enum VecOp { V_APPLY_DOT = 0, V_CG_UPDATE = 1, V_DOT_RZ = 2, V_XPAY = 3, V_RESID64 = 4,
             V_TO32 = 5, V_COPY_ZP = 7, V_ZERO_X = 8, V_COPY_X0 = 9, V_RR = 10 };
// (V_APPLY_DOT and V_RESID64 are row kernels: apply_dot_tab.)
void vec(hipStream_t s, Prec p, int op, const BVec* vecs, const BItem* items, int n,
         double* partials, double* maxpart);
// Level-0 strip items: inner q = A p with partials p.q (operator and scalars of tab0), or the
// outer FP64 residual r64 = b64 - A64 x64 with partials r.r (all: also for finished problems).
void apply_dot_tab(hipStream_t s, Prec p, const BVec* vecs, const BEnt* tab0, const BItem* rows,
                   int n, double* partials, bool outer, bool all = false);
// Inner scalars reset for a refinement (outer metric kept as the correction scale).
void reset_inner(hipStream_t s, const BVec* vecs, int nprob);
// Mixed solve with residual replacement: inner scalars fresh, inner scale = initial residual.
void ru_start(hipStream_t s, const BVec* vecs, int nprob);
// Residual replacement over level-0 strip items for problems whose inner scalars flag it:
// r32 = (b64 - A64 x64) / snorm, partials |b - A x|^2.
void replace_residual(hipStream_t s, const BVec* vecs, const BItem* rows, int n, double* partials);
void init_scalars(hipStream_t s, const BVec* vecs, int nprob, double eps);
// Segmented finalize: problem q sums partials[first[q] .. first[q]+cnt[q]) and applies FinOp
// `op` to its scalars (inner when inner != 0).
void finalize(hipStream_t s, const BVec* vecs, const int* first, const int* cnt, int nprob,
              const double* partials, const double* maxpart, int op, int inner, double tol,
              int tol_mode);
// Count of problems whose selected scalars are not done (written to *active).
void active_count(hipStream_t s, const BVec* vecs, int nprob, int inner, int* active);
// Problem p's outer and inner scalars marked broken down with no restart pending.
void cancel(hipStream_t s, const BVec* vecs, int p);
// Finish: boundary closures of every problem's FP64 solution (phase 0/1: Neumann/Robin x/y
// faces, 2/3: Dirichlet x/y faces, launched in that order), the reference residual norms over
// element items (partials per item), and the per-problem gather into BReport.
void closures(hipStream_t s, const BVec* vecs, const BFace* faces, int nprob, int phase);
void resnorm(hipStream_t s, const BVec* vecs, const BFace* faces, const BItem* items, int n,
             double* sums, double* maxs);
void finish_reduce(hipStream_t s, const BVec* vecs, const int* rfirst, const int* rcnt,
                   const double* rr_part, const int* efirst, const int* ecnt, const double* sums,
                   const double* maxs, int nprob, BReport* out);
// ---- batched BiCGStab (k_batch.hip).  Row-block items (p, j0, j1, 0) for the vector kernels,
// level-0 strip items for the applies; partials per item, segmented finalize per problem.
enum KkVec { KK_RH = 0, KK_RHO = 1, KK_P = 2, KK_S = 3, KK_XS = 4, KK_X = 5, KK_RESTART = 6,
             KK_X0 = 7 };
// rh = r / partials rh.r / p = r + beta (p - omega v), ph = D^-1 p / s = r - alpha v,
// sh = D^-1 s, partials s.s / x += alpha ph (problems whose |s| met the target) /
// x += alpha ph + omega sh / rh = r for problems with a restart pending after a breakdown /
// x = x0.
void kvec(hipStream_t s, Prec p, int op, const BKVec* vecs, const BItem* items, int n,
          double* partials);
// mode 0: v = A ph, partials rh.v; 1: t = A sh, partials t.t and t.s (skipped for problems whose
// |s| met the target); 2: r = b - A64 x in FP64, partials |r|^2.
void kapply(hipStream_t s, Prec p, int mode, const BKVec* vecs, const BItem* items, int n,
            double* p1, double* p2);
// KK_FR with tol > 0: stall window in iterations (see kk_finalize); KK_R0K: the completion
// pass's start, keeping the first pass's r0 and target.
enum KkFin { KK_R0 = 0, KK_FRHO = 1, KK_FALPHA = 2, KK_FS = 3, KK_FOMEGA = 4, KK_FR = 5,
             KK_R0K = 6 };
// FP32 pass of a mixed BiCGStab batch: iterations without a halving of the true residual after
// which a problem leaves the pass for the FP64 completion.
constexpr int kMixedStallWindow = 100;
// Scalars of a fresh start for problems whose restart is pending (after KK_RESTART).
void krestart(hipStream_t s, const BKVec* vecs, int nprob);
// Mixed batches: every problem that did not converge gets fresh scalars (eps) and steps = -1,
// keeping r0 and the target, for the FP64 completion pass; converged problems keep theirs.
void kreset_failed(hipStream_t s, const BKVec* vecs, int nprob, double eps);
// Count of problems that did not converge, written to *out.
void kfailed(hipStream_t s, const BKVec* vecs, int nprob, int* out);
void kfinalize(hipStream_t s, const BKVec* vecs, const int* first, const int* cnt, int nprob,
               const double* p1, const double* p2, int op, double tol, int tol_mode);
// ---- batched GMRES(m) (k_batch.hip).  Element items for the vector kernels, level-0 strips
// for the applies; `col` is the Arnoldi step of the current cycle, the same for every problem.
// KG_DOT: partials[i * n + item] = V_i . w for i <= col (launched over groups of 8 vectors);
// KG_AXPY1 / KG_AXPY2: w -= sum_i hc_pass[i] V_i (AXPY2 also writes the partials w . w to
// partials2); KG_NORMALIZE: V_{col+1} = w / |w| (|w| from hc[2 m + 1]); KG_START: V_0 = r / |r|;
// KG_UPDATE: x += M^-1 sum_{i < steps} y_i V_i.
enum KgVec { KG_DOT = 0, KG_AXPY1 = 1, KG_AXPY2 = 2, KG_NORMALIZE = 3, KG_START = 4, KG_UPDATE = 5 };
// The vector kernels' items are element ranges of at most this many nodes.
constexpr int kGmresItemNodes = 512;
void gvec(hipStream_t s, int op, const BGVec* vecs, const BItem* items, int n, int col, int mm,
          double* partials, double* partials2);
// mode 0: w = A M^-1 V_col; mode 1: r = b - A x, partials |r|^2 per item.
void gapply(hipStream_t s, int mode, const BGVec* vecs, const BItem* items, int n, int col,
            double* partials);
// Segmented finalize per problem: GF_H1 / GF_H2 sum the CGS coefficients of step col (GF_H2
// also sets H[i][col] = h1[i] + h2[i]); GF_NORM sets H[col+1][col] = |w| and applies the Givens
// step; GF_R0 / GF_R: the (initial) true residual, target and stop test.
enum KgFin { GF_H1 = 0, GF_H2 = 1, GF_NORM = 2, GF_R0 = 3, GF_R = 4 };
void gfinalize(hipStream_t s, const BGVec* vecs, const int* first, const int* cnt, int nitems,
               int nprob, const double* partials, int op, int col, int mm, double tol,
               int tol_mode);
// GC_START: a new cycle (g = [|r|, 0, ...], H = 0, cycle flags cleared); GC_BACKSUB: y from the
// triangular system of the cycle's steps.
enum KgCycle { GC_START = 0, GC_BACKSUB = 1 };
void gcycle(hipStream_t s, const BGVec* vecs, int nprob, int op, int mm);
}  // namespace bk

}  // namespace pdehip
