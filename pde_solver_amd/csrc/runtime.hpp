// runtime.hpp — host runtime of libpdehip: device context, HBM-budgeted buffers, problem
// assembly, multigrid hierarchy and the solver drivers behind the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <memory>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "kernels_ext.hpp"
#include "kernels_krylov.hpp"

namespace pdehip {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char* what);
#define HIPCHK(x) ::pdehip::hip_check((x), #x)

// ------------------------------------------------------------------ context
// Mixed precision: coarse multigrid levels stored in FP16 (default) or FP32; read when a plan
// is created.
void set_mixed_fp16(bool on);
bool mixed_fp16();
// mixed MG-PCG on 3D plane-march levels: residual replacement in one launch (else three)
void set_fused_replace(bool on);
bool fused_replace();
// mixed MG-PCG on 3D FP32 plane-march levels: direction update p = z + beta p inside the apply
void set_fused_xpay(bool on);
bool fused_xpay();

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  size_t budget = 0;                 // bytes this context may allocate
  std::atomic<size_t> used{0};
  std::atomic<size_t> peak{0};
  int cus = 0;
  std::string arch;
  std::mutex mu;                     // serialises one-shot solves on the context stream
  std::string last_error;
  std::mutex err_mu;
  std::mutex tune_mu;                // batch work-item shapes picked by the autotuner
  std::map<std::string, std::vector<int>> batch_shapes;
  // Released device blocks kept for reuse by size (up to pool_cap bytes; 0 = no pooling).
  // Every buffer of a context is used on its stream only, so a block handed out again is
  // ordered after its previous user's work; hipFree would synchronise the whole device.
  std::mutex pool_mu;
  std::map<size_t, std::vector<void*>> pool;
  size_t pooled = 0;
  size_t pool_cap = 0;
  void trim_pool();
  ~Ctx() { trim_pool(); }
  void set_error(const std::string& e) {
    std::lock_guard<std::mutex> g(err_mu);
    last_error = e;
  }
};

// Device buffer charged against the context budget.
class DBuf {
 public:
  DBuf() = default;
  DBuf(Ctx* c, size_t bytes) { alloc(c, bytes); }
  ~DBuf() { release(); }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept { *this = std::move(o); }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) {
      release();
      ctx_ = o.ctx_; p_ = o.p_; bytes_ = o.bytes_;
      o.p_ = nullptr; o.bytes_ = 0; o.ctx_ = nullptr;
    }
    return *this;
  }
  void alloc(Ctx* c, size_t bytes);
  void release();
  void* get() const { return p_; }
  template <typename T> T* as() const { return static_cast<T*>(p_); }
  size_t bytes() const { return bytes_; }
  explicit operator bool() const { return p_ != nullptr; }

 private:
  Ctx* ctx_ = nullptr;
  void* p_ = nullptr;
  size_t bytes_ = 0;
};

// ------------------------------------------------------------------ geometry
struct Geo {
  int nx = 0, ny = 0, nz = 1;
  bool is3d = false;
  double xmin = 0, xmax = 1, ymin = 0, ymax = 1, zmin = 0, zmax = 1;
  double dx = 0, dy = 0, dz = 1;
  long long n = 0;
  long long sxy() const { return (long long)nx * ny; }
};
Geo make_geo(const pdehip_problem& p);
Geo coarsen(const Geo& f);
int level_count(const Geo& g, int max_levels);

// Host copy of a face's boundary data sampled per face node.
struct HostFaces {
  int kind[6] = {0, 0, 0, 0, 0, 0};
  std::vector<double> value[6], alpha[6], beta[6], gamma[6];
  double h[6] = {1, 1, 1, 1, 1, 1};
  int convention = 0;
};
HostFaces sample_faces(const pdehip_problem& p, const Geo& g);
// Folded closure factor theta of each face at the level's spacing.
void fold_thetas(const pdehip_problem& p, const Geo& g, bool folded, double theta[6]);
// CONST weights (C, W, E, S, N, B, F) at a level's spacing.
void const_weights(const pdehip_problem& p, const Geo& g, double w[7]);

// Device copy of HostFaces.
struct DevFaces {
  DBuf buf;
  FaceData fd{};
  void upload(Ctx* c, const HostFaces& h, const Geo& g, hipStream_t s);
};

// ------------------------------------------------------------------ multigrid level
struct Level {
  Geo g;
  DevOp op64{}, op32{};
  DBuf kap64, kap32, pw;
  DBuf u, b, r;
  DBuf alt;               // second smoother buffer (out-of-place fused sweeps)
  Prec prec = P64;
  double sigma = 1.0;     // stored correction = u / sigma (mixed, P16 levels)
  DBuf ainv, map;         // coarsest level dense inverse
  int m = 0;
};

// ------------------------------------------------------------------ plan
struct Plan {
  Ctx* ctx = nullptr;
  pdehip_problem prob{};   // host pointers valid only during creation
  pdehip_options opt{};
  Geo g;
  bool folded = false;
  bool mixed = false;
  HostFaces hfaces;
  DevFaces faces;
  NonlinearTerms nl{};
  DBuf f;                  // per-node source (or empty: constant prob.f)
  double f_const = 0.0;
  DBuf bvec, x, x0;        // FP64 right-hand side, solution, initial guess (folded: boundary data)
  DBuf bnd;                // reference-row plans: Dirichlet datum for rebuilding b (on demand)
  std::vector<Level> L;    // L[0] is the fine level (operator for every method)
  // Krylov work vectors
  DBuf r, p, q, z, w1, w2, w3, w4;
  DBuf r32, x32, p32, q32, z32;
  DBuf xb;                         // mixed MG-PCG: the second FP64 solution buffer of the fused replacement
  DBuf p32b;                       // ... 3D: the second direction buffer of the fused update + apply
  bool k32 = false;                // mixed BiCGStab/GMRES: FP32 Krylov vectors, FP64 x and residual
  DBuf v32[8], V32;                // their FP32 vectors (GMRES: basis in V32)
  DBuf V, H, cs, sn, gv, yv, hcb;   // GMRES (hcb: two staged Hessenberg columns)
  DBuf part, part2, part3;     // reduction partials
  int npart = 0;
  DBuf scal;                   // KScalars x 2 (outer/inner)
  KScalars* hsc = nullptr;     // pinned host mirror (3 scalar blocks, then 2 x 3 trailing-check slots)
  double* hstage = nullptr;    // pinned staging of the report's reductions
  bool fin_resid = false;      // plan_finish_enqueue staged the true-residual partials
  bool resid_fresh = false;    // the solver left r = b - A x of the final x and its partials
  // set by solvers that leave these decisions on the device: the final true residual runs
  // gated by resid_gate (skipped when the solver left it), and the report's counts and status
  // come from the scalar blocks staged with the report (outer block 0, inner block 1)
  KScalars* resid_gate = nullptr;
  bool dev_report = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t evq[2] = {nullptr, nullptr};   // trailing convergence checks

  ~Plan();
  KScalars* sc(int i = 0) const { return scal.as<KScalars>() + i; }
  hipStream_t stream() const { return ctx->stream; }
  const DevOp& op() const { return L[0].op64; }
};

std::unique_ptr<Plan> plan_create(Ctx* ctx, const pdehip_problem& p, const pdehip_options& o);
// Replace the source (f_host: per-node field, or NULL for the constant f_const) and rebuild b.
void plan_set_source(Plan& P, const double* f_host, double f_const);
void plan_solve(Plan& P, pdehip_report* rep);
// The tail of plan_solve split in two: enqueue the post-solve work (true residual, boundary
// reconstruction, residual norms, staging copies) without waiting, then fill the report after
// the stream is synchronised.  Lets a batch finish many plans with one synchronisation.
void plan_finish_enqueue(Plan& P);
void plan_finish_collect(Plan& P, pdehip_report* rep);
// Scalars of a plan (device) and the plan's KScalars initialisation, for batched drivers.
void plan_init_scalars(Plan& P, int i, double eps);
void plan_download(Plan& P, double* u_out);
// Device bytes the plan's own buffers hold now.
size_t plan_bytes(const Plan& P);
void plan_residual_norms(Plan& P, const double* u_dev, double* l2, double* linf);

// Allocation-free residual norms (noalloc.hip): looks up the static scratch array's address
// on `device` (call with that device current), and the entry itself.
void noalloc_init(int device);
void residual_norms_noalloc(Ctx* ctx, const pdehip_problem& p, const double* u_dev, double* l2,
                            double* linf);

// Standalone entry points.
void op_apply(Ctx* ctx, const pdehip_problem& p, const double* x, double* y);
void time_series(Ctx* ctx, const pdehip_problem& p, const pdehip_time_config& tc,
                 pdehip_frame_fn cb, void* user, double* u_out, pdehip_report* rep);
double time_kernel(Plan& P, int kernel, int precision, int reps, double* bytes);

}  // namespace pdehip
