// hip_solver.cpp — the C++ plug-in of the HIP backend (include/hip_solver.h): converts a
// SolveInput into the C ABI's problem description, samples the LaTeX fields on the host (as
// the CUDA plug-in does, src/backends/cuda/solvers/relaxation.cu:186-253), calls
// libpdehip and fills SolveOutput.  Shaped like SolvePDECuda (src/backends/cuda/CudaSolve.cu:31-65).
#include "hip_solver.h"

#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "pdehip.h"

#ifdef PDEHIP_REFERENCE_EVALUATOR
#include "expression_eval.h"
namespace {
struct Sampler {
  ExpressionEvaluator ev;
  explicit Sampler(const std::string& s) : ev(ExpressionEvaluator::ParseLatex(s)) {}
  bool ok() const { return ev.ok(); }
  std::string error() const { return ev.error(); }
  double operator()(double x, double y, double z, double t) const { return ev.Eval(x, y, z, t); }
};
}  // namespace
#else
#include "latex_expr.h"
namespace {
struct Sampler {
  pdehip_host::LatexExpr ev;
  explicit Sampler(const std::string& s) : ev(pdehip_host::LatexExpr::Parse(s)) {}
  bool ok() const { return ev.ok(); }
  std::string error() const { return ev.error(); }
  double operator()(double x, double y, double z, double t) const { return ev.Eval(x, y, z, t); }
};
}  // namespace
#endif

namespace {

std::mutex g_ctx_mu;
pdehip_ctx* g_ctx = nullptr;

std::mutex g_opt_mu;
HipSolverOptions g_opt;

pdehip_ctx* context(std::string* err) {
  std::lock_guard<std::mutex> g(g_ctx_mu);
  if (!g_ctx) {
    const int st = pdehip_create(0, 0, &g_ctx);
    if (st != PDEHIP_OK) {
      if (err) *err = std::string("HIP context creation failed: ") + pdehip_last_error(nullptr);
      g_ctx = nullptr;
    }
  }
  return g_ctx;
}

struct Grid {
  int nx, ny, nz;
  double dx, dy, dz;
  size_t n;
  double X(int i, const Domain& d) const { return d.xmin + i * dx; }
  double Y(int j, const Domain& d) const { return d.ymin + j * dy; }
  double Z(int k, const Domain& d) const { return d.zmin + k * dz; }
};

// Owned arrays behind a pdehip_problem.
struct Assembled {
  pdehip_problem p{};
  std::vector<double> coef[PDEHIP_NCOEF];
  std::vector<double> f, initial, velocity;
  std::vector<double> face[6][4];
  std::string error;
};

// BoundaryCondition::Expression -> pdehip_expr: linear form, or per-node samples of the LaTeX
// with the linear fallback of EvalExpr (reference cpu_utils.cpp:521-530).
bool convert_expr(const BoundaryCondition::Expression& e, const Grid& g, const Domain& d, int face,
                  pdehip_expr* out, std::vector<double>* store, std::string* err) {
  out->c = e.constant;
  out->cx = e.x;
  out->cy = e.y;
  out->cz = e.z;
  out->samples = nullptr;
  if (e.latex.empty()) return true;
  Sampler s(e.latex);
  const bool ok = s.ok();
  const size_t m = face < 2 ? (size_t)g.ny * g.nz : (face < 4 ? (size_t)g.nx * g.nz : (size_t)g.nx * g.ny);
  store->resize(m);
  for (size_t q = 0; q < m; ++q) {
    double x, y, z;
    if (face < 2) { x = face == 0 ? d.xmin : d.xmax; y = g.Y(q % g.ny, d); z = g.nz > 1 ? g.Z(q / g.ny, d) : 0.0; }
    else if (face < 4) { x = g.X(q % g.nx, d); y = face == 2 ? d.ymin : d.ymax; z = g.nz > 1 ? g.Z(q / g.nx, d) : 0.0; }
    else { x = g.X(q % g.nx, d); y = g.Y(q / g.nx, d); z = face == 4 ? d.zmin : d.zmax; }
    double v = ok ? s(x, y, z, 0.0) : NAN;
    if (!std::isfinite(v)) v = e.constant + e.x * x + e.y * y + e.z * z;
    (*store)[q] = v;
  }
  out->samples = store->data();
  (void)err;
  return true;
}

bool sample_field(const std::string& latex, double constant, const Grid& g, const Domain& d,
                  std::vector<double>* out, std::string* err, const char* what) {
  Sampler s(latex);
  if (!s.ok()) {
    *err = std::string("invalid ") + what + " expression: " + s.error();
    return false;
  }
  out->resize(g.n);
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const double v = s(g.X(i, d), g.Y(j, d), g.nz > 1 ? g.Z(k, d) : 0.0, 0.0);
        (*out)[((size_t)k * g.ny + j) * g.nx + i] = std::isfinite(v) ? v : constant;
      }
  return true;
}

bool assemble(const SolveInput& in, const HipSolverOptions& hopt, Assembled* A) {
  const Domain& d = in.domain;
  if (d.nx < 3 || d.ny < 3) { A->error = "grid must be at least 3x3"; return false; }
  if (d.nz > 1 && d.nz < 3) { A->error = "grid must be at least 3x3x3"; return false; }
  if (d.xmax <= d.xmin || d.ymax <= d.ymin || (d.nz > 1 && d.zmax <= d.zmin)) {
    A->error = "domain bounds are invalid";
    return false;
  }
  if (d.coord_system != CoordinateSystem::Cartesian) { A->error = "HIP backend supports Cartesian coordinates only"; return false; }
  if (HasMixedDerivatives(in.pde) || HasHigherOrderDerivatives(in.pde)) {
    A->error = "mixed or higher-order derivatives are not supported on this backend";
    return false;
  }
  if (!in.integrals.empty()) { A->error = "integral terms are not supported"; return false; }
  if (HasImplicitShape(in)) { A->error = "implicit domain shapes are not supported"; return false; }
  if (!in.nonlinear_derivatives.empty()) { A->error = "nonlinear derivative terms are not supported on this backend"; return false; }
  Grid g;
  g.nx = d.nx; g.ny = d.ny; g.nz = std::max(1, d.nz);
  g.dx = (d.xmax - d.xmin) / (g.nx - 1);
  g.dy = (d.ymax - d.ymin) / (g.ny - 1);
  g.dz = g.nz > 1 ? (d.zmax - d.zmin) / (g.nz - 1) : 1.0;
  g.n = (size_t)g.nx * g.ny * g.nz;
  pdehip_problem& p = A->p;
  pdehip_default_problem(&p);
  p.nx = g.nx; p.ny = g.ny; p.nz = g.nz;
  p.xmin = d.xmin; p.xmax = d.xmax; p.ymin = d.ymin; p.ymax = d.ymax; p.zmin = d.zmin; p.zmax = d.zmax;
  const PDECoefficients& c = in.pde;
  p.a = c.a; p.b = c.b; p.az = c.az; p.c = c.c; p.d = c.d; p.dz = c.dz; p.e = c.e; p.f = c.f;
  p.ut = c.ut; p.utt = c.utt;
  p.op_kind = PDEHIP_OP_CONST;
  p.convection = in.solver.advection_scheme == AdvectionScheme::Upwind ? PDEHIP_CONV_UPWIND
                                                                      : PDEHIP_CONV_CENTERED;
  // variable coefficients: EvalCoefficient per node (coefficient_evaluator.h:34)
  const std::string* latex[PDEHIP_NCOEF] = {&c.a_latex, &c.b_latex, &c.az_latex, &c.c_latex,
                                            &c.d_latex, &c.dz_latex, &c.e_latex};
  const double cst[PDEHIP_NCOEF] = {c.a, c.b, c.az, c.c, c.d, c.dz, c.e};
  const char* names[PDEHIP_NCOEF] = {"a", "b", "az", "c", "d", "dz", "e"};
  for (int m = 0; m < PDEHIP_NCOEF; ++m) {
    if (latex[m]->empty()) continue;
    if (!sample_field(*latex[m], cst[m], g, d, &A->coef[m], &A->error, names[m])) return false;
    p.coef[m] = A->coef[m].data();
    p.op_kind = PDEHIP_OP_POINTWISE;
  }
  // RHS f(x,y,z) at t = 0 (BuildRhs2D, solver.cpp:193-202)
  if (!c.rhs_latex.empty()) {
    if (!sample_field(c.rhs_latex, c.f, g, d, &A->f, &A->error, "RHS")) return false;
    p.f_field = A->f.data();
  }
  // boundary conditions
  const BoundaryCondition* faces[6] = {&in.bc.left, &in.bc.right, &in.bc.bottom, &in.bc.top,
                                       &in.bc.front, &in.bc.back};
  const std::vector<PiecewiseBoundaryCondition>* pw[6] = {&in.bc.left_piecewise, &in.bc.right_piecewise,
      &in.bc.bottom_piecewise, &in.bc.top_piecewise, &in.bc.front_piecewise, &in.bc.back_piecewise};
  for (int f = 0; f < 6; ++f) {
    if (!pw[f]->empty()) { A->error = "piecewise boundary conditions are not supported"; return false; }
    const BoundaryCondition& bc = *faces[f];
    switch (bc.kind) {
      case BCKind::Dirichlet: p.bc[f].kind = PDEHIP_BC_DIRICHLET; break;
      case BCKind::Neumann: p.bc[f].kind = PDEHIP_BC_NEUMANN; break;
      case BCKind::Robin: p.bc[f].kind = PDEHIP_BC_ROBIN; break;
      default:
        if (f < 4 || g.nz > 1) { A->error = "hyperbolic boundary kinds are not supported on this backend"; return false; }
        p.bc[f].kind = PDEHIP_BC_DIRICHLET;
    }
    if (f >= 4 && g.nz <= 1) continue;
    const BoundaryCondition::Expression* ex[4] = {&bc.value, &bc.alpha, &bc.beta, &bc.gamma};
    pdehip_expr* dst[4] = {&p.bc[f].value, &p.bc[f].alpha, &p.bc[f].beta, &p.bc[f].gamma};
    for (int t = 0; t < 4; ++t)
      if (!convert_expr(*ex[t], g, d, f, dst[t], &A->face[f][t], &A->error)) return false;
  }
  p.bc_convention = hopt.outward_robin ? PDEHIP_BCCONV_OUTWARD : PDEHIP_BCCONV_AXIS;
  p.n_nonlinear = 0;
  for (const NonlinearTerm& t : in.nonlinear) {
    if (p.n_nonlinear >= PDEHIP_MAX_NONLINEAR) { A->error = "too many nonlinear terms"; return false; }
    pdehip_nonlinear& q = p.nonlinear[p.n_nonlinear++];
    q.kind = static_cast<int>(t.kind);
    q.coeff = t.coeff;
    q.power = t.power;
  }
  if (!in.initial_grid.empty()) {
    if (in.initial_grid.size() != g.n) { A->error = "initial grid size mismatch"; return false; }
    A->initial = in.initial_grid;
    p.initial = A->initial.data();
  }
  if (!in.initial_velocity.empty()) {
    if (in.initial_velocity.size() != g.n) { A->error = "initial velocity size mismatch"; return false; }
    A->velocity = in.initial_velocity;
    p.initial_velocity = A->velocity.data();
  }
  return true;
}

bool all_dirichlet(const pdehip_problem& p) {
  for (int f = 0; f < (p.nz > 1 ? 6 : 4); ++f)
    if (p.bc[f].kind != PDEHIP_BC_DIRICHLET) return false;
  return true;
}

// Mirrors SolveInput.cancel (std::atomic<bool>) into the C ABI's int flag while a call runs.
class CancelBridge {
 public:
  explicit CancelBridge(std::atomic<bool>* src) : src_(src) {
    if (src_) worker_ = std::thread([this] {
      while (!stop_.load()) {
        if (src_->load()) flag_ = 1;
        std::this_thread::sleep_for(std::chrono::milliseconds(1));
      }
    });
  }
  ~CancelBridge() {
    stop_ = true;
    if (worker_.joinable()) worker_.join();
  }
  const volatile int* flag() const { return src_ ? &flag_ : nullptr; }

 private:
  std::atomic<bool>* src_;
  volatile int flag_ = 0;
  std::atomic<bool> stop_{false};
  std::thread worker_;
};

const char* method_failure(SolveMethod m) {
  switch (m) {
    case SolveMethod::CG:
    case SolveMethod::BiCGStab:
    case SolveMethod::GMRES: return "HIP Krylov solver failed to converge";
    case SolveMethod::MultigridVcycle: return "HIP multigrid solver failed to converge";
    default: return "HIP relaxation solver failed";
  }
}

}  // namespace

void HipSetSolverOptions(const HipSolverOptions& options) {
  std::lock_guard<std::mutex> g(g_opt_mu);
  g_opt = options;
}

HipSolverOptions HipGetSolverOptions() {
  std::lock_guard<std::mutex> g(g_opt_mu);
  return g_opt;
}

bool HipWriteVtk(const std::string& path, const Domain& domain, const std::vector<double>& grid,
                 int digits, std::string* error) {
  auto fail = [&](const std::string& m) {
    if (error) *error = m;
    return false;
  };
  const int nx = domain.nx, ny = domain.ny, nz = std::max(1, domain.nz);
  const size_t n = static_cast<size_t>(nx) * ny * nz;
  if (nx < 1 || ny < 1 || grid.size() != n) return fail("grid size does not match domain");
  std::FILE* f = std::fopen(path.c_str(), "w");
  if (!f) return fail("failed to open output file");
  const int d = std::max(1, std::min(digits, 17));
  auto step = [](double lo, double hi, int m) { return m > 1 ? (hi - lo) / (m - 1) : 1.0; };
  std::fprintf(f, "# vtk DataFile Version 3.0\npde_solution\nASCII\nDATASET STRUCTURED_POINTS\n");
  std::fprintf(f, "DIMENSIONS %d %d %d\n", nx, ny, nz);
  std::fprintf(f, "ORIGIN %.17g %.17g %.17g\n", domain.xmin, domain.ymin, domain.zmin);
  std::fprintf(f, "SPACING %.17g %.17g %.17g\n", step(domain.xmin, domain.xmax, nx),
               step(domain.ymin, domain.ymax, ny), step(domain.zmin, domain.zmax, nz));
  std::fprintf(f, "POINT_DATA %zu\nSCALARS solution double 1\nLOOKUP_TABLE default\n", n);
  for (size_t row = 0; row < n / nx; ++row) {
    for (int i = 0; i < nx; ++i)
      std::fprintf(f, i + 1 < nx ? "%.*g " : "%.*g\n", d, grid[row * nx + i]);
  }
  const bool ok = std::ferror(f) == 0;
  if (std::fclose(f) != 0 || !ok) return fail("failed to write output file");
  return true;
}

bool HipIsAvailable(std::string* note) {
  char buf[256] = {0};
  const int ok = pdehip_is_available(buf, sizeof(buf));
  if (note) *note = buf;
  return ok != 0;
}

namespace {

// SolverConfig (+ the HIP options) -> pdehip_options, with the plug-in's method checks.
// Returns an error message, or an empty string.
std::string build_options(const SolveInput& input, const HipSolverOptions& hopt,
                          const Assembled& A, pdehip_options* out) {
  const SolverConfig& s = input.solver;
  pdehip_options& o = *out;
  pdehip_default_options(&o);
  o.max_iter = std::max(1, s.max_iter);
  o.tol = s.tol;
  o.tol_mode = hopt.relative_tol ? PDEHIP_TOL_REL : PDEHIP_TOL_ABS;
  o.sor_omega = s.sor_omega;
  o.gmres_restart = s.gmres_restart;
  o.mg_pre = s.mg_pre_smooth;
  o.mg_post = s.mg_post_smooth;
  o.mg_coarse_iters = s.mg_coarse_iters;
  o.mg_max_levels = s.mg_max_levels;
  o.precision = hopt.mixed_precision ? PDEHIP_MIXED : PDEHIP_FP64;
  o.precond = hopt.jacobi_precond ? PDEHIP_PC_JACOBI : PDEHIP_PC_NONE;
  const bool dirichlet = all_dirichlet(A.p);
  switch (s.method) {
    case SolveMethod::Jacobi: o.method = PDEHIP_JACOBI; break;
    case SolveMethod::GaussSeidel: o.method = PDEHIP_GAUSS_SEIDEL; break;
    case SolveMethod::SOR:
      o.method = PDEHIP_SOR;
      if (s.sor_omega <= 0.0 || s.sor_omega >= 2.0) return "SOR omega must be in (0,2)";
      break;
    case SolveMethod::CG:
      // Folded rows: identical iterates to the reference rows for zero Dirichlet data, and a
      // symmetric system (so CG converges) for non-zero Dirichlet, Neumann and Robin data.
      o.method = PDEHIP_CG;
      o.bc_rows = PDEHIP_ROWS_FOLDED;
      if (A.p.op_kind != PDEHIP_OP_CONST || std::abs(A.p.c) > 1e-12 || std::abs(A.p.d) > 1e-12 ||
          std::abs(A.p.dz) > 1e-12)
        return "CG requires symmetric operators with constant coefficients (no convection/mixed/higher-order terms)";
      break;
    case SolveMethod::BiCGStab:
      o.method = PDEHIP_BICGSTAB;
      o.bc_rows = dirichlet ? PDEHIP_ROWS_REFERENCE : PDEHIP_ROWS_FOLDED;
      break;
    case SolveMethod::GMRES:
      o.method = PDEHIP_GMRES;
      o.bc_rows = dirichlet ? PDEHIP_ROWS_REFERENCE : PDEHIP_ROWS_FOLDED;
      break;
    case SolveMethod::MultigridVcycle:
      o.method = hopt.mg_pcg ? PDEHIP_MGPCG : PDEHIP_MG;
      o.mg_single_cycle = hopt.mg_single_cycle ? 1 : 0;
      o.bc_rows = PDEHIP_ROWS_FOLDED;
      o.precond = PDEHIP_PC_MG;
      break;
  }
  if (!input.nonlinear.empty() && o.method > PDEHIP_SOR) return "Krylov solvers require linear PDEs";
  return std::string();
}

// Status of a finished solve -> SolveOutput (error string or residual norms).
void finish_output(int st, SolveMethod m, const pdehip_report& rep, pdehip_ctx* ctx, SolveOutput* out) {
  if (st == PDEHIP_ERR_CANCELLED) { *out = {"solve cancelled", {}}; return; }
  if (st == PDEHIP_ERR_NOT_CONVERGED) { *out = {method_failure(m), {}}; return; }
  if (st != PDEHIP_OK) {
    const char* e = pdehip_last_error(ctx);
    *out = {std::string(e && *e ? e : "HIP solve failed"), {}};
    return;
  }
  out->residual_l2 = rep.residual_l2;
  out->residual_linf = rep.residual_linf;
}

bool same_options(const pdehip_options& a, const pdehip_options& b) {
  return a.method == b.method && a.max_iter == b.max_iter && a.tol == b.tol &&
         a.tol_mode == b.tol_mode && a.sor_omega == b.sor_omega &&
         a.gmres_restart == b.gmres_restart && a.mg_pre == b.mg_pre && a.mg_post == b.mg_post &&
         a.mg_coarse_iters == b.mg_coarse_iters && a.mg_max_levels == b.mg_max_levels &&
         a.mg_single_cycle == b.mg_single_cycle && a.precond == b.precond &&
         a.precision == b.precision && a.bc_rows == b.bc_rows;
}

}  // namespace

SolveOutput SolvePDEHip(const SolveInput& input) {
  if (std::abs(input.pde.ut) > 1e-12 || std::abs(input.pde.utt) > 1e-12)
    return {"time-dependent PDEs require the time-series solver", {}};
  std::string err;
  pdehip_ctx* ctx = context(&err);
  if (!ctx) return {"HIP unavailable: " + err, {}};
  const HipSolverOptions hopt = HipGetSolverOptions();
  Assembled A;
  if (!assemble(input, hopt, &A)) return {A.error, {}};
  pdehip_options o;
  err = build_options(input, hopt, A, &o);
  if (!err.empty()) return {err, {}};
  CancelBridge cancel(input.cancel);
  o.cancel = cancel.flag();
  SolveOutput out;
  out.grid.assign(static_cast<size_t>(A.p.nx) * A.p.ny * A.p.nz, 0.0);
  pdehip_report rep{};
  const int st = pdehip_solve(ctx, &A.p, &o, out.grid.data(), &rep);
  finish_output(st, input.solver.method, rep, ctx, &out);
  return out;
}

std::vector<SolveOutput> SolvePDEBatchHip(const std::vector<SolveInput>& inputs) {
  std::vector<SolveOutput> outs(inputs.size());
  std::string err;
  pdehip_ctx* ctx = context(&err);
  if (!ctx) {
    for (auto& o : outs) o = {"HIP unavailable: " + err, {}};
    return outs;
  }
  const HipSolverOptions hopt = HipGetSolverOptions();
  std::vector<Assembled> A(inputs.size());
  std::vector<pdehip_options> opt(inputs.size());
  std::vector<char> ok(inputs.size(), 0);
  for (size_t i = 0; i < inputs.size(); ++i) {
    const SolveInput& in = inputs[i];
    if (std::abs(in.pde.ut) > 1e-12 || std::abs(in.pde.utt) > 1e-12) {
      outs[i] = {"time-dependent PDEs require the time-series solver", {}};
      continue;
    }
    if (!assemble(in, hopt, &A[i])) { outs[i] = {A[i].error, {}}; continue; }
    err = build_options(in, hopt, A[i], &opt[i]);
    if (!err.empty()) { outs[i] = {err, {}}; continue; }
    ok[i] = 1;
  }
  // runs of consecutive assembled inputs with equal options: one pdehip_solve_batch each
  for (size_t i = 0; i < inputs.size();) {
    if (!ok[i]) { ++i; continue; }
    size_t j = i + 1;
    while (j < inputs.size() && ok[j] && same_options(opt[j], opt[i])) ++j;
    const int count = static_cast<int>(j - i);
    std::vector<pdehip_problem> probs(count);
    std::vector<double*> grids(count);
    std::vector<pdehip_report> reps(count);
    for (int t = 0; t < count; ++t) {
      probs[t] = A[i + t].p;
      outs[i + t].grid.assign(static_cast<size_t>(probs[t].nx) * probs[t].ny * probs[t].nz, 0.0);
      grids[t] = outs[i + t].grid.data();
    }
    for (auto& r : reps) r.status = PDEHIP_PENDING;   // left so if the batch stops early
    const int st = pdehip_solve_batch(ctx, count, probs.data(), &opt[i], nullptr, grids.data(),
                                      reps.data());
    for (int t = 0; t < count; ++t) {
      const int sti = reps[t].status != PDEHIP_PENDING ? reps[t].status
                                                       : (st != PDEHIP_OK ? st : PDEHIP_ERR_DEVICE);
      finish_output(sti, inputs[i + t].solver.method, reps[t], ctx, &outs[i + t]);
    }
    i = j;
  }
  return outs;
}

namespace {
struct FrameBridge {
  const FrameCallback* cb;
  size_t n;
  std::vector<double> grid, vel;
};
int frame_thunk(void* user, int frame, double t, const double* grid, const double* velocity) {
  FrameBridge* b = static_cast<FrameBridge*>(user);
  b->grid.assign(grid, grid + b->n);
  const std::vector<double>* vp = nullptr;
  if (velocity) {
    b->vel.assign(velocity, velocity + b->n);
    vp = &b->vel;
  }
  return (*b->cb)(frame, t, b->grid, vp) ? 1 : 0;
}
}  // namespace

SolveOutput SolvePDETimeSeriesHip(const SolveInput& input, const FrameCallback& on_frame,
                                  const ProgressCallback& progress) {
  if (input.domain.nz > 1) return {"3D Time Series Solver not yet refactored.", {}};
  std::string err;
  pdehip_ctx* ctx = context(&err);
  if (!ctx) return {"HIP unavailable: " + err, {}};
  Assembled A;
  if (!assemble(input, HipGetSolverOptions(), &A)) return {A.error, {}};
  pdehip_time_config tc{};
  switch (input.time.integrator) {
    case TimeIntegrator::ForwardEuler: tc.integrator = PDEHIP_TI_FORWARD_EULER; break;
    case TimeIntegrator::RK2: tc.integrator = PDEHIP_TI_RK2; break;
    case TimeIntegrator::RK4: tc.integrator = PDEHIP_TI_RK4; break;
    case TimeIntegrator::SSPRK2: tc.integrator = PDEHIP_TI_SSPRK2; break;
    case TimeIntegrator::SSPRK3: tc.integrator = PDEHIP_TI_SSPRK3; break;
    default: return {"implicit and IMEX integrators are not supported on this backend", {}};
  }
  tc.t_start = input.time.t_start;
  tc.dt = input.time.dt;
  tc.frames = std::max(1, input.time.frames);
  tc.precision = PDEHIP_FP64;
  const size_t n = static_cast<size_t>(A.p.nx) * A.p.ny;
  // frame buffering by TimeConfig.buffer_mb (pde_types.h:341-344)
  const double frame_mb = n * sizeof(double) / (1024.0 * 1024.0);
  tc.buffer_frames = std::max(1, static_cast<int>(input.time.buffer_mb / std::max(frame_mb, 1e-9)));
  if (progress) {
    progress("solve_total", static_cast<double>(tc.frames));
    progress("time", 0.0);
  }
  FrameBridge bridge{&on_frame, n, {}, {}};
  SolveOutput out;
  out.grid.assign(n, 0.0);
  pdehip_report rep{};
  const int st = pdehip_time_series(ctx, &A.p, &tc, on_frame ? frame_thunk : nullptr, &bridge,
                                    out.grid.data(), &rep);
  if (st != PDEHIP_OK) {
    const char* e = pdehip_last_error(ctx);
    return {std::string(e && *e ? e : "HIP time series failed"), {}};
  }
  if (progress) progress("time", 1.0);
  return out;
}
