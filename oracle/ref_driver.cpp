// ref_driver.cpp — glue that drives the REFERENCE CPU path (compiled from the reference's own
// sources into oracle/_ref by oracle/Makefile) through the pdehip_problem description, so
// tests/golden/make_golden.py can record its outputs as golden vectors.  TEST INFRASTRUCTURE
// ONLY: it exists in this container (where the reference sources are) and never in the product.
//
// Everything solver-related below is the reference's own code: this file only converts the
// plain-C problem description into a SolveInput (include/pde_types.h:385-414) and calls
// SolvePDE (src/backends/cpu/solver.cpp:431), SolvePDETimeSeries (solver.cpp:989) and
// ComputeResidualNorms (src/residual.cpp:153).

#include <cstring>
#include <string>
#include <vector>

#include "pdehip.h"
#include "pde_types.h"
#include "residual.h"
#include "solver.h"
#include "backends/cpu/cpu_utils.h"

namespace {
std::string g_err;

BoundaryCondition::Expression ToExpr(const pdehip_expr& e) {
  BoundaryCondition::Expression x;
  x.constant = e.c;
  x.x = e.cx;
  x.y = e.cy;
  x.z = e.cz;
  return x;
}

BoundaryCondition ToBc(const pdehip_face_bc& f) {
  BoundaryCondition bc;
  bc.kind = f.kind == PDEHIP_BC_DIRICHLET ? BCKind::Dirichlet
          : f.kind == PDEHIP_BC_NEUMANN   ? BCKind::Neumann
                                          : BCKind::Robin;
  bc.value = ToExpr(f.value);
  bc.alpha = ToExpr(f.alpha);
  bc.beta = ToExpr(f.beta);
  bc.gamma = ToExpr(f.gamma);
  return bc;
}

SolveMethod ToMethod(int m) {
  switch (m) {
    case PDEHIP_JACOBI: return SolveMethod::Jacobi;
    case PDEHIP_GAUSS_SEIDEL: return SolveMethod::GaussSeidel;
    case PDEHIP_SOR: return SolveMethod::SOR;
    case PDEHIP_CG: return SolveMethod::CG;
    case PDEHIP_BICGSTAB: return SolveMethod::BiCGStab;
    case PDEHIP_GMRES: return SolveMethod::GMRES;
    default: return SolveMethod::MultigridVcycle;
  }
}

// coef_latex: 7 strings (a, b, az, c, d, dz, e) or NULL entries.
SolveInput ToInput(const pdehip_problem& p, const char* rhs_latex, const char* const* coef_latex) {
  SolveInput in;
  in.domain.nx = p.nx;
  in.domain.ny = p.ny;
  in.domain.nz = p.nz < 1 ? 1 : p.nz;
  in.domain.xmin = p.xmin; in.domain.xmax = p.xmax;
  in.domain.ymin = p.ymin; in.domain.ymax = p.ymax;
  in.domain.zmin = p.zmin; in.domain.zmax = p.zmax;
  in.pde.a = p.a; in.pde.b = p.b; in.pde.az = p.az;
  in.pde.c = p.c; in.pde.d = p.d; in.pde.dz = p.dz;
  in.pde.e = p.e; in.pde.f = p.f;
  in.pde.ut = p.ut; in.pde.utt = p.utt;
  if (rhs_latex) in.pde.rhs_latex = rhs_latex;
  if (coef_latex) {
    std::string* dst[7] = {&in.pde.a_latex, &in.pde.b_latex, &in.pde.az_latex, &in.pde.c_latex,
                           &in.pde.d_latex, &in.pde.dz_latex, &in.pde.e_latex};
    for (int m = 0; m < 7; ++m)
      if (coef_latex[m]) *dst[m] = coef_latex[m];
  }
  in.bc.left = ToBc(p.bc[PDEHIP_LEFT]);
  in.bc.right = ToBc(p.bc[PDEHIP_RIGHT]);
  in.bc.bottom = ToBc(p.bc[PDEHIP_BOTTOM]);
  in.bc.top = ToBc(p.bc[PDEHIP_TOP]);
  in.bc.front = ToBc(p.bc[PDEHIP_FRONT]);
  in.bc.back = ToBc(p.bc[PDEHIP_BACK]);
  for (int t = 0; t < p.n_nonlinear; ++t) {
    NonlinearTerm term;
    term.kind = static_cast<NonlinearKind>(p.nonlinear[t].kind);
    term.coeff = p.nonlinear[t].coeff;
    term.power = p.nonlinear[t].power;
    in.nonlinear.push_back(term);
  }
  return in;
}

void ToSolver(const pdehip_options& o, SolverConfig* s) {
  s->method = ToMethod(o.method);
  s->max_iter = o.max_iter;
  s->tol = o.tol;
  s->sor_omega = o.sor_omega;
  s->gmres_restart = o.gmres_restart;
  s->mg_pre_smooth = o.mg_pre;
  s->mg_post_smooth = o.mg_post;
  s->mg_coarse_iters = o.mg_coarse_iters;
  s->mg_max_levels = o.mg_max_levels;
}
}  // namespace

extern "C" {

const char* ref_last_error(void) { return g_err.c_str(); }

// SolvePDE on the reference CPU backend.  Returns 0 and fills u_out on success; 1 with
// ref_last_error() holding SolveOutput.error otherwise.
int ref_solve(const pdehip_problem* p, const pdehip_options* o, const char* rhs_latex,
              const char* const* coef_latex, double* u_out, double* l2, double* linf) {
  SolveInput in = ToInput(*p, rhs_latex, coef_latex);
  ToSolver(*o, &in.solver);
  SolveOutput out = SolvePDE(in);
  g_err = out.error;
  if (!out.error.empty()) return 1;
  std::memcpy(u_out, out.grid.data(), out.grid.size() * sizeof(double));
  ResidualNorms norms;
  std::string err;
  if (l2 && linf && ComputeResidualNorms(in, out.grid, &norms, &err)) {
    *l2 = norms.l2;
    *linf = norms.linf;
  }
  return 0;
}

// LinearOperator2D::Apply (src/backends/cpu/cpu_utils.cpp:324) through BuildLinearOperator2D.
int ref_apply(const pdehip_problem* p, const char* const* coef_latex, const double* x, double* y) {
  SolveInput in = ToInput(*p, nullptr, coef_latex);
  LinearOperator2D op = BuildLinearOperator2D(in, in.domain, nullptr, nullptr);
  if (!op.ok) { g_err = op.error; return 1; }
  const size_t n = static_cast<size_t>(p->nx) * p->ny;
  std::vector<double> xv(x, x + n), yv;
  op.Apply(xv, &yv);
  std::memcpy(y, yv.data(), n * sizeof(double));
  return 0;
}

int ref_residual_norms(const pdehip_problem* p, const char* rhs_latex,
                       const char* const* coef_latex, const double* u, double* l2, double* linf) {
  SolveInput in = ToInput(*p, rhs_latex, coef_latex);
  const size_t n = static_cast<size_t>(p->nx) * p->ny * (p->nz < 1 ? 1 : p->nz);
  std::vector<double> grid(u, u + n);
  ResidualNorms norms;
  std::string err;
  if (!ComputeResidualNorms(in, grid, &norms, &err)) { g_err = err; return 1; }
  *l2 = norms.l2;
  *linf = norms.linf;
  return 0;
}

// SolvePDETimeSeries: frames_out receives frames * N values (frame 0 = initial state).
int ref_time_series(const pdehip_problem* p, const char* rhs_latex, const char* const* coef_latex,
                    const pdehip_time_config* tc, double* frames_out) {
  SolveInput in = ToInput(*p, rhs_latex, coef_latex);
  in.time.enabled = true;
  in.time.t_start = tc->t_start;
  in.time.dt = tc->dt;
  in.time.frames = tc->frames;
  in.time.integrator = static_cast<TimeIntegrator>(tc->integrator);
  const size_t n = static_cast<size_t>(p->nx) * p->ny;
  if (p->initial) in.initial_grid.assign(p->initial, p->initial + n);
  if (p->initial_velocity) in.initial_velocity.assign(p->initial_velocity, p->initial_velocity + n);
  FrameCallback cb = [&](int frame, double, const std::vector<double>& grid,
                         const std::vector<double>*) {
    std::memcpy(frames_out + static_cast<size_t>(frame) * n, grid.data(), n * sizeof(double));
    return true;
  };
  SolveOutput out = SolvePDETimeSeries(in, cb);
  g_err = out.error;
  return out.error.empty() ? 0 : 1;
}

}  // extern "C"
