// pde_types.h — the subset of pde_solver's request/result types that the HIP plug-in reads
// and writes, restated for standalone builds of this repository.  Field names, order of
// meaning and defaults follow the reference header (include/pde_types.h:10-58, 148-159,
// 183-238, 282-345, 385-430) so the plug-in compiles unchanged against the reference's own
// header in-tree.  Members of the reference the HIP path never reads (shapes, multi-field
// coupling, conservation-law settings) are omitted here.  Options of the HIP backend that the
// reference has no member for are set through HipSetSolverOptions (include/hip_solver.h).
#ifndef PDE_TYPES_H
#define PDE_TYPES_H

#include <atomic>
#include <cmath>
#include <string>
#include <utility>
#include <vector>

struct PDECoefficients {
  double a = 0.0, b = 0.0, az = 0.0;     // u_xx, u_yy, u_zz
  double c = 0.0, d = 0.0, dz = 0.0;     // u_x, u_y, u_z
  double e = 0.0;                        // u
  double f = 0.0;                        // constant term
  double ut = 0.0, utt = 0.0;            // time derivatives
  double ab = 0.0, ac = 0.0, bc = 0.0;   // mixed derivatives
  double a3 = 0.0, b3 = 0.0, az3 = 0.0;  // third order
  double a4 = 0.0, b4 = 0.0, az4 = 0.0;  // fourth order
  std::string rhs_latex;
  std::string a_latex, b_latex, az_latex, c_latex, d_latex, dz_latex, e_latex;
  std::string ab_latex, ac_latex, bc_latex;
  std::string a3_latex, b3_latex, az3_latex, a4_latex, b4_latex, az4_latex;
};

inline bool HasVariableCoefficients(const PDECoefficients& p) {
  for (const std::string* s : {&p.a_latex, &p.b_latex, &p.az_latex, &p.c_latex, &p.d_latex,
                               &p.dz_latex, &p.e_latex, &p.ab_latex, &p.ac_latex, &p.bc_latex,
                               &p.a3_latex, &p.b3_latex, &p.az3_latex, &p.a4_latex, &p.b4_latex,
                               &p.az4_latex})
    if (!s->empty()) return true;
  return false;
}

inline bool HasMixedDerivatives(const PDECoefficients& p) {
  return std::abs(p.ab) > 1e-12 || std::abs(p.ac) > 1e-12 || std::abs(p.bc) > 1e-12 ||
         !p.ab_latex.empty() || !p.ac_latex.empty() || !p.bc_latex.empty();
}

inline bool HasHigherOrderDerivatives(const PDECoefficients& p) {
  for (double v : {p.a3, p.b3, p.az3, p.a4, p.b4, p.az4})
    if (std::abs(v) > 1e-12) return true;
  return !p.a3_latex.empty() || !p.b3_latex.empty() || !p.az3_latex.empty() ||
         !p.a4_latex.empty() || !p.b4_latex.empty() || !p.az4_latex.empty();
}

struct IntegralTerm {
  double coeff = 0.0;
  std::string kernel_latex;
};

enum class NonlinearKind { Power, Sin, Cos, Exp, Abs };

struct NonlinearTerm {
  NonlinearKind kind = NonlinearKind::Power;
  double coeff = 0.0;
  int power = 2;
};

enum class NonlinearDerivativeKind { UUx, UUy, UUz, UxUx, UyUy, UzUz, GradSquared };
struct NonlinearDerivativeTerm {
  NonlinearDerivativeKind kind = NonlinearDerivativeKind::UUx;
  double coeff = 0.0;
};

enum class CoordinateSystem {
  Cartesian, Polar, Axisymmetric, Cylindrical, SphericalSurface, SphericalVolume,
  ToroidalSurface, ToroidalVolume,
};

struct Domain {
  double xmin = 0.0, xmax = 1.0, ymin = 0.0, ymax = 1.0, zmin = 0.0, zmax = 1.0;
  int nx = 64, ny = 64, nz = 1;
  CoordinateSystem coord_system = CoordinateSystem::Cartesian;
};

enum class BCKind { Dirichlet, Neumann, Robin, Inflow, Outflow, Transmissive };

struct BoundaryCondition {
  BCKind kind = BCKind::Dirichlet;
  struct Expression {
    double constant = 0.0;
    double x = 0.0, y = 0.0, z = 0.0;
    std::string latex;
  };
  Expression value, alpha, beta, gamma;
  std::string condition_latex;
};

struct PiecewiseBoundaryCondition {
  std::vector<std::pair<std::string, BoundaryCondition>> segments;
  BoundaryCondition default_bc;
};

struct BoundarySet {
  BoundaryCondition left, right, bottom, top, front, back;
  std::vector<PiecewiseBoundaryCondition> left_piecewise, right_piecewise, bottom_piecewise,
      top_piecewise, front_piecewise, back_piecewise;
};

enum class SolveMethod { Jacobi, GaussSeidel, SOR, CG, BiCGStab, GMRES, MultigridVcycle };

enum class AdvectionScheme { Centered, Upwind, LaxWendroff, BeamWarming, Fromm, MinMod, Superbee,
                             VanLeer, MC };

struct SolverConfig {
  SolveMethod method = SolveMethod::Jacobi;
  int max_iter = 10000;
  double tol = 1e-6;
  int thread_count = 0;
  int metal_reduce_interval = 10;
  int metal_tg_x = 0;
  int metal_tg_y = 0;
  int residual_interval = 0;
  double sor_omega = 1.5;
  int gmres_restart = 30;
  int mg_pre_smooth = 2;
  int mg_post_smooth = 2;
  int mg_coarse_iters = 50;
  int mg_max_levels = 10;
  AdvectionScheme advection_scheme = AdvectionScheme::Centered;
  bool advection_cfl_check = false;
};

enum class TimeIntegrator { ForwardEuler, RK2, RK4, SSPRK2, SSPRK3, BackwardEuler, CrankNicolson, IMEX };

struct TimeConfig {
  bool enabled = false;
  double t_start = 0.0;
  double t_end = 1.0;
  double dt = 0.01;
  int frames = 1;
  TimeIntegrator integrator = TimeIntegrator::ForwardEuler;
  bool use_cfl = false;
  double cfl_target = 0.5;
  bool adaptive_dt = false;
  double dt_min = 1e-12;
  double dt_max = 1.0;
  double error_tol = 1e-5;
  int buffer_mb = 256;
};

struct ShapeMask {
  Domain domain;
  std::vector<double> values;
};

struct SolveInput {
  PDECoefficients pde;
  Domain domain;
  BoundarySet bc;
  SolverConfig solver;
  std::string domain_shape;
  ShapeMask shape_mask;
  TimeConfig time;
  std::vector<IntegralTerm> integrals;
  std::vector<NonlinearTerm> nonlinear;
  std::vector<NonlinearDerivativeTerm> nonlinear_derivatives;
  std::vector<double> initial_grid;
  std::vector<double> initial_velocity;
  std::atomic<bool>* cancel = nullptr;
};

struct SolveOutput {
  std::string error;
  std::vector<double> grid;
  double residual_l2 = 0.0;
  double residual_linf = 0.0;
  std::vector<int> residual_iters;
  std::vector<double> residual_l2_history;
  std::vector<double> residual_linf_history;
};

inline bool HasImplicitShape(const SolveInput& input) {
  return !input.domain_shape.empty() || !input.shape_mask.values.empty();
}

#endif  // PDE_TYPES_H
