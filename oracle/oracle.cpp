// oracle.cpp — CPU oracle for the HIP solver backend.  TEST INFRASTRUCTURE ONLY.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this
// library, and only as the checker or the timed CPU baseline; the product path never
// calls it.
//
// It restates, in code written for this repository, the algorithms of the reference CPU
// backend (nathanielasun/pde_solver), each function citing the reference file:line it
// follows, over the same plain-C problem description as the product's C ABI (pdehip.h).
// Parts with no reference counterpart (FLUX operator, folded Robin/Neumann rows, upwind
// convection, MG-preconditioned CG, iterated V-cycles, Jacobi-preconditioned Krylov,
// 3D Krylov) are marked "extension"; they are pinned by dense direct solves, closed forms
// and manufactured solutions in tests/ rather than by the reference.
//
// Each function is pinned by golden vectors from the reference build (oracle/_ref), by closed
// forms and dense direct solves in tests/, or both.

#include "pdehip.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_error;

// ----------------------------------------------------------------------------- geometry
struct Geo {
  int nx = 0, ny = 0, nz = 1;
  bool is3d = false;
  double xmin = 0, xmax = 1, ymin = 0, ymax = 1, zmin = 0, zmax = 1;
  double dx = 0, dy = 0, dz = 1;
  size_t n = 0;
  size_t idx(int i, int j, int k) const { return (static_cast<size_t>(k) * ny + j) * nx + i; }
  double X(int i) const { return xmin + i * dx; }
  double Y(int j) const { return ymin + j * dy; }
  double Z(int k) const { return zmin + k * dz; }
};

// Grid spacing (reference solver.cpp:462-467, 633-639).
Geo MakeGeo(const pdehip_problem& p) {
  Geo g;
  g.nx = p.nx;
  g.ny = p.ny;
  g.nz = std::max(1, p.nz);
  g.is3d = g.nz > 1;
  g.xmin = p.xmin; g.xmax = p.xmax; g.ymin = p.ymin; g.ymax = p.ymax;
  g.zmin = p.zmin; g.zmax = p.zmax;
  g.dx = (p.xmax - p.xmin) / static_cast<double>(g.nx - 1);
  g.dy = (p.ymax - p.ymin) / static_cast<double>(g.ny - 1);
  g.dz = g.is3d ? (p.zmax - p.zmin) / static_cast<double>(g.nz - 1) : 1.0;
  g.n = static_cast<size_t>(g.nx) * g.ny * g.nz;
  return g;
}

Geo Coarsen(const Geo& f) {
  Geo c = f;
  c.nx = (f.nx + 1) / 2;
  c.ny = (f.ny + 1) / 2;
  c.nz = f.is3d ? (f.nz + 1) / 2 : 1;
  c.dx = (c.xmax - c.xmin) / static_cast<double>(c.nx - 1);
  c.dy = (c.ymax - c.ymin) / static_cast<double>(c.ny - 1);
  c.dz = c.is3d ? (c.zmax - c.zmin) / static_cast<double>(c.nz - 1) : 1.0;
  c.n = static_cast<size_t>(c.nx) * c.ny * c.nz;
  return c;
}

// Index of a boundary node inside its face's sample array (pdehip_expr.samples order).
size_t FaceIndex(const Geo& g, int face, int i, int j, int k) {
  switch (face) {
    case PDEHIP_LEFT: case PDEHIP_RIGHT: return static_cast<size_t>(k) * g.ny + j;
    case PDEHIP_BOTTOM: case PDEHIP_TOP: return static_cast<size_t>(k) * g.nx + i;
    default: return static_cast<size_t>(j) * g.nx + i;
  }
}

// BoundaryCondition::Expression evaluation, linear form (reference cpu_utils.cpp:521-530).
double EvalExpr(const pdehip_expr& e, size_t fidx, double x, double y, double z) {
  if (e.samples) return e.samples[fidx];
  return e.c + e.cx * x + e.cy * y + e.cz * z;
}

// EvalNonlinear (reference relaxation.cpp:27-49).
double EvalNonlinear(const pdehip_problem& p, double u) {
  double v = 0.0;
  for (int t = 0; t < p.n_nonlinear; ++t) {
    const pdehip_nonlinear& term = p.nonlinear[t];
    switch (term.kind) {
      case PDEHIP_NL_POWER: v += term.coeff * std::pow(u, static_cast<double>(term.power)); break;
      case PDEHIP_NL_SIN: v += term.coeff * std::sin(u); break;
      case PDEHIP_NL_COS: v += term.coeff * std::cos(u); break;
      case PDEHIP_NL_EXP: v += term.coeff * std::exp(u); break;
      case PDEHIP_NL_ABS: v += term.coeff * std::abs(u); break;
    }
  }
  return v;
}

double SourceAt(const pdehip_problem& p, size_t idx) { return p.f_field ? p.f_field[idx] : p.f; }

// ----------------------------------------------------------------------------- stencil weights
// Seven-point weights of one row: y_i = C u_i + W u_{i-1} + E u_{i+1} + S u_{j-1} + N u_{j+1}
//                                        + B u_{k-1} + F u_{k+1}.
struct W7 {
  double C = 0, W = 0, E = 0, S = 0, N = 0, B = 0, F = 0;
};

double CoefAt(const pdehip_problem& p, int m, size_t idx, double cst) {
  return p.coef[m] ? p.coef[m][idx] : cst;
}

// Reference coefficient formulas (cpu_utils.cpp:293-297, 404-408; solver.cpp:723-729,854-862).
// Upwind first-order convection is an extension.
W7 WeightsAt(const pdehip_problem& p, const Geo& g, int i, int j, int k) {
  const size_t id = g.idx(i, j, k);
  W7 w;
  if (p.op_kind == PDEHIP_OP_FLUX) {
    // extension: a*div(kappa grad u) + e*u, face kappa = harmonic mean of the two nodes.
    const double kc = p.kappa[id];
    auto face = [&](size_t nb, double h) {
      const double kn = p.kappa[nb];
      const double kf = (kc + kn) > 0.0 ? 2.0 * kc * kn / (kc + kn) : 0.0;
      return p.a * kf / (h * h);
    };
    const double e_val = CoefAt(p, PDEHIP_COEF_E, id, p.e);
    w.W = face(g.idx(i - 1, j, k), g.dx);
    w.E = face(g.idx(i + 1, j, k), g.dx);
    w.S = face(g.idx(i, j - 1, k), g.dy);
    w.N = face(g.idx(i, j + 1, k), g.dy);
    double sum = w.W + w.E + w.S + w.N;
    if (g.is3d) {
      w.B = face(g.idx(i, j, k - 1), g.dz);
      w.F = face(g.idx(i, j, k + 1), g.dz);
      sum += w.B + w.F;
    }
    w.C = e_val - sum;
    return w;
  }
  const double a = CoefAt(p, PDEHIP_COEF_A, id, p.a);
  const double b = CoefAt(p, PDEHIP_COEF_B, id, p.b);
  const double az = CoefAt(p, PDEHIP_COEF_AZ, id, p.az);
  const double c = CoefAt(p, PDEHIP_COEF_C, id, p.c);
  const double d = CoefAt(p, PDEHIP_COEF_D, id, p.d);
  const double dzc_coef = CoefAt(p, PDEHIP_COEF_DZ, id, p.dz);
  const double e = CoefAt(p, PDEHIP_COEF_E, id, p.e);
  const double ax = a / (g.dx * g.dx);
  const double by = b / (g.dy * g.dy);
  const double cz = g.is3d ? az / (g.dz * g.dz) : 0.0;
  w.C = g.is3d ? (-2.0 * ax - 2.0 * by - 2.0 * cz + e) : (-2.0 * ax - 2.0 * by + e);
  if (p.convection == PDEHIP_CONV_UPWIND) {
    w.E = ax + std::min(c, 0.0) / g.dx;
    w.W = ax - std::max(c, 0.0) / g.dx;
    w.N = by + std::min(d, 0.0) / g.dy;
    w.S = by - std::max(d, 0.0) / g.dy;
    w.C += std::abs(c) / g.dx + std::abs(d) / g.dy;
    if (g.is3d) {
      w.F = cz + std::min(dzc_coef, 0.0) / g.dz;
      w.B = cz - std::max(dzc_coef, 0.0) / g.dz;
      w.C += std::abs(dzc_coef) / g.dz;
    }
  } else {
    const double cx = c / (2.0 * g.dx);
    const double dyc = d / (2.0 * g.dy);
    w.E = ax + cx;
    w.W = ax - cx;
    w.N = by + dyc;
    w.S = by - dyc;
    if (g.is3d) {
      const double dzc = dzc_coef / (2.0 * g.dz);
      w.F = cz + dzc;
      w.B = cz - dzc;
    }
  }
  return w;
}

// Neighbour sum in the reference evaluation order (cpu_utils.cpp:420-424, solver.cpp:892-898).
inline double NeighbourSum(const W7& w, const Geo& g, const double* u, int i, int j, int k) {
  double s = w.E * u[g.idx(i + 1, j, k)] + w.W * u[g.idx(i - 1, j, k)] +
             w.N * u[g.idx(i, j + 1, k)] + w.S * u[g.idx(i, j - 1, k)];
  if (g.is3d) s = s + w.F * u[g.idx(i, j, k + 1)] + w.B * u[g.idx(i, j, k - 1)];
  return s;
}

bool IsBoundary(const Geo& g, int i, int j, int k) {
  return i == 0 || j == 0 || i == g.nx - 1 || j == g.ny - 1 ||
         (g.is3d && (k == 0 || k == g.nz - 1));
}

// ----------------------------------------------------------------------------- BC closures
// ApplyDirichletCPU / ApplyDirichletCPU3D (reference cpu_utils.cpp:575-628, 722-772).
void ApplyDirichlet(const pdehip_problem& p, const Geo& g, std::vector<double>& u) {
  auto face = [&](int f, int i, int j, int k, double x, double y, double z) {
    const pdehip_face_bc& bc = p.bc[f];
    if (bc.kind == PDEHIP_BC_DIRICHLET)
      u[g.idx(i, j, k)] = EvalExpr(bc.value, FaceIndex(g, f, i, j, k), x, y, z);
  };
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j) {
      const double z = g.is3d ? g.Z(k) : 0.0;
      face(PDEHIP_LEFT, 0, j, k, g.xmin, g.Y(j), z);
      face(PDEHIP_RIGHT, g.nx - 1, j, k, g.xmax, g.Y(j), z);
    }
  for (int k = 0; k < g.nz; ++k)
    for (int i = 0; i < g.nx; ++i) {
      const double z = g.is3d ? g.Z(k) : 0.0;
      face(PDEHIP_BOTTOM, i, 0, k, g.X(i), g.ymin, z);
      face(PDEHIP_TOP, i, g.ny - 1, k, g.X(i), g.ymax, z);
    }
  if (g.is3d)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        face(PDEHIP_FRONT, i, j, 0, g.X(i), g.Y(j), g.zmin);
        face(PDEHIP_BACK, i, j, g.nz - 1, g.X(i), g.Y(j), g.zmax);
      }
}

// Closure of one boundary node from its inward neighbour.  Returns false when the Robin
// denominator vanishes (the reference then leaves the node untouched, cpu_utils.cpp:651).
// AXIS: cpu_utils.cpp:643-653 (low) and 664-674 (high).  OUTWARD: extension, alpha*u+beta*du/dn.
bool Closure(const pdehip_problem& p, int f, double h, double inner, size_t fidx, double x,
             double y, double z, double* out) {
  const pdehip_face_bc& bc = p.bc[f];
  const bool high = (f == PDEHIP_RIGHT || f == PDEHIP_TOP || f == PDEHIP_BACK);
  const bool axis = p.bc_convention == PDEHIP_BCCONV_AXIS;
  if (bc.kind == PDEHIP_BC_NEUMANN) {
    const double gval = EvalExpr(bc.value, fidx, x, y, z);
    if (axis) *out = high ? inner + h * gval : inner - h * gval;
    else *out = inner + h * gval;
    return true;
  }
  if (bc.kind == PDEHIP_BC_ROBIN) {
    const double alpha = EvalExpr(bc.alpha, fidx, x, y, z);
    const double beta = EvalExpr(bc.beta, fidx, x, y, z);
    const double gamma = EvalExpr(bc.gamma, fidx, x, y, z);
    if (axis && high) {
      const double denom = alpha - beta / h;
      if (std::abs(denom) > 1e-12) { *out = (gamma - (beta / h) * inner) / denom; return true; }
      return false;
    }
    const double denom = alpha + beta / h;
    if (std::abs(denom) > 1e-12) { *out = (gamma + (beta / h) * inner) / denom; return true; }
    return false;
  }
  return false;
}

// ApplyNeumannRobinCPU / ApplyNeumannRobinCPU3D (reference cpu_utils.cpp:630-720, 774-901).
void ApplyNeumannRobin(const pdehip_problem& p, const Geo& g, std::vector<double>& u) {
  auto node = [&](int f, int i, int j, int k, int ii, int jj, int kk, double h, double x, double y,
                  double z) {
    if (p.bc[f].kind == PDEHIP_BC_DIRICHLET) return;
    double v;
    if (Closure(p, f, h, u[g.idx(ii, jj, kk)], FaceIndex(g, f, i, j, k), x, y, z, &v))
      u[g.idx(i, j, k)] = v;
  };
  if (!g.is3d) {
    for (int j = 0; j < g.ny; ++j) node(PDEHIP_LEFT, 0, j, 0, 1, j, 0, g.dx, g.xmin, g.Y(j), 0.0);
    for (int j = 0; j < g.ny; ++j)
      node(PDEHIP_RIGHT, g.nx - 1, j, 0, g.nx - 2, j, 0, g.dx, g.xmax, g.Y(j), 0.0);
    for (int i = 0; i < g.nx; ++i) node(PDEHIP_BOTTOM, i, 0, 0, i, 1, 0, g.dy, g.X(i), g.ymin, 0.0);
    for (int i = 0; i < g.nx; ++i)
      node(PDEHIP_TOP, i, g.ny - 1, 0, i, g.ny - 2, 0, g.dy, g.X(i), g.ymax, 0.0);
    return;
  }
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j) {
      node(PDEHIP_LEFT, 0, j, k, 1, j, k, g.dx, g.xmin, g.Y(j), g.Z(k));
      node(PDEHIP_RIGHT, g.nx - 1, j, k, g.nx - 2, j, k, g.dx, g.xmax, g.Y(j), g.Z(k));
    }
  for (int k = 0; k < g.nz; ++k)
    for (int i = 0; i < g.nx; ++i) {
      node(PDEHIP_BOTTOM, i, 0, k, i, 1, k, g.dy, g.X(i), g.ymin, g.Z(k));
      node(PDEHIP_TOP, i, g.ny - 1, k, i, g.ny - 2, k, g.dy, g.X(i), g.ymax, g.Z(k));
    }
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i) {
      node(PDEHIP_FRONT, i, j, 0, i, j, 1, g.dz, g.X(i), g.Y(j), g.zmin);
      node(PDEHIP_BACK, i, j, g.nz - 1, i, j, g.nz - 2, g.dz, g.X(i), g.Y(j), g.zmax);
    }
}

bool AllDirichlet(const pdehip_problem& p, const Geo& g) {
  const int nf = g.is3d ? 6 : 4;
  for (int f = 0; f < nf; ++f)
    if (p.bc[f].kind != PDEHIP_BC_DIRICHLET) return false;
  return true;
}

// ----------------------------------------------------------------------------- operator
// Folded-row data (extension): per face the closure factor theta with u_b = theta*u_in + c0.
struct Fold {
  double theta[6] = {0, 0, 0, 0, 0, 0};
};

// theta of a Neumann/Robin face at spacing h (AXIS / OUTWARD as in Closure()).
double ThetaOf(const pdehip_problem& p, int f, double h) {
  const pdehip_face_bc& bc = p.bc[f];
  if (bc.kind == PDEHIP_BC_NEUMANN) return 1.0;
  if (bc.kind != PDEHIP_BC_ROBIN) return 0.0;
  const double alpha = bc.alpha.c, beta = bc.beta.c;
  const bool high = (f == PDEHIP_RIGHT || f == PDEHIP_TOP || f == PDEHIP_BACK);
  if (p.bc_convention == PDEHIP_BCCONV_AXIS && high) return -(beta / h) / (alpha - beta / h);
  return (beta / h) / (alpha + beta / h);
}

Fold MakeFold(const pdehip_problem& p, const Geo& g, bool folded) {
  Fold fo;
  if (!folded) return fo;
  const double hs[6] = {g.dx, g.dx, g.dy, g.dy, g.dz, g.dz};
  for (int f = 0; f < (g.is3d ? 6 : 4); ++f) fo.theta[f] = ThetaOf(p, f, hs[f]);
  return fo;
}

// Effective weights of an interior row after folding: a neighbour on a folded face adds
// w*theta to the centre (its stored value carries the affine part c0).
W7 FoldedWeights(const W7& w0, const Fold& fo, const Geo& g, int i, int j, int k) {
  W7 w = w0;
  if (i == 1) w.C += w.W * fo.theta[PDEHIP_LEFT];
  if (i == g.nx - 2) w.C += w.E * fo.theta[PDEHIP_RIGHT];
  if (j == 1) w.C += w.S * fo.theta[PDEHIP_BOTTOM];
  if (j == g.ny - 2) w.C += w.N * fo.theta[PDEHIP_TOP];
  if (g.is3d) {
    if (k == 1) w.C += w.B * fo.theta[PDEHIP_FRONT];
    if (k == g.nz - 2) w.C += w.F * fo.theta[PDEHIP_BACK];
  }
  return w;
}

// A level operator: problem coefficients sampled at a (possibly coarsened) geometry.
struct LevelOp {
  Geo g;
  const pdehip_problem* p = nullptr;
  pdehip_problem pl;             // copy with coarsened kappa / coefficient pointers
  std::vector<double> kappa;     // coarsened kappa (FLUX)
  Fold fold;
  std::vector<W7> w;             // per-node folded weights (interior nodes)
  void Build(const pdehip_problem& prob, const Geo& geo, bool folded, const double* kap) {
    p = &prob;
    pl = prob;
    g = geo;
    if (prob.op_kind == PDEHIP_OP_FLUX) {
      kappa.assign(kap, kap + g.n);
      pl.kappa = kappa.data();
    }
    fold = MakeFold(prob, g, folded);
    w.assign(g.n, W7());
#pragma omp parallel for schedule(static)
    for (int k = (g.is3d ? 1 : 0); k < (g.is3d ? g.nz - 1 : 1); ++k)
      for (int j = 1; j < g.ny - 1; ++j)
        for (int i = 1; i < g.nx - 1; ++i)
          w[g.idx(i, j, k)] = FoldedWeights(WeightsAt(pl, g, i, j, k), fold, g, i, j, k);
  }
  // y = A x; interior rows from the weights, boundary rows identity
  // (LinearOperator2D::Apply boundary rule, cpu_utils.cpp:368-381, 497-512).
  void Apply(const std::vector<double>& x, std::vector<double>& y) const {
    y.assign(g.n, 0.0);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < g.nz; ++k)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {
          const size_t id = g.idx(i, j, k);
          if (IsBoundary(g, i, j, k)) { y[id] = x[id]; continue; }
          const W7& ww = w[id];
          y[id] = ww.C * x[id] + NeighbourSum(ww, g, x.data(), i, j, k);
        }
  }
};

double Dot(const std::vector<double>& a, const std::vector<double>& b) {
  // Dot (reference cpu_utils.cpp:30-40).
  double s = 0.0;
  const size_t n = std::min(a.size(), b.size());
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}
double Norm2(const std::vector<double>& v) { return std::sqrt(std::max(0.0, Dot(v, v))); }

// ----------------------------------------------------------------------------- RHS
// BuildRhs2D (reference solver.cpp:167-232), 3D by the same rule: b = -f on interior
// rows, the Dirichlet value on Dirichlet boundary nodes, 0 elsewhere on the boundary.
std::vector<double> BuildRhs(const pdehip_problem& p, const Geo& g) {
  std::vector<double> boundary(g.n, 0.0);
  ApplyDirichlet(p, g, boundary);
  std::vector<double> b(g.n, 0.0);
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const size_t id = g.idx(i, j, k);
        b[id] = IsBoundary(g, i, j, k) ? boundary[id] : -SourceAt(p, id);
      }
  return b;
}

// Folded initial vector (extension): Dirichlet values on Dirichlet nodes, the affine part
// c0 of the closure (closure of a zero inner value) on Neumann/Robin face nodes.
std::vector<double> FoldedBoundaryData(const pdehip_problem& p, const Geo& g) {
  std::vector<double> u(g.n, 0.0);
  // c0 via the closures with a zero interior, then Dirichlet overwrite (reference order).
  ApplyNeumannRobin(p, g, u);
  ApplyDirichlet(p, g, u);
  return u;
}

// Final boundary reconstruction of a folded solve: closures then Dirichlet, reference
// order (solver.cpp:1321-1322 / relaxation.cpp:551-552).
void ReconstructBoundary(const pdehip_problem& p, const Geo& g, std::vector<double>& u) {
  ApplyNeumannRobin(p, g, u);
  ApplyDirichlet(p, g, u);
}

// ----------------------------------------------------------------------------- Krylov
struct KrylovResult {
  bool ok = false;
  int status = PDEHIP_ERR_NOT_CONVERGED;
  int iters = 0;
  double metric = 0.0;
};

// Jacobi preconditioner (extension): z = r / diag on interior rows, identity on the boundary.
struct Precond {
  const LevelOp* op = nullptr;
  bool jacobi = false;
  void Apply(const std::vector<double>& r, std::vector<double>& z) const {
    z = r;
    if (!jacobi) return;
    const Geo& g = op->g;
    for (int k = 0; k < g.nz; ++k)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {
          if (IsBoundary(g, i, j, k)) continue;
          const size_t id = g.idx(i, j, k);
          z[id] = r[id] / op->w[id].C;
        }
  }
};

// CgSolve (reference krylov.cpp:14-80): x0 = 0, stop ||r||_2 < tol (absolute).
KrylovResult Cg(const LevelOp& op, const std::vector<double>& b, std::vector<double>& x,
                int max_iter, double tol) {
  KrylovResult res;
  const size_t n = b.size();
  std::vector<double> r(n), p(n), Ap;
  op.Apply(x, Ap);
  for (size_t i = 0; i < n; ++i) { r[i] = b[i] - Ap[i]; p[i] = r[i]; }
  double rr = Dot(r, r);
  res.metric = std::sqrt(rr);
  if (std::sqrt(rr) < tol) { res.ok = true; res.status = PDEHIP_OK; return res; }
  for (int iter = 0; iter < max_iter; ++iter) {
    op.Apply(p, Ap);
    const double pAp = Dot(p, Ap);
    if (std::abs(pAp) < 1e-30) { res.status = PDEHIP_ERR_BREAKDOWN; res.iters = iter; return res; }
    const double alpha = rr / pAp;
    for (size_t i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * Ap[i]; }
    const double rr_new = Dot(r, r);
    res.iters = iter + 1;
    res.metric = std::sqrt(rr_new);
    if (std::sqrt(rr_new) < tol) { res.ok = true; res.status = PDEHIP_OK; return res; }
    const double beta = rr_new / rr;
    for (size_t i = 0; i < n; ++i) p[i] = r[i] + beta * p[i];
    rr = rr_new;
  }
  return res;
}

// BiCGStabSolve (reference krylov.cpp:82-185), with an optional right preconditioner
// (extension; identity reproduces the reference exactly).
KrylovResult BiCGStab(const LevelOp& op, const Precond& M, const std::vector<double>& b,
                      std::vector<double>& x, int max_iter, double tol) {
  KrylovResult res;
  const size_t n = b.size();
  std::vector<double> r(n), rhat(n), p(n, 0.0), v(n, 0.0), s(n), t(n), Ax, ph, sh;
  op.Apply(x, Ax);
  for (size_t i = 0; i < n; ++i) { r[i] = b[i] - Ax[i]; rhat[i] = r[i]; }
  double rho_prev = 1.0, alpha = 1.0, omega = 1.0;
  const double resid0 = Norm2(r);
  res.metric = resid0;
  if (resid0 < tol) { res.ok = true; res.status = PDEHIP_OK; return res; }
  for (int iter = 0; iter < max_iter; ++iter) {
    const double rho = Dot(rhat, r);
    if (std::abs(rho) < 1e-30) { res.status = PDEHIP_ERR_BREAKDOWN; return res; }
    if (iter == 0) {
      p = r;
    } else {
      const double beta = (rho / rho_prev) * (alpha / omega);
      for (size_t i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * v[i]);
    }
    M.Apply(p, ph);
    op.Apply(ph, v);
    const double denom = Dot(rhat, v);
    if (std::abs(denom) < 1e-30) { res.status = PDEHIP_ERR_BREAKDOWN; return res; }
    alpha = rho / denom;
    for (size_t i = 0; i < n; ++i) s[i] = r[i] - alpha * v[i];
    const double s_norm = Norm2(s);
    if (s_norm < tol) {
      for (size_t i = 0; i < n; ++i) x[i] += alpha * ph[i];
      res.iters = iter + 1; res.metric = s_norm; res.ok = true; res.status = PDEHIP_OK;
      return res;
    }
    M.Apply(s, sh);
    op.Apply(sh, t);
    const double tt = Dot(t, t);
    if (std::abs(tt) < 1e-30) { res.status = PDEHIP_ERR_BREAKDOWN; return res; }
    omega = Dot(t, s) / tt;
    for (size_t i = 0; i < n; ++i) x[i] += alpha * ph[i] + omega * sh[i];
    op.Apply(x, Ax);
    for (size_t i = 0; i < n; ++i) r[i] = b[i] - Ax[i];
    const double resid = Norm2(r);
    res.iters = iter + 1;
    res.metric = resid;
    if (resid < tol) { res.ok = true; res.status = PDEHIP_OK; return res; }
    if (std::abs(omega) < 1e-30) { res.status = PDEHIP_ERR_BREAKDOWN; return res; }
    rho_prev = rho;
  }
  return res;
}

// GmresSolve (reference krylov.cpp:187-333): restarted GMRES(m), modified Gram-Schmidt,
// Givens rotations with the CPU sign convention (krylov.cpp:240-245, 277-279).  The
// back-substitution dimension is the number of Arnoldi steps of the current cycle; the
// reference uses min(m, cumulative steps) (krylov.cpp:297), which equals it in the first
// cycle and, in later cycles that stop early, divides by the zero diagonal of an unfilled
// column and reports failure.  Optional right preconditioner: extension.
KrylovResult Gmres(const LevelOp& op, const Precond& M, const std::vector<double>& b,
                   std::vector<double>& x, int max_iter, int restart, double tol) {
  KrylovResult res;
  const size_t n = b.size();
  restart = std::max(1, restart);
  max_iter = std::max(1, max_iter);
  std::vector<double> Ax, r(n), w(n), zt;
  op.Apply(x, Ax);
  for (size_t i = 0; i < n; ++i) r[i] = b[i] - Ax[i];
  double beta = Norm2(r);
  res.metric = beta;
  if (beta < tol) { res.ok = true; res.status = PDEHIP_OK; return res; }
  int iter_total = 0;
  while (iter_total < max_iter) {
    const int m = std::min(restart, max_iter - iter_total);
    std::vector<std::vector<double>> V(m + 1, std::vector<double>(n, 0.0));
    std::vector<std::vector<double>> H(m + 1, std::vector<double>(m, 0.0));
    std::vector<double> cs(m, 0.0), sn(m, 0.0), gv(m + 1, 0.0);
    gv[0] = beta;
    for (size_t i = 0; i < n; ++i) V[0][i] = r[i] / beta;
    auto givens = [&](double v1, double v2, int k, double* o1, double* o2) {
      *o1 = cs[k] * v1 - sn[k] * v2;
      *o2 = sn[k] * v1 + cs[k] * v2;
    };
    int steps = 0;
    for (int j = 0; j < m; ++j) {
      M.Apply(V[j], zt);
      op.Apply(zt, w);
      for (int i = 0; i <= j; ++i) {
        H[i][j] = Dot(w, V[i]);
        for (size_t q = 0; q < n; ++q) w[q] -= H[i][j] * V[i][q];
      }
      H[j + 1][j] = Norm2(w);
      if (H[j + 1][j] > 0.0)
        for (size_t q = 0; q < n; ++q) V[j + 1][q] = w[q] / H[j + 1][j];
      for (int i = 0; i < j; ++i) {
        double t1, t2;
        givens(H[i][j], H[i + 1][j], i, &t1, &t2);
        H[i][j] = t1;
        H[i + 1][j] = t2;
      }
      const double h1 = H[j + 1][j], h0 = H[j][j];
      const double denom = std::sqrt(h0 * h0 + h1 * h1);
      if (denom < 1e-30) { res.status = PDEHIP_ERR_BREAKDOWN; return res; }
      cs[j] = h0 / denom;
      sn[j] = -h1 / denom;
      H[j][j] = cs[j] * h0 - sn[j] * h1;
      H[j + 1][j] = 0.0;
      double t1, t2;
      givens(gv[j], gv[j + 1], j, &t1, &t2);
      gv[j] = t1;
      gv[j + 1] = t2;
      ++iter_total;
      ++steps;
      res.metric = std::abs(gv[j + 1]);
      if (std::abs(gv[j + 1]) < tol) break;
    }
    const int y_dim = steps;
    std::vector<double> y(y_dim, 0.0);
    for (int i = y_dim - 1; i >= 0; --i) {
      double sum = gv[i];
      for (int q = i + 1; q < y_dim; ++q) sum -= H[i][q] * y[q];
      if (std::abs(H[i][i]) < 1e-30) { res.status = PDEHIP_ERR_BREAKDOWN; return res; }
      y[i] = sum / H[i][i];
    }
    // x += M (V y)
    std::vector<double> vy(n, 0.0);
    for (int j = 0; j < y_dim; ++j)
      for (size_t q = 0; q < n; ++q) vy[q] += y[j] * V[j][q];
    M.Apply(vy, zt);
    for (size_t q = 0; q < n; ++q) x[q] += zt[q];
    op.Apply(x, Ax);
    for (size_t i = 0; i < n; ++i) r[i] = b[i] - Ax[i];
    beta = Norm2(r);
    res.iters = iter_total;
    res.metric = beta;
    if (beta < tol) { res.ok = true; res.status = PDEHIP_OK; return res; }
    if (iter_total >= max_iter) break;
  }
  return res;
}

// ----------------------------------------------------------------------------- multigrid
// The reference V-cycle (multigrid.cpp:10-222): 2D, constant a/b/e, Dirichlet, lexicographic
// Gauss-Seidel, full weighting, bilinear prolongation, one cycle.
bool CanCoarsenOdd(int n) { return n >= 5 && (n % 2 == 1); }  // multigrid.cpp:10-12

void RefRestrictFW(const std::vector<double>& fine, int nxf, std::vector<double>& coarse, int nxc,
                   int nyc) {  // multigrid.cpp:14-36
  coarse.assign(static_cast<size_t>(nxc) * nyc, 0.0);
  for (int jc = 1; jc < nyc - 1; ++jc) {
    const int jf = 2 * jc;
    for (int ic = 1; ic < nxc - 1; ++ic) {
      const int ifx = 2 * ic;
      auto F = [&](int i, int j) { return fine[static_cast<size_t>(j) * nxf + i]; };
      const double c = F(ifx, jf) * 0.25;
      const double e = (F(ifx - 1, jf) + F(ifx + 1, jf) + F(ifx, jf - 1) + F(ifx, jf + 1)) * 0.125;
      const double d =
          (F(ifx - 1, jf - 1) + F(ifx + 1, jf - 1) + F(ifx - 1, jf + 1) + F(ifx + 1, jf + 1)) * 0.0625;
      coarse[static_cast<size_t>(jc) * nxc + ic] = c + e + d;
    }
  }
}

void RefProlongAdd(const std::vector<double>& coarse, int nxc, int nyc, std::vector<double>& fine,
                   int nxf, int nyf) {  // multigrid.cpp:38-67
  auto C = [&](int i, int j) { return coarse[static_cast<size_t>(j) * nxc + i]; };
  auto Fr = [&](int i, int j) -> double& { return fine[static_cast<size_t>(j) * nxf + i]; };
  for (int jc = 0; jc < nyc; ++jc)
    for (int ic = 0; ic < nxc; ++ic) {
      const double vc = C(ic, jc);
      const int ifx = 2 * ic, jf = 2 * jc;
      if (ifx < nxf && jf < nyf) Fr(ifx, jf) += vc;
      if (ic + 1 < nxc && ifx + 1 < nxf && jf < nyf) Fr(ifx + 1, jf) += 0.5 * (vc + C(ic + 1, jc));
      if (jc + 1 < nyc && ifx < nxf && jf + 1 < nyf) Fr(ifx, jf + 1) += 0.5 * (vc + C(ic, jc + 1));
      if (ic + 1 < nxc && jc + 1 < nyc && ifx + 1 < nxf && jf + 1 < nyf)
        Fr(ifx + 1, jf + 1) += 0.25 * (vc + C(ic + 1, jc) + C(ic, jc + 1) + C(ic + 1, jc + 1));
    }
}

void RefSmoothGs(int nx, int ny, double ax, double by, double center, const std::vector<double>& b,
                 std::vector<double>& u, int iters) {  // multigrid.cpp:69-94
  for (int it = 0; it < iters; ++it)
    for (int j = 1; j < ny - 1; ++j)
      for (int i = 1; i < nx - 1; ++i) {
        const size_t id = static_cast<size_t>(j) * nx + i;
        const double rhs = b[id] - ax * (u[id - 1] + u[id + 1]) - by * (u[id - nx] + u[id + nx]);
        u[id] = rhs / center;
      }
}

void RefResidual(int nx, int ny, double ax, double by, double center, const std::vector<double>& b,
                 const std::vector<double>& u, std::vector<double>& r) {  // multigrid.cpp:96-119
  r.assign(static_cast<size_t>(nx) * ny, 0.0);
  for (int j = 1; j < ny - 1; ++j)
    for (int i = 1; i < nx - 1; ++i) {
      const size_t id = static_cast<size_t>(j) * nx + i;
      const double Au = center * u[id] + ax * (u[id - 1] + u[id + 1]) + by * (u[id - nx] + u[id + nx]);
      r[id] = b[id] - Au;
    }
}

// MultigridVcyclePoisson (multigrid.cpp:121-222).
int RefMgVcycle(const Geo& fine, double a, double bcoef, double e, const std::vector<double>& bfine,
                std::vector<double>& ufine, int pre, int post, int coarse_iters, int max_levels) {
  if (!CanCoarsenOdd(fine.nx) || !CanCoarsenOdd(fine.ny)) return PDEHIP_ERR_UNSUPPORTED;
  struct Level { int nx, ny; double dx, dy; std::vector<double> u, b, r; };
  std::vector<Level> L;
  L.push_back({fine.nx, fine.ny, fine.dx, fine.dy, ufine, bfine, {}});
  const double lx = fine.xmax - fine.xmin, ly = fine.ymax - fine.ymin;
  while (static_cast<int>(L.size()) < max_levels) {
    const Level& prev = L.back();
    if (!CanCoarsenOdd(prev.nx) || !CanCoarsenOdd(prev.ny)) break;
    const int nxc = (prev.nx + 1) / 2, nyc = (prev.ny + 1) / 2;
    if (nxc < 3 || nyc < 3) break;
    Level c{nxc, nyc, lx / (nxc - 1), ly / (nyc - 1),
            std::vector<double>(static_cast<size_t>(nxc) * nyc, 0.0),
            std::vector<double>(static_cast<size_t>(nxc) * nyc, 0.0), {}};
    L.push_back(std::move(c));
    if (nxc <= 5 || nyc <= 5) break;
  }
  auto coeffs = [&](const Level& l, double* ax, double* by, double* center) {
    *ax = a / (l.dx * l.dx);
    *by = bcoef / (l.dy * l.dy);
    *center = -2.0 * (*ax) - 2.0 * (*by) + e;
  };
  const int nl = static_cast<int>(L.size());
  for (int l = 0; l < nl - 1; ++l) {
    double ax, by, center;
    coeffs(L[l], &ax, &by, &center);
    if (std::abs(center) < 1e-12) return PDEHIP_ERR_DEGENERATE;
    RefSmoothGs(L[l].nx, L[l].ny, ax, by, center, L[l].b, L[l].u, std::max(0, pre));
    RefResidual(L[l].nx, L[l].ny, ax, by, center, L[l].b, L[l].u, L[l].r);
    RefRestrictFW(L[l].r, L[l].nx, L[l + 1].b, L[l + 1].nx, L[l + 1].ny);
    std::fill(L[l + 1].u.begin(), L[l + 1].u.end(), 0.0);
  }
  {
    Level& c = L.back();
    double ax, by, center;
    coeffs(c, &ax, &by, &center);
    if (std::abs(center) < 1e-12) return PDEHIP_ERR_DEGENERATE;
    RefSmoothGs(c.nx, c.ny, ax, by, center, c.b, c.u, std::max(1, coarse_iters));
  }
  for (int l = nl - 2; l >= 0; --l) {
    RefProlongAdd(L[l + 1].u, L[l + 1].nx, L[l + 1].ny, L[l].u, L[l].nx, L[l].ny);
    double ax, by, center;
    coeffs(L[l], &ax, &by, &center);
    if (std::abs(center) < 1e-12) return PDEHIP_ERR_DEGENERATE;
    RefSmoothGs(L[l].nx, L[l].ny, ax, by, center, L[l].b, L[l].u, std::max(0, post));
  }
  ufine = L[0].u;
  return PDEHIP_OK;
}

// --- extension: geometric multigrid hierarchy for any 2D/3D folded operator.
// Levels halve every axis ((n+1)/2) while all axes are odd >= 5, up to max_levels, and stop
// after a level with an axis <= 5 (the reference rule, multigrid.cpp:150-170).
// Smoother: red-black Gauss-Seidel (red = (i+j+k) even); pre-smoothing red then black,
// post-smoothing black then red, so V(n,n) is a symmetric preconditioner.  Restriction is full
// weighting (tensor product of [1/4 1/2 1/4]), prolongation is (bi/tri)linear interpolation.
// Coarsest level: exact dense solve when it has <= 1000 interior unknowns, else
// coarse_iters symmetric red-black sweeps.  kappa (FLUX) is restricted by full weighting.
struct MgLevel {
  LevelOp op;
  std::vector<double> u, b, r;
  std::vector<double> dense_inv;   // coarsest level only
  std::vector<size_t> dense_map;   // interior node ids of the coarsest level
};

struct Hierarchy {
  std::vector<MgLevel> L;
  int pre = 2, post = 2, coarse_iters = 50;
};

// Full weighting.  Interior coarse nodes always; coarse boundary nodes either by injection
// (coefficient fields) or, with boundary_fw, by the weights of the fine nodes that exist (the
// transpose of interpolation, folded inward afterwards by FoldTranspose).
void RestrictFW(const Geo& gf, const std::vector<double>& fine, const Geo& gc,
                std::vector<double>& coarse, bool keep_boundary_injection, bool boundary_fw = false) {
  coarse.assign(gc.n, 0.0);
  static const double w1[3] = {0.25, 0.5, 0.25};
#pragma omp parallel for schedule(static)
  for (int kc = 0; kc < gc.nz; ++kc)
    for (int jc = 0; jc < gc.ny; ++jc)
      for (int ic = 0; ic < gc.nx; ++ic) {
        const size_t cid = gc.idx(ic, jc, kc);
        const int fi = 2 * ic, fj = 2 * jc, fk = gc.is3d ? 2 * kc : 0;
        if (IsBoundary(gc, ic, jc, kc) && !boundary_fw) {
          coarse[cid] = keep_boundary_injection ? fine[gf.idx(fi, fj, fk)] : 0.0;
          continue;
        }
        double s = 0.0;
        for (int dk = (gc.is3d ? -1 : 0); dk <= (gc.is3d ? 1 : 0); ++dk)
          for (int dj = -1; dj <= 1; ++dj)
            for (int di = -1; di <= 1; ++di) {
              const int ii = fi + di, jj = fj + dj, kk = fk + dk;
              if (ii < 0 || jj < 0 || kk < 0 || ii >= gf.nx || jj >= gf.ny || kk >= gf.nz) continue;
              const double wgt = w1[di + 1] * w1[dj + 1] * (gc.is3d ? w1[dk + 1] : 1.0);
              s += wgt * fine[gf.idx(ii, jj, kk)];
            }
        coarse[cid] = s;
      }
}

// Homogeneous closure extension of a correction: face nodes = theta_f * inward neighbour
// (0 on Dirichlet faces), faces in the reference closure order (left, right, bottom, top,
// front, back) so edge and corner nodes take the later faces' values.
void ExtendClosure(const LevelOp& op, std::vector<double>& u) {
  const Geo& g = op.g;
  const double* th = op.fold.theta;
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j) {
      u[g.idx(0, j, k)] = th[PDEHIP_LEFT] * u[g.idx(1, j, k)];
      u[g.idx(g.nx - 1, j, k)] = th[PDEHIP_RIGHT] * u[g.idx(g.nx - 2, j, k)];
    }
  for (int k = 0; k < g.nz; ++k)
    for (int i = 0; i < g.nx; ++i) {
      u[g.idx(i, 0, k)] = th[PDEHIP_BOTTOM] * u[g.idx(i, 1, k)];
      u[g.idx(i, g.ny - 1, k)] = th[PDEHIP_TOP] * u[g.idx(i, g.ny - 2, k)];
    }
  if (g.is3d)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        u[g.idx(i, j, 0)] = th[PDEHIP_FRONT] * u[g.idx(i, j, 1)];
        u[g.idx(i, j, g.nz - 1)] = th[PDEHIP_BACK] * u[g.idx(i, j, g.nz - 2)];
      }
}

// Transpose of ExtendClosure: faces in reverse order, each boundary value folded into its
// inward neighbour with theta_f and then cleared.
void FoldTranspose(const LevelOp& op, std::vector<double>& r) {
  const Geo& g = op.g;
  const double* th = op.fold.theta;
  if (g.is3d)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        r[g.idx(i, j, g.nz - 2)] += th[PDEHIP_BACK] * r[g.idx(i, j, g.nz - 1)];
        r[g.idx(i, j, g.nz - 1)] = 0.0;
        r[g.idx(i, j, 1)] += th[PDEHIP_FRONT] * r[g.idx(i, j, 0)];
        r[g.idx(i, j, 0)] = 0.0;
      }
  for (int k = 0; k < g.nz; ++k)
    for (int i = 0; i < g.nx; ++i) {
      r[g.idx(i, g.ny - 2, k)] += th[PDEHIP_TOP] * r[g.idx(i, g.ny - 1, k)];
      r[g.idx(i, g.ny - 1, k)] = 0.0;
      r[g.idx(i, 1, k)] += th[PDEHIP_BOTTOM] * r[g.idx(i, 0, k)];
      r[g.idx(i, 0, k)] = 0.0;
    }
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j) {
      r[g.idx(g.nx - 2, j, k)] += th[PDEHIP_RIGHT] * r[g.idx(g.nx - 1, j, k)];
      r[g.idx(g.nx - 1, j, k)] = 0.0;
      r[g.idx(1, j, k)] += th[PDEHIP_LEFT] * r[g.idx(0, j, k)];
      r[g.idx(0, j, k)] = 0.0;
    }
}

void ProlongAdd(const Geo& gc, const std::vector<double>& coarse, const Geo& gf,
                std::vector<double>& fine) {
#pragma omp parallel for schedule(static)
  for (int k = 0; k < gf.nz; ++k)
    for (int j = 0; j < gf.ny; ++j)
      for (int i = 0; i < gf.nx; ++i) {
        if (IsBoundary(gf, i, j, k)) continue;
        const int ic = i / 2, jc = j / 2, kc = k / 2;
        const int oi = i & 1, oj = j & 1, ok = gf.is3d ? (k & 1) : 0;
        double s = 0.0;
        for (int a = 0; a <= ok; ++a)
          for (int bb = 0; bb <= oj; ++bb)
            for (int c = 0; c <= oi; ++c) s += coarse[gc.idx(ic + c, jc + bb, kc + a)];
        const int cnt = (1 + oi) * (1 + oj) * (1 + ok);
        fine[gf.idx(i, j, k)] += s / cnt;
      }
}

void ResidualInterior(const LevelOp& op, const std::vector<double>& b, const std::vector<double>& u,
                      std::vector<double>& r) {
  const Geo& g = op.g;
  r.assign(g.n, 0.0);
#pragma omp parallel for schedule(static)
  for (int k = (g.is3d ? 1 : 0); k < (g.is3d ? g.nz - 1 : 1); ++k)
    for (int j = 1; j < g.ny - 1; ++j)
      for (int i = 1; i < g.nx - 1; ++i) {
        const size_t id = g.idx(i, j, k);
        const W7& w = op.w[id];
        r[id] = b[id] - (w.C * u[id] + NeighbourSum(w, g, u.data(), i, j, k));
      }
}

void RbSweep(const LevelOp& op, const std::vector<double>& b, std::vector<double>& u, int color) {
  const Geo& g = op.g;
#pragma omp parallel for schedule(static)
  for (int k = (g.is3d ? 1 : 0); k < (g.is3d ? g.nz - 1 : 1); ++k)
    for (int j = 1; j < g.ny - 1; ++j)
      for (int i = 1 + ((1 + j + k + color) & 1); i < g.nx - 1; i += 2) {
        const size_t id = g.idx(i, j, k);
        const W7& w = op.w[id];
        u[id] = (b[id] - NeighbourSum(w, g, u.data(), i, j, k)) / w.C;
      }
}

bool BuildDenseInverse(MgLevel& lv) {
  const Geo& g = lv.op.g;
  lv.dense_map.clear();
  std::vector<long> pos(g.n, -1);
  for (int k = (g.is3d ? 1 : 0); k < (g.is3d ? g.nz - 1 : 1); ++k)
    for (int j = 1; j < g.ny - 1; ++j)
      for (int i = 1; i < g.nx - 1; ++i) {
        pos[g.idx(i, j, k)] = static_cast<long>(lv.dense_map.size());
        lv.dense_map.push_back(g.idx(i, j, k));
      }
  const size_t m = lv.dense_map.size();
  if (m == 0 || m > 1000) return false;
  std::vector<double> A(m * m, 0.0);
  for (size_t r = 0; r < m; ++r) {
    const size_t id = lv.dense_map[r];
    const int i = static_cast<int>(id % g.nx);
    const int j = static_cast<int>((id / g.nx) % g.ny);
    const int k = static_cast<int>(id / (static_cast<size_t>(g.nx) * g.ny));
    const W7& w = lv.op.w[id];
    A[r * m + r] = w.C;
    auto put = [&](int ii, int jj, int kk, double wv) {
      const long c = pos[g.idx(ii, jj, kk)];
      if (c >= 0) A[r * m + static_cast<size_t>(c)] += wv;
    };
    put(i - 1, j, k, w.W); put(i + 1, j, k, w.E); put(i, j - 1, k, w.S); put(i, j + 1, k, w.N);
    if (g.is3d) { put(i, j, k - 1, w.B); put(i, j, k + 1, w.F); }
  }
  // Gauss-Jordan inverse with partial pivoting.
  std::vector<double> inv(m * m, 0.0);
  for (size_t r = 0; r < m; ++r) inv[r * m + r] = 1.0;
  for (size_t c = 0; c < m; ++c) {
    size_t piv = c;
    for (size_t r = c + 1; r < m; ++r)
      if (std::abs(A[r * m + c]) > std::abs(A[piv * m + c])) piv = r;
    if (std::abs(A[piv * m + c]) < 1e-300) return false;
    if (piv != c)
      for (size_t q = 0; q < m; ++q) { std::swap(A[c * m + q], A[piv * m + q]); std::swap(inv[c * m + q], inv[piv * m + q]); }
    const double d = 1.0 / A[c * m + c];
    for (size_t q = 0; q < m; ++q) { A[c * m + q] *= d; inv[c * m + q] *= d; }
    for (size_t r = 0; r < m; ++r) {
      if (r == c) continue;
      const double fct = A[r * m + c];
      if (fct == 0.0) continue;
      for (size_t q = 0; q < m; ++q) { A[r * m + q] -= fct * A[c * m + q]; inv[r * m + q] -= fct * inv[c * m + q]; }
    }
  }
  lv.dense_inv.swap(inv);
  return true;
}

int LevelCount(const Geo& g0, int max_levels) {
  int levels = 1;
  Geo g = g0;
  while (levels < max_levels) {
    if (!CanCoarsenOdd(g.nx) || !CanCoarsenOdd(g.ny) || (g.is3d && !CanCoarsenOdd(g.nz))) break;
    Geo c = Coarsen(g);
    if (c.nx < 3 || c.ny < 3 || (g.is3d && c.nz < 3)) break;
    ++levels;
    g = c;
    if (c.nx <= 5 || c.ny <= 5 || (g.is3d && c.nz <= 5)) break;
  }
  return levels;
}

void BuildHierarchy(const pdehip_problem& p, const Geo& g0, const pdehip_options& o, Hierarchy& H) {
  const int nl = LevelCount(g0, std::max(1, o.mg_max_levels));
  H.L.assign(nl, MgLevel());
  H.pre = std::max(0, o.mg_pre);
  H.post = std::max(0, o.mg_post);
  H.coarse_iters = std::max(1, o.mg_coarse_iters);
  Geo g = g0;
  std::vector<double> kap;
  if (p.op_kind == PDEHIP_OP_FLUX) kap.assign(p.kappa, p.kappa + g0.n);
  for (int l = 0; l < nl; ++l) {
    H.L[l].op.Build(p, g, true, kap.empty() ? nullptr : kap.data());
    H.L[l].u.assign(g.n, 0.0);
    H.L[l].b.assign(g.n, 0.0);
    H.L[l].r.assign(g.n, 0.0);
    if (l + 1 < nl) {
      Geo c = Coarsen(g);
      if (!kap.empty()) {
        std::vector<double> kc;
        RestrictFW(g, kap, c, kc, true);
        kap.swap(kc);
      }
      g = c;
    }
  }
  BuildDenseInverse(H.L.back());
}

void CoarseSolve(Hierarchy& H) {
  MgLevel& c = H.L.back();
  if (!c.dense_inv.empty()) {
    const size_t m = c.dense_map.size();
    std::vector<double> x(m, 0.0);
    for (size_t r = 0; r < m; ++r) {
      double s = 0.0;
      for (size_t q = 0; q < m; ++q) s += c.dense_inv[r * m + q] * c.b[c.dense_map[q]];
      x[r] = s;
    }
    for (size_t r = 0; r < m; ++r) c.u[c.dense_map[r]] = x[r];
    return;
  }
  const int half = std::max(1, H.coarse_iters / 2);
  for (int it = 0; it < half; ++it) { RbSweep(c.op, c.b, c.u, 0); RbSweep(c.op, c.b, c.u, 1); }
  for (int it = 0; it < half; ++it) { RbSweep(c.op, c.b, c.u, 1); RbSweep(c.op, c.b, c.u, 0); }
}

// One V-cycle on level 0 with L[0].b as right-hand side and L[0].u as initial guess.
void VCycle(Hierarchy& H) {
  const int nl = static_cast<int>(H.L.size());
  for (int l = 0; l < nl - 1; ++l) {
    MgLevel& f = H.L[l];
    for (int s = 0; s < H.pre; ++s) { RbSweep(f.op, f.b, f.u, 0); RbSweep(f.op, f.b, f.u, 1); }
    ResidualInterior(f.op, f.b, f.u, f.r);
    RestrictFW(f.op.g, f.r, H.L[l + 1].op.g, H.L[l + 1].b, false, true);
    FoldTranspose(H.L[l + 1].op, H.L[l + 1].b);
    std::fill(H.L[l + 1].u.begin(), H.L[l + 1].u.end(), 0.0);
  }
  CoarseSolve(H);
  for (int l = nl - 2; l >= 0; --l) {
    MgLevel& f = H.L[l];
    ExtendClosure(H.L[l + 1].op, H.L[l + 1].u);
    ProlongAdd(H.L[l + 1].op.g, H.L[l + 1].u, f.op.g, f.u);
    for (int s = 0; s < H.post; ++s) { RbSweep(f.op, f.b, f.u, 1); RbSweep(f.op, f.b, f.u, 0); }
  }
}

// z = M r: one V-cycle from a zero guess (boundary rows of z are zero).
void MgPrecond(Hierarchy& H, const std::vector<double>& r, std::vector<double>& z) {
  MgLevel& f = H.L[0];
  f.b = r;
  std::fill(f.u.begin(), f.u.end(), 0.0);
  VCycle(H);
  z = f.u;
}

// --- extension: MG-preconditioned CG on the folded operator.  r0 = b - A x0 with x0 holding
// the boundary data, so boundary rows of r stay zero.  Stop rule: ||r||_2 < tol (ABS) or
// ||r||_2 < tol*||r0||_2 (REL).
KrylovResult MgPcg(const LevelOp& op, Hierarchy& H, const std::vector<double>& b,
                   std::vector<double>& x, int max_iter, double tol, int tol_mode, bool use_mg) {
  KrylovResult res;
  const size_t n = b.size();
  std::vector<double> r(n), z(n), p(n), q;
  op.Apply(x, q);
  for (size_t i = 0; i < n; ++i) r[i] = b[i] - q[i];
  const double r0 = Norm2(r);
  const double target = tol_mode == PDEHIP_TOL_REL ? tol * r0 : tol;
  res.metric = r0;
  if (r0 < target || r0 == 0.0) { res.ok = true; res.status = PDEHIP_OK; return res; }
  if (use_mg) MgPrecond(H, r, z); else z = r;
  p = z;
  double rz = Dot(r, z);
  for (int it = 0; it < max_iter; ++it) {
    op.Apply(p, q);
    const double pq = Dot(p, q);
    if (std::abs(pq) < 1e-300) { res.status = PDEHIP_ERR_BREAKDOWN; res.iters = it; return res; }
    const double alpha = rz / pq;
    for (size_t i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
    const double rn = Norm2(r);
    res.iters = it + 1;
    res.metric = rn;
    if (rn < target) { res.ok = true; res.status = PDEHIP_OK; return res; }
    if (use_mg) MgPrecond(H, r, z); else z = r;
    const double rz_new = Dot(r, z);
    const double beta = rz_new / rz;
    for (size_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    rz = rz_new;
  }
  return res;
}

// ----------------------------------------------------------------------------- relaxation
// Interior update of the reference relaxation loops (relaxation.cpp:179-287, 437-545;
// solver.cpp:827-959): u_new = (-(f + N(u_old)) - sum_nb w u_nb) / centre.
inline bool RelaxValue(const pdehip_problem& p, const Geo& g, const double* u, int i, int j, int k,
                       double* out) {
  const W7 w = WeightsAt(p, g, i, j, k);
  if (std::abs(w.C) < 1e-12) return false;
  const size_t id = g.idx(i, j, k);
  const double nl = p.n_nonlinear ? EvalNonlinear(p, u[id]) : 0.0;
  const double rhs = -(SourceAt(p, id) + nl) - NeighbourSum(w, g, u, i, j, k);
  *out = rhs / w.C;
  return true;
}

// SolveJacobi2D (relaxation.cpp:310-574) and the Jacobi branch of SolvePDE3D (solver.cpp:751-981).
int Jacobi(const pdehip_problem& p, const Geo& g, const pdehip_options& o, std::vector<double>& grid,
           int* iters, double* metric) {
  grid.assign(g.n, 0.0);
  std::vector<double> next(g.n, 0.0);
  ApplyDirichlet(p, g, grid);
  if (g.is3d) { ApplyNeumannRobin(p, g, grid); ApplyDirichlet(p, g, grid); }
  const int max_iter = std::max(1, o.max_iter);
  for (int it = 0; it < max_iter; ++it) {
    if (o.cancel && *o.cancel) return PDEHIP_ERR_CANCELLED;
    ApplyNeumannRobin(p, g, grid);
    double md = 0.0;
    bool degenerate = false;
#pragma omp parallel for reduction(max : md) reduction(|| : degenerate) schedule(static)
    for (int k = (g.is3d ? 1 : 0); k < (g.is3d ? g.nz - 1 : 1); ++k)
      for (int j = 1; j < g.ny - 1; ++j)
        for (int i = 1; i < g.nx - 1; ++i) {
          double v;
          if (!RelaxValue(p, g, grid.data(), i, j, k, &v)) { degenerate = true; continue; }
          const size_t id = g.idx(i, j, k);
          next[id] = v;
          md = std::max(md, std::abs(v - grid[id]));
        }
    if (degenerate) return PDEHIP_ERR_DEGENERATE;
    ApplyNeumannRobin(p, g, next);
    ApplyDirichlet(p, g, next);
    grid.swap(next);
    *iters = it + 1;
    *metric = md;
    if (md < o.tol) break;
  }
  return PDEHIP_OK;
}

// SolveGaussSeidelSor2D (relaxation.cpp:52-308) and the GS/SOR branch of SolvePDE3D:
// lexicographic in-place sweep; 2D applies closures only at the start of an iteration,
// 3D re-applies closures and Dirichlet values after the sweep (solver.cpp:973-974).
int GaussSeidelSor(const pdehip_problem& p, const Geo& g, const pdehip_options& o,
                   std::vector<double>& grid, int* iters, double* metric) {
  const double omega = o.method == PDEHIP_SOR ? o.sor_omega : 1.0;
  if (omega <= 0.0 || omega >= 2.0) return PDEHIP_ERR_INVALID;
  grid.assign(g.n, 0.0);
  ApplyDirichlet(p, g, grid);
  if (g.is3d) { ApplyNeumannRobin(p, g, grid); ApplyDirichlet(p, g, grid); }
  const int max_iter = std::max(1, o.max_iter);
  for (int it = 0; it < max_iter; ++it) {
    if (o.cancel && *o.cancel) return PDEHIP_ERR_CANCELLED;
    ApplyNeumannRobin(p, g, grid);
    double md = 0.0;
    for (int k = (g.is3d ? 1 : 0); k < (g.is3d ? g.nz - 1 : 1); ++k)
      for (int j = 1; j < g.ny - 1; ++j)
        for (int i = 1; i < g.nx - 1; ++i) {
          double v;
          if (!RelaxValue(p, g, grid.data(), i, j, k, &v)) return PDEHIP_ERR_DEGENERATE;
          const size_t id = g.idx(i, j, k);
          const double old = grid[id];
          const double upd = (1.0 - omega) * old + omega * v;
          grid[id] = upd;
          md = std::max(md, std::abs(upd - old));
        }
    if (g.is3d) { ApplyNeumannRobin(p, g, grid); ApplyDirichlet(p, g, grid); }
    *iters = it + 1;
    *metric = md;
    if (md < o.tol) break;
  }
  return PDEHIP_OK;
}

// ----------------------------------------------------------------------------- residual norms
struct Kahan {  // residual.cpp:54-63
  double sum = 0.0, c = 0.0;
  void Add(double v) { const double y = v - c; const double t = sum + y; c = (t - sum) - y; sum = t; }
};

// ComputeResidualNorms (residual.cpp:153-583).  2D includes the boundary closures with the
// residual's own convention (Neumann along the axis, Robin alpha+beta/h on every face,
// residual.cpp:286-383); 3D sums interior rows only (residual.cpp:591-740).
int ResidualNorms(const pdehip_problem& p, const Geo& g, const double* u, double* l2, double* linf) {
  Kahan s;
  double mx = 0.0;
  auto acc = [&](double r) { s.Add(r * r); mx = std::max(mx, std::abs(r)); };
  if (!g.is3d) {
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const size_t id = g.idx(i, j, 0);
        const double x = g.X(i), y = g.Y(j);
        if (IsBoundary(g, i, j, 0)) {
          int f; double inner; double h;
          if (i == 0) { f = PDEHIP_LEFT; inner = u[g.idx(1, j, 0)]; h = g.dx; }
          else if (i == g.nx - 1) { f = PDEHIP_RIGHT; inner = u[g.idx(g.nx - 2, j, 0)]; h = g.dx; }
          else if (j == 0) { f = PDEHIP_BOTTOM; inner = u[g.idx(i, 1, 0)]; h = g.dy; }
          else { f = PDEHIP_TOP; inner = u[g.idx(i, g.ny - 2, 0)]; h = g.dy; }
          const bool high = (f == PDEHIP_RIGHT || f == PDEHIP_TOP);
          const pdehip_face_bc& bc = p.bc[f];
          const double bx = (f == PDEHIP_LEFT) ? g.xmin : (f == PDEHIP_RIGHT ? g.xmax : x);
          const double byy = (f == PDEHIP_BOTTOM) ? g.ymin : (f == PDEHIP_TOP ? g.ymax : y);
          const size_t fi = FaceIndex(g, f, i, j, 0);
          if (bc.kind == PDEHIP_BC_DIRICHLET) {
            acc(u[id] - EvalExpr(bc.value, fi, bx, byy, 0.0));
          } else if (bc.kind == PDEHIP_BC_NEUMANN) {
            const double gv = EvalExpr(bc.value, fi, bx, byy, 0.0);
            acc(u[id] - (high ? inner + h * gv : inner - h * gv));
          } else {
            const double al = EvalExpr(bc.alpha, fi, bx, byy, 0.0);
            const double be = EvalExpr(bc.beta, fi, bx, byy, 0.0);
            const double ga = EvalExpr(bc.gamma, fi, bx, byy, 0.0);
            const double den = al + be / h;
            if (std::abs(den) > 1e-12) acc(u[id] - (ga + (be / h) * inner) / den);
          }
          continue;
        }
        const W7 w = WeightsAt(p, g, i, j, 0);
        if (std::abs(w.C) < 1e-12) return PDEHIP_ERR_DEGENERATE;
        const double nl = p.n_nonlinear ? EvalNonlinear(p, u[id]) : 0.0;
        const double base = NeighbourSum(w, g, u, i, j, 0);
        acc(w.C * u[id] + base + (SourceAt(p, id) + nl));
      }
  } else {
    for (int k = 1; k < g.nz - 1; ++k)
      for (int j = 1; j < g.ny - 1; ++j)
        for (int i = 1; i < g.nx - 1; ++i) {
          const size_t id = g.idx(i, j, k);
          const W7 w = WeightsAt(p, g, i, j, k);
          if (std::abs(w.C) < 1e-12) return PDEHIP_ERR_DEGENERATE;
          acc(w.C * u[id] + NeighbourSum(w, g, u, i, j, k) + SourceAt(p, id));
        }
  }
  *l2 = std::sqrt(std::max(0.0, s.sum));
  *linf = mx;
  return PDEHIP_OK;
}

// ----------------------------------------------------------------------------- time series
double SafeDivClamped(double num, double den) {  // include/safe_math.h:48-58
  if (std::abs(den) < 1e-12) return 0.0;
  const double r = num / den;
  return std::isfinite(r) ? r : 0.0;
}

// fill_dudt of SolvePDETimeSeries (solver.cpp:1182-1195): du/dt = -(L u + f + N(u))/ut on
// interior nodes, 0 on the boundary.
void FillDudt(const pdehip_problem& p, const Geo& g, const std::vector<double>& u,
              std::vector<double>& dudt) {
  dudt.assign(g.n, 0.0);
#pragma omp parallel for schedule(static)
  for (int j = 1; j < g.ny - 1; ++j)
    for (int i = 1; i < g.nx - 1; ++i) {
      const size_t id = g.idx(i, j, 0);
      const W7 w = WeightsAt(p, g, i, j, 0);
      const double opv = w.C * u[id] + NeighbourSum(w, g, u.data(), i, j, 0);
      const double nl = p.n_nonlinear ? EvalNonlinear(p, u[id]) : 0.0;
      const double rhs = opv + SourceAt(p, id) + nl;
      dudt[id] = SafeDivClamped(-rhs, p.ut);
    }
}

// One explicit step of TimeStep (time_integrator.cpp:49-216, 299-333).
void ExplicitStep(const pdehip_problem& p, const Geo& g, int integrator, double dt,
                  std::vector<double>& u) {
  const size_t n = u.size();
  std::vector<double> k1, k2, k3, k4, tmp(n);
  switch (integrator) {
    case PDEHIP_TI_RK2:
      FillDudt(p, g, u, k1);
      for (size_t i = 0; i < n; ++i) tmp[i] = u[i] + dt * k1[i];
      FillDudt(p, g, tmp, k2);
      for (size_t i = 0; i < n; ++i) u[i] += 0.5 * dt * (k1[i] + k2[i]);
      break;
    case PDEHIP_TI_RK4:
      FillDudt(p, g, u, k1);
      for (size_t i = 0; i < n; ++i) tmp[i] = u[i] + 0.5 * dt * k1[i];
      FillDudt(p, g, tmp, k2);
      for (size_t i = 0; i < n; ++i) tmp[i] = u[i] + 0.5 * dt * k2[i];
      FillDudt(p, g, tmp, k3);
      for (size_t i = 0; i < n; ++i) tmp[i] = u[i] + dt * k3[i];
      FillDudt(p, g, tmp, k4);
      for (size_t i = 0; i < n; ++i) u[i] += dt / 6.0 * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]);
      break;
    case PDEHIP_TI_SSPRK2: {
      std::vector<double> un(u);
      FillDudt(p, g, u, k1);
      for (size_t i = 0; i < n; ++i) u[i] = un[i] + dt * k1[i];
      FillDudt(p, g, u, k1);
      for (size_t i = 0; i < n; ++i) u[i] = 0.5 * un[i] + 0.5 * (u[i] + dt * k1[i]);
      break;
    }
    case PDEHIP_TI_SSPRK3: {
      std::vector<double> un(u), u1(n), u2(n);
      FillDudt(p, g, un, k1);
      for (size_t i = 0; i < n; ++i) u1[i] = un[i] + dt * k1[i];
      FillDudt(p, g, u1, k1);
      for (size_t i = 0; i < n; ++i) u2[i] = 0.75 * un[i] + 0.25 * (u1[i] + dt * k1[i]);
      FillDudt(p, g, u2, k1);
      for (size_t i = 0; i < n; ++i) u[i] = (1.0 / 3.0) * un[i] + (2.0 / 3.0) * (u2[i] + dt * k1[i]);
      break;
    }
    default: {  // forward Euler (time_integrator.cpp:49-69)
      FillDudt(p, g, u, k1);
      for (size_t i = 0; i < n; ++i) u[i] += dt * k1[i];
    }
  }
}

}  // namespace

// =============================================================================== C API
extern "C" {

const char* oracle_last_error(void) { return g_error.c_str(); }

void oracle_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int oracle_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// y = A x with reference rows (LinearOperator2D::Apply, cpu_utils.cpp:324-513; 3D rows
// by the SolvePDE3D stencil).
int oracle_apply(const pdehip_problem* p, const double* x, double* y) {
  const Geo g = MakeGeo(*p);
  LevelOp op;
  op.Build(*p, g, false, p->kappa);
  std::vector<double> xv(x, x + g.n), yv;
  op.Apply(xv, yv);
  std::copy(yv.begin(), yv.end(), y);
  return PDEHIP_OK;
}

// y = A x with folded Robin/Neumann rows (extension).
int oracle_apply_folded(const pdehip_problem* p, const double* x, double* y) {
  const Geo g = MakeGeo(*p);
  LevelOp op;
  op.Build(*p, g, true, p->kappa);
  std::vector<double> xv(x, x + g.n), yv;
  op.Apply(xv, yv);
  std::copy(yv.begin(), yv.end(), y);
  return PDEHIP_OK;
}

int oracle_build_rhs(const pdehip_problem* p, double* b) {
  const Geo g = MakeGeo(*p);
  const std::vector<double> bv = BuildRhs(*p, g);
  std::copy(bv.begin(), bv.end(), b);
  return PDEHIP_OK;
}

// which: 1 = Neumann/Robin closures, 2 = Dirichlet values, 3 = closures then Dirichlet.
int oracle_apply_bc(const pdehip_problem* p, double* u, int which) {
  const Geo g = MakeGeo(*p);
  std::vector<double> v(u, u + g.n);
  if (which & 1) ApplyNeumannRobin(*p, g, v);
  if (which & 2) ApplyDirichlet(*p, g, v);
  std::copy(v.begin(), v.end(), u);
  return PDEHIP_OK;
}

int oracle_residual_norms(const pdehip_problem* p, const double* u, double* l2, double* linf) {
  const Geo g = MakeGeo(*p);
  return ResidualNorms(*p, g, u, l2, linf);
}

// Steady solve, dispatch as SolvePDE (solver.cpp:491-595) plus the extensions.
int oracle_solve(const pdehip_problem* pp, const pdehip_options* oo, double* u_out,
                 pdehip_report* rep) {
  const pdehip_problem& p = *pp;
  const pdehip_options& o = *oo;
  pdehip_report dummy;
  if (!rep) rep = &dummy;
  std::memset(rep, 0, sizeof(*rep));
  const Geo g = MakeGeo(p);
  std::vector<double> x;
  int st = PDEHIP_OK;
  switch (o.method) {
    case PDEHIP_JACOBI:
      st = Jacobi(p, g, o, x, &rep->iterations, &rep->final_metric);
      rep->converged = rep->final_metric < o.tol;
      break;
    case PDEHIP_GAUSS_SEIDEL:
    case PDEHIP_SOR:
      st = GaussSeidelSor(p, g, o, x, &rep->iterations, &rep->final_metric);
      rep->converged = rep->final_metric < o.tol;
      break;
    case PDEHIP_CG:
    case PDEHIP_BICGSTAB:
    case PDEHIP_GMRES: {
      const bool folded = o.bc_rows == PDEHIP_ROWS_FOLDED;
      if (!folded && !AllDirichlet(p, g)) { g_error = "Krylov methods require Dirichlet boundaries on all sides"; return PDEHIP_ERR_UNSUPPORTED; }
      LevelOp op;
      op.Build(p, g, folded, p.kappa);
      std::vector<double> b = BuildRhs(p, g);
      if (folded) {
        x = FoldedBoundaryData(p, g);
        for (int k = 0; k < g.nz; ++k)
          for (int j = 0; j < g.ny; ++j)
            for (int i = 0; i < g.nx; ++i)
              if (IsBoundary(g, i, j, k)) b[g.idx(i, j, k)] = x[g.idx(i, j, k)];
      } else {
        x.assign(g.n, 0.0);
      }
      double tol = o.tol;
      if (o.tol_mode == PDEHIP_TOL_REL) {
        std::vector<double> ax;
        op.Apply(x, ax);
        double s = 0.0;
        for (size_t i = 0; i < g.n; ++i) s += (b[i] - ax[i]) * (b[i] - ax[i]);
        tol = o.tol * std::sqrt(s);
      }
      Precond M;
      M.op = &op;
      M.jacobi = o.precond == PDEHIP_PC_JACOBI;
      KrylovResult kr;
      const int max_iter = std::max(1, o.max_iter);
      if (o.method == PDEHIP_CG) {
        if (M.jacobi) { g_error = "CG preconditioning: use PDEHIP_MGPCG"; return PDEHIP_ERR_UNSUPPORTED; }
        kr = Cg(op, b, x, max_iter, tol);
      } else if (o.method == PDEHIP_BICGSTAB) {
        kr = BiCGStab(op, M, b, x, max_iter, tol);
      } else {
        kr = Gmres(op, M, b, x, max_iter, o.gmres_restart, tol);
      }
      rep->iterations = kr.iters;
      rep->final_metric = kr.metric;
      rep->converged = kr.ok;
      if (folded) ReconstructBoundary(p, g, x);
      st = kr.ok ? PDEHIP_OK : kr.status;
      break;
    }
    case PDEHIP_MG: {
      if (o.mg_single_cycle) {
        // SolvePDE MultigridVcycle branch (solver.cpp:549-592).
        if (g.is3d || !AllDirichlet(p, g) || p.op_kind != PDEHIP_OP_CONST) {
          g_error = "reference multigrid: 2D, constant coefficients, Dirichlet only";
          return PDEHIP_ERR_UNSUPPORTED;
        }
        std::vector<double> b = BuildRhs(p, g);
        x.assign(g.n, 0.0);
        ApplyDirichlet(p, g, x);
        st = RefMgVcycle(g, p.a, p.b, p.e, b, x, o.mg_pre, o.mg_post, o.mg_coarse_iters,
                         o.mg_max_levels);
        rep->iterations = 1;
        rep->converged = st == PDEHIP_OK;
        break;
      }
      // extension: V-cycles (RB-GS) iterated until ||b - A u||_2 < tol.
      Hierarchy H;
      BuildHierarchy(p, g, o, H);
      std::vector<double> b = BuildRhs(p, g);
      x = FoldedBoundaryData(p, g);
      for (int k = 0; k < g.nz; ++k)
        for (int j = 0; j < g.ny; ++j)
          for (int i = 0; i < g.nx; ++i)
            if (IsBoundary(g, i, j, k)) b[g.idx(i, j, k)] = x[g.idx(i, j, k)];
      const LevelOp& op = H.L[0].op;
      std::vector<double> r;
      ResidualInterior(op, b, x, r);
      const double r0 = Norm2(r);
      const double target = o.tol_mode == PDEHIP_TOL_REL ? o.tol * r0 : o.tol;
      double rn = r0;
      int it = 0;
      const int max_iter = std::max(1, o.max_iter);
      while (rn >= target && it < max_iter) {
        MgLevel& f = H.L[0];
        f.b = b;
        f.u = x;
        VCycle(H);
        x = f.u;
        ResidualInterior(op, b, x, r);
        rn = Norm2(r);
        ++it;
      }
      rep->iterations = it;
      rep->final_metric = rn;
      rep->converged = rn < target;
      ReconstructBoundary(p, g, x);
      st = rep->converged ? PDEHIP_OK : PDEHIP_ERR_NOT_CONVERGED;
      break;
    }
    case PDEHIP_MGPCG: {
      Hierarchy H;
      BuildHierarchy(p, g, o, H);
      std::vector<double> b = BuildRhs(p, g);
      x = FoldedBoundaryData(p, g);
      for (int k = 0; k < g.nz; ++k)
        for (int j = 0; j < g.ny; ++j)
          for (int i = 0; i < g.nx; ++i)
            if (IsBoundary(g, i, j, k)) b[g.idx(i, j, k)] = x[g.idx(i, j, k)];
      const auto t0 = std::chrono::steady_clock::now();
      KrylovResult kr = MgPcg(H.L[0].op, H, b, x, std::max(1, o.max_iter), o.tol, o.tol_mode, true);
      rep->solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      rep->iterations = kr.iters;
      rep->final_metric = kr.metric;
      rep->converged = kr.ok;
      ReconstructBoundary(p, g, x);
      st = kr.ok ? PDEHIP_OK : kr.status;
      break;
    }
    default:
      g_error = "unknown method";
      return PDEHIP_ERR_INVALID;
  }
  if (st == PDEHIP_ERR_DEGENERATE) g_error = "degenerate PDE center coefficient";
  if (!x.empty() && u_out) std::copy(x.begin(), x.end(), u_out);
  if (!x.empty()) ResidualNorms(p, g, x.data(), &rep->residual_l2, &rep->residual_linf);
  rep->status = st;
  return st;
}

// Number of multigrid levels the extension hierarchy builds.
int oracle_mg_levels(const pdehip_problem* p, int max_levels) {
  return LevelCount(MakeGeo(*p), max_levels);
}

// One V-cycle of the extension hierarchy applied to b from a zero guess (z = M b).
int oracle_mg_precond(const pdehip_problem* p, const pdehip_options* o, const double* b,
                      double* z) {
  const Geo g = MakeGeo(*p);
  Hierarchy H;
  BuildHierarchy(*p, g, *o, H);
  std::vector<double> bv(b, b + g.n), zv;
  MgPrecond(H, bv, zv);
  std::copy(zv.begin(), zv.end(), z);
  return PDEHIP_OK;
}

// SolvePDETimeSeries (solver.cpp:989-1329), 2D explicit integrators, no u_tt with
// integrators (the reference routes u_tt to its inline semi-implicit Euler, solver.cpp:1307-1311).
// frames_out (frames * N, may be NULL) receives every frame; u_out the last one.
int oracle_time_series(const pdehip_problem* pp, const pdehip_time_config* tc, double* frames_out,
                       double* u_out) {
  const pdehip_problem& p = *pp;
  const Geo g = MakeGeo(p);
  if (g.is3d) { g_error = "3D Time Series Solver not yet refactored."; return PDEHIP_ERR_UNSUPPORTED; }
  if (tc->dt <= 0.0) { g_error = "time step dt must be positive"; return PDEHIP_ERR_INVALID; }
  const bool has_ut = std::abs(p.ut) > 1e-12, has_utt = std::abs(p.utt) > 1e-12;
  if (!has_ut && !has_utt) { g_error = "time-dependent solve requires u_t or u_tt term"; return PDEHIP_ERR_INVALID; }
  const int frames = std::max(1, tc->frames);
  std::vector<double> grid(g.n, 0.0), vel;
  if (p.initial) grid.assign(p.initial, p.initial + g.n);
  if (has_utt) {
    vel.assign(g.n, 0.0);
    if (p.initial_velocity) vel.assign(p.initial_velocity, p.initial_velocity + g.n);
  }
  ApplyDirichlet(p, g, grid);
  ApplyNeumannRobin(p, g, grid);
  ApplyDirichlet(p, g, grid);
  const bool use_integrator = !has_utt && tc->integrator != PDEHIP_TI_FORWARD_EULER;
  std::vector<double> next(g.n);
  for (int frame = 0; frame < frames; ++frame) {
    if (frames_out) std::copy(grid.begin(), grid.end(), frames_out + static_cast<size_t>(frame) * g.n);
    if (frame + 1 >= frames) break;
    if (use_integrator) {
      ExplicitStep(p, g, tc->integrator, tc->dt, grid);
    } else {
      // inline forward Euler / u_tt update (solver.cpp:1283-1318)
      next = grid;
      for (int j = 1; j < g.ny - 1; ++j)
        for (int i = 1; i < g.nx - 1; ++i) {
          const size_t id = g.idx(i, j, 0);
          const W7 w = WeightsAt(p, g, i, j, 0);
          const double opv = w.C * grid[id] + NeighbourSum(w, g, grid.data(), i, j, 0);
          const double nl = p.n_nonlinear ? EvalNonlinear(p, grid[id]) : 0.0;
          const double rhs = opv + SourceAt(p, id) + nl;
          if (has_utt) {
            const double accel = SafeDivClamped(-(rhs + p.ut * vel[id]), p.utt);
            vel[id] += tc->dt * accel;
            next[id] = grid[id] + tc->dt * vel[id];
          } else {
            next[id] = grid[id] + tc->dt * SafeDivClamped(-rhs, p.ut);
          }
        }
      grid.swap(next);
    }
    ApplyNeumannRobin(p, g, grid);
    ApplyDirichlet(p, g, grid);
  }
  if (u_out) std::copy(grid.begin(), grid.end(), u_out);
  return PDEHIP_OK;
}

}  // extern "C"
