// hip_self_test.cpp — the reference's self-test cases (src/self_test.cpp:168-411) run through
// SolveWithBackend(..., BackendKind::HIP), plus HIP-only cases (3D Krylov/multigrid with Robin
// faces, the time-series entry).  Prints one [OK]/[FAIL] line per case; exit code 0 when all pass.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <functional>
#include <iostream>
#include <string>
#include <vector>

#include "backend.h"
#include "backend_capability_matrix.h"
#include "hip_solver.h"
#include "solver.h"

namespace {

BoundaryCondition::Expression Linear(double c, double x, double y, double z = 0.0) {
  BoundaryCondition::Expression e;
  e.constant = c; e.x = x; e.y = y; e.z = z;
  return e;
}
BoundaryCondition::Expression Latex(const std::string& s) {
  BoundaryCondition::Expression e;
  e.latex = s;
  return e;
}
BoundarySet DirichletAll(const BoundaryCondition::Expression& e) {
  BoundarySet bc;
  for (BoundaryCondition* f : {&bc.left, &bc.right, &bc.bottom, &bc.top, &bc.front, &bc.back}) {
    f->kind = BCKind::Dirichlet;
    f->value = e;
  }
  return bc;
}
Domain Unit(int n, int nz = 1) {
  Domain d;
  d.nx = n; d.ny = n; d.nz = nz;
  return d;
}
void Configure(SolveInput* in, SolveMethod m, int max_iter, double tol) {
  in->solver = {};
  in->solver.method = m;
  in->solver.max_iter = max_iter;
  in->solver.tol = tol;
  in->solver.sor_omega = 1.7;      // self_test.cpp:121
  in->solver.gmres_restart = 30;
}

// ComputeErrorNorms2D/3D-like max norm against the exact solution.
double ErrInf(const Domain& d, const std::vector<double>& u,
              const std::function<double(double, double, double)>& exact) {
  const double dx = (d.xmax - d.xmin) / (d.nx - 1), dy = (d.ymax - d.ymin) / (d.ny - 1);
  const double dz = d.nz > 1 ? (d.zmax - d.zmin) / (d.nz - 1) : 0.0;
  double m = 0.0;
  for (int k = 0; k < std::max(1, d.nz); ++k)
    for (int j = 0; j < d.ny; ++j)
      for (int i = 0; i < d.nx; ++i) {
        const size_t id = ((size_t)k * d.ny + j) * d.nx + i;
        m = std::max(m, std::abs(u[id] - exact(d.xmin + i * dx, d.ymin + j * dy, d.zmin + k * dz)));
      }
  return m;
}

int g_fail = 0, g_total = 0;

void Run(const std::string& name, const SolveInput& in,
         const std::function<double(double, double, double)>& exact, double max_err) {
  ++g_total;
  BackendKind sel = BackendKind::Auto;
  std::string note;
  SolveOutput out = SolveWithBackend(in, BackendKind::HIP, &sel, &note);
  if (!out.error.empty()) {
    std::cout << "[FAIL] " << name << ": " << out.error << "\n";
    ++g_fail;
    return;
  }
  const double e = ErrInf(in.domain, out.grid, exact);
  const bool pass = e <= max_err && sel == BackendKind::HIP;
  std::cout << (pass ? "[OK] " : "[FAIL] ") << name << " backend=" << BackendKindName(sel)
            << " errinf=" << e << " tol=" << max_err << " resid2=" << out.residual_l2
            << " residinf=" << out.residual_linf << "\n";
  if (!pass) ++g_fail;
}

}  // namespace

int main(int argc, char** argv) {
  const std::string vtk_path = argc > 1 ? argv[1] : "/tmp/pdehip_cfg1_129.vtk";
  std::string note;
  if (!HipIsAvailable(&note)) {
    std::cout << "[FAIL] HIP unavailable: " << note << "\n";
    return 1;
  }
  std::cout << "HIP: " << note << "\n";
  const auto harmonic = [](double x, double y, double) { return x + 2.0 * y; };
  const auto zero = [](double, double, double) { return 0.0; };

  // Solver coverage (self_test.cpp:241-277): 33^2, tol 1e-8.
  SolveInput h;
  h.domain = Unit(33);
  h.pde.a = -1.0; h.pde.b = -1.0;
  h.bc = DirichletAll(Linear(0.0, 1.0, 2.0));
  SolveInput z = h;
  z.bc = DirichletAll(Linear(0.0, 0.0, 0.0));
  struct M { SolveMethod m; const char* name; } methods[] = {
      {SolveMethod::SOR, "SOR"}, {SolveMethod::CG, "CG"}, {SolveMethod::BiCGStab, "BiCGStab"},
      {SolveMethod::GMRES, "GMRES"}, {SolveMethod::MultigridVcycle, "MG"}};
  for (const M& m : methods) {
    const bool krylov = m.m == SolveMethod::CG || m.m == SolveMethod::BiCGStab || m.m == SolveMethod::GMRES;
    SolveInput in = krylov ? z : h;
    Configure(&in, m.m, m.m == SolveMethod::MultigridVcycle ? 120 : 4000, 1e-8);
    const double thr = krylov ? 1e-8 : (m.m == SolveMethod::MultigridVcycle ? 5e-6 : 1e-6);
    Run(std::string(krylov ? "Poisson zero (" : "Poisson harmonic (") + m.name + ")", in,
        krylov ? zero : harmonic, thr);
  }
  // CG on non-zero Dirichlet data (the reference's CG diverges on these rows)
  {
    SolveInput in = h;
    Configure(&in, SolveMethod::CG, 4000, 1e-10);
    Run("Poisson harmonic (CG, folded Dirichlet rows)", in, harmonic, 1e-9);
  }
  // Regression suite (self_test.cpp:279-340): 21^2, SOR, tol 1e-10.
  {
    SolveInput in;
    in.domain = Unit(21);
    in.pde.a = -1.0; in.pde.b = -1.0; in.pde.e = 1.0; in.pde.f = -1.0;
    in.bc = DirichletAll(Linear(1.0, 0.0, 0.0));
    Configure(&in, SolveMethod::SOR, 6000, 1e-10);
    Run("Helmholtz constant (SOR)", in, [](double, double, double) { return 1.0; }, 1e-7);
  }
  {
    SolveInput in;
    in.domain = Unit(21);
    in.pde.a = -2.0; in.pde.b = -3.0; in.pde.f = 10.0;
    in.bc = DirichletAll(Latex("x^2 + y^2"));
    Configure(&in, SolveMethod::SOR, 6000, 1e-10);
    Run("Anisotropic diffusion (SOR)", in, [](double x, double y, double) { return x * x + y * y; }, 1e-7);
  }
  // MMS convergence (self_test.cpp:168-221): CG, 17^2 vs 33^2, ratio > 3.
  {
    ++g_total;
    SolveInput in;
    in.pde.a = -1.0; in.pde.b = -1.0;
    in.bc = DirichletAll(Latex("\\sin(\\pi*x)*\\sin(\\pi*y)"));
    in.pde.rhs_latex = "-(-(-\\pi^2*\\sin(\\pi*x)*\\sin(\\pi*y))-(-\\pi^2*\\sin(\\pi*x)*\\sin(\\pi*y)))";
    Configure(&in, SolveMethod::CG, 8000, 1e-10);
    const auto ex = [](double x, double y, double) { return std::sin(M_PI * x) * std::sin(M_PI * y); };
    in.domain = Unit(17);
    SolveOutput c = SolveWithBackend(in, BackendKind::HIP, nullptr, nullptr);
    in.domain = Unit(33);
    SolveOutput f = SolveWithBackend(in, BackendKind::HIP, nullptr, nullptr);
    if (!c.error.empty() || !f.error.empty()) {
      std::cout << "[FAIL] MMS convergence (CG): " << c.error << f.error << "\n";
      ++g_fail;
    } else {
      const double ec = ErrInf(Unit(17), c.grid, ex), ef = ErrInf(Unit(33), f.grid, ex);
      const double ratio = ec / std::max(1e-14, ef);
      std::cout << (ratio > 3.0 ? "[OK] " : "[FAIL] ") << "MMS convergence (CG) errinf_coarse=" << ec
                << " errinf_fine=" << ef << " ratio=" << ratio << "\n";
      if (!(ratio > 3.0)) ++g_fail;
    }
  }
  // HIP-only: 3D Krylov and multigrid with Robin faces (outward), manufactured solution.
  {
    SolveInput in;
    in.domain = Unit(33, 33);
    in.pde.a = in.pde.b = in.pde.az = -1.0;
    in.pde.rhs_latex = "-3*\\pi^2*\\sin(\\pi*x)*\\sin(\\pi*y)*\\sin(\\pi*z)";
    in.bc = DirichletAll(Linear(0.0, 0.0, 0.0));
    const auto ex = [](double x, double y, double zz) {
      return std::sin(M_PI * x) * std::sin(M_PI * y) * std::sin(M_PI * zz);
    };
    Configure(&in, SolveMethod::CG, 4000, 1e-10);
    Run("3D Poisson (CG)", in, ex, 2e-3);
    Configure(&in, SolveMethod::MultigridVcycle, 100, 1e-10);
    HipSolverOptions ho;
    ho.mg_pcg = true;
    HipSetSolverOptions(ho);
    Run("3D Poisson (MG-PCG)", in, ex, 2e-3);
    ho.mixed_precision = true;
    HipSetSolverOptions(ho);
    Run("3D Poisson (MG-PCG mixed precision)", in, ex, 2e-3);
    HipSetSolverOptions(HipSolverOptions());
  }
  {
    // Robin (outward) faces: u = cos(pi x/2)... use a closed form u = 1 + x with
    // du/dn on x faces handled by Robin alpha*u + beta*du/dn = gamma
    SolveInput in;
    in.domain = Unit(33);
    in.pde.a = -1.0; in.pde.b = -1.0;
    in.bc = DirichletAll(Linear(1.0, 1.0, 0.0));
    in.bc.left.kind = BCKind::Robin;     // at x=0: u=1, du/dn = -1 -> 1*u + 1*(-1) = 0
    in.bc.left.alpha = Linear(1.0, 0, 0); in.bc.left.beta = Linear(1.0, 0, 0); in.bc.left.gamma = Linear(0.0, 0, 0);
    in.bc.right.kind = BCKind::Neumann;  // du/dx = 1 (axis convention = outward on the high face)
    in.bc.right.value = Linear(1.0, 0, 0);
    Configure(&in, SolveMethod::CG, 4000, 1e-11);
    HipSolverOptions ho;
    ho.outward_robin = true;
    HipSetSolverOptions(ho);
    Run("Linear solution, Robin + Neumann faces (CG folded rows)", in,
        [](double x, double, double) { return 1.0 + x; }, 1e-8);
    // absolute L2 tolerance: 1e-11 sits at the FP64 rounding floor of rows scaled by 1/h^2
    Configure(&in, SolveMethod::MultigridVcycle, 200, 1e-10);
    Run("Linear solution, Robin + Neumann faces (MG)", in, [](double x, double, double) { return 1.0 + x; }, 1e-8);
    HipSetSolverOptions(HipSolverOptions());
  }
  // Capability matrix: HIP lifts the 2D/Dirichlet-only gates, keeps the shape gate.
  {
    ++g_total;
    SolveInput in;
    in.domain = Unit(9, 9);
    in.bc.front.kind = BCKind::Robin;
    std::string r;
    const bool hip_ok = BackendSupportsMethodForInput(BackendKind::HIP, SolveMethod::CG, in, &r);
    const bool cuda_ok = BackendSupportsMethodForInput(BackendKind::CUDA, SolveMethod::CG, in, &r);
    in.domain_shape = "x^2+y^2-0.25";
    const bool shape_ok = BackendSupportsMethodForInput(BackendKind::HIP, SolveMethod::Jacobi, in, &r);
    const bool pass = hip_ok && !cuda_ok && !shape_ok && ParseBackendKind("hip") == BackendKind::HIP;
    std::cout << (pass ? "[OK] " : "[FAIL] ") << "capability matrix HIP row\n";
    if (!pass) ++g_fail;
  }
  // Time series: heat equation decay of sin(pi x) sin(pi y): u(t) ~ exp(-2 pi^2 t) (FE, RK4).
  for (TimeIntegrator ti : {TimeIntegrator::ForwardEuler, TimeIntegrator::RK4}) {
    ++g_total;
    SolveInput in;
    in.domain = Unit(33);
    in.pde.a = -1.0; in.pde.b = -1.0; in.pde.ut = 1.0;
    in.bc = DirichletAll(Linear(0.0, 0.0, 0.0));
    const int n = 33;
    in.initial_grid.resize(n * n);
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i)
        in.initial_grid[j * n + i] = std::sin(M_PI * i / 32.0) * std::sin(M_PI * j / 32.0);
    in.time.enabled = true;
    in.time.dt = 0.2 / (32.0 * 32.0);
    in.time.frames = 201;
    in.time.integrator = ti;
    int frames_seen = 0;
    std::vector<double> last;
    SolveOutput out = SolvePDETimeSeriesHip(in, [&](int, double, const std::vector<double>& g,
                                                    const std::vector<double>*) {
      ++frames_seen;
      last = g;
      return true;
    });
    const double t = in.time.dt * 200;
    const double h = 1.0 / 32.0;
    const double lam = 8.0 / (h * h) * std::pow(std::sin(M_PI * h / 2.0), 2);  // discrete eigenvalue
    const double expect = std::exp(-lam * t);
    const double mid = last.empty() ? 0.0 : last[16 * n + 16];
    const bool pass = out.error.empty() && frames_seen == 201 && std::abs(mid - expect) < 5e-3 * expect;
    std::cout << (pass ? "[OK] " : "[FAIL] ") << "heat decay time series ("
              << (ti == TimeIntegrator::RK4 ? "RK4" : "FE") << ") u_mid=" << mid << " expect=" << expect
              << " frames=" << frames_seen << " " << out.error << "\n";
    if (!pass) ++g_fail;
  }
  // The plain CPU workload end to end on HIP: 129^2 Poisson, f = 2 pi^2 sin sin, CG to rel-resid
  // 1e-10, written as .vtk and read back.  Max error against sin sin is the closed form
  // 2 pi^2 / lambda_h - 1 = 5.020092e-05; 17 digits round-trip the grid exactly.
  {
    ++g_total;
    SolveInput in;
    in.domain = Unit(129);
    in.pde.a = -1.0; in.pde.b = -1.0;
    in.pde.rhs_latex = "-2*\\pi^2*\\sin(\\pi*x)*\\sin(\\pi*y)";
    in.bc = DirichletAll(Linear(0.0, 0.0, 0.0));
    Configure(&in, SolveMethod::CG, 2000, 1e-10);
    HipSolverOptions ho;
    ho.relative_tol = true;
    HipSetSolverOptions(ho);
    BackendKind sel = BackendKind::Auto;
    std::string sel_note;
    const SolveOutput out = SolveWithBackend(in, BackendKind::HIP, &sel, &sel_note);
    HipSetSolverOptions(HipSolverOptions());
    const std::string path = vtk_path;
    std::string err;
    bool ok = out.error.empty() && HipWriteVtk(path, in.domain, out.grid, 17, &err);
    std::vector<double> back;
    int dims[3] = {0, 0, 0};
    if (ok) {
      std::ifstream f(path);
      std::string line, word;
      while (std::getline(f, line) && line.rfind("LOOKUP_TABLE", 0) != 0) {
        std::istringstream ls(line);
        ls >> word;
        if (word == "DIMENSIONS") ls >> dims[0] >> dims[1] >> dims[2];
      }
      double v;
      while (f >> v) back.push_back(v);
    }
    ok = ok && dims[0] == 129 && dims[1] == 129 && dims[2] == 1 && back == out.grid;
    const double e = ok ? ErrInf(in.domain, back, [](double x, double y, double) {
      return std::sin(M_PI * x) * std::sin(M_PI * y);
    }) : -1.0;
    const bool pass = ok && std::abs(e - 5.020092e-05) < 1e-10;
    std::cout << (pass ? "[OK] " : "[FAIL] ") << "129^2 Poisson CG on HIP, .vtk round trip (" << path
              << ") MMS Linf=" << e << " (closed form 5.020092e-05) " << out.error << err << "\n";
    if (!pass) ++g_fail;
  }
  // Batch entry (the batch runner's replacement, src/main.cpp:2071-2097): uneven 2D MG-PCG
  // inputs with shared options go through the lock-step engine together, a GMRES input on its
  // own; each answer matches the input solved alone to the solve tolerance.
  {
    ++g_total;
    std::vector<SolveInput> ins;
    for (int n : {33, 65, 17, 129}) {
      SolveInput in;
      in.domain = Unit(n);
      in.pde.a = -1.0; in.pde.b = -1.0; in.pde.e = 0.1 * n;
      in.pde.rhs_latex = "-2*\\pi^2*\\sin(\\pi*x)*\\sin(\\pi*y)";
      in.bc = DirichletAll(Linear(0.0, 0.0, 0.0));
      Configure(&in, SolveMethod::MultigridVcycle, 100, 1e-10);
      ins.push_back(in);
    }
    Configure(&ins.back(), SolveMethod::GMRES, 4000, 1e-9);
    HipSolverOptions ho;
    ho.mg_pcg = true;
    ho.relative_tol = true;
    HipSetSolverOptions(ho);
    const std::vector<SolveOutput> outs = SolvePDEBatchHip(ins);
    double worst = 0.0;
    bool ok = outs.size() == ins.size();
    for (size_t i = 0; ok && i < ins.size(); ++i) {
      const SolveOutput one = SolvePDEHip(ins[i]);
      ok = outs[i].error.empty() && one.error.empty() && outs[i].grid.size() == one.grid.size();
      double num = 0.0, den = 0.0;
      for (size_t q = 0; ok && q < one.grid.size(); ++q) {
        num += (outs[i].grid[q] - one.grid[q]) * (outs[i].grid[q] - one.grid[q]);
        den += one.grid[q] * one.grid[q];
      }
      if (ok) worst = std::max(worst, std::sqrt(num / std::max(den, 1e-300)));
      if (!ok) std::cout << "  batch input " << i << ": " << outs[i].error << one.error << "\n";
    }
    HipSetSolverOptions(HipSolverOptions());
    const bool pass = ok && worst < 1e-7;
    std::cout << (pass ? "[OK] " : "[FAIL] ") << "batch entry (3 MG-PCG in lock step + 1 GMRES) "
              << "max rel diff vs single solves " << worst << "\n";
    if (!pass) ++g_fail;
  }
  // Restart equivalence (tests/checkpoint_equivalence.cpp): a series restarted from frame 2's
  // grid (and velocity, for u_tt) at t_start = that frame's time reproduces the remaining frames
  // bit for bit.  Heat with RK4 and Neumann/Robin faces; wave (u_tt) with forward Euler.
  for (int wave = 0; wave < 2; ++wave) {
    ++g_total;
    const int n = 17, total = 7, ck = 2;
    SolveInput in;
    in.domain = Unit(n);
    in.pde.a = -1.0; in.pde.b = -1.0;
    if (wave) in.pde.utt = 1.0; else in.pde.ut = 1.0;
    in.pde.f = 0.5;
    in.bc = DirichletAll(Linear(0.1, 0.2, 0.0));
    in.bc.left.kind = BCKind::Neumann;
    in.bc.left.value = Linear(0.3, 0.0, 0.0);
    in.bc.right.kind = BCKind::Robin;
    in.bc.right.alpha = Linear(1.0, 0.0, 0.0);
    in.bc.right.beta = Linear(0.5, 0.0, 0.0);
    in.bc.right.gamma = Linear(0.2, 0.0, 0.0);
    in.initial_grid.resize(n * n);
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i)
        in.initial_grid[j * n + i] = std::sin(M_PI * i / (n - 1.0)) * std::cos(1.3 * j / (n - 1.0));
    // an initial grid for u_tt needs a velocity as well (solver.cpp:1111)
    if (wave) in.initial_velocity.assign(in.initial_grid.begin(), in.initial_grid.end());
    in.time.enabled = true;
    in.time.t_start = 0.25;
    in.time.dt = wave ? 0.2 / (n - 1.0) : 0.2 / ((n - 1.0) * (n - 1.0));
    in.time.frames = total;
    in.time.integrator = wave ? TimeIntegrator::ForwardEuler : TimeIntegrator::RK4;
    std::vector<std::vector<double>> frames;
    std::vector<double> ck_vel;
    double ck_t = 0.0;
    SolveOutput base = SolvePDETimeSeriesHip(in, [&](int fr, double t, const std::vector<double>& g,
                                                     const std::vector<double>* v) {
      frames.push_back(g);
      if (fr == ck) {
        ck_t = t;
        if (v) ck_vel = *v;
      }
      return true;
    });
    SolveInput re = in;
    re.time.t_start = ck_t;
    re.time.frames = total - ck;
    re.initial_grid = frames.size() > ck ? frames[ck] : std::vector<double>();
    re.initial_velocity = ck_vel;
    std::vector<std::vector<double>> again;
    SolveOutput out = SolvePDETimeSeriesHip(re, [&](int, double, const std::vector<double>& g,
                                                    const std::vector<double>*) {
      again.push_back(g);
      return true;
    });
    bool same = base.error.empty() && out.error.empty() && (int)frames.size() == total &&
                (int)again.size() == total - ck && (!wave || ck_vel.size() == (size_t)n * n);
    double moved = 0.0;
    for (size_t q = 0; same && q < again.size(); ++q) {
      same = again[q] == frames[ck + q];
      for (size_t i = 0; i < again[q].size(); ++i) moved = std::max(moved, std::abs(again[q][i] - frames[0][i]));
    }
    const bool pass = same && moved > 1e-6;   // the state does evolve
    std::cout << (pass ? "[OK] " : "[FAIL] ") << "restart equivalence (" << (wave ? "u_tt, FE" : "u_t, RK4")
              << ") frames " << frames.size() << "/" << again.size() << " " << base.error << out.error << "\n";
    if (!pass) ++g_fail;
  }
  std::cout << (g_total - g_fail) << "/" << g_total << " passed\n";
  return g_fail == 0 ? 0 : 1;
}
