// backend_hip.cpp — backend selection and capability gates with the HIP backend registered.
//
// Restates the reference's selection logic (src/backend_detect.cpp:25-90,
// src/backend_solve.cpp:53-211, src/solver_tokens.cpp:16-35) and capability matrix
// (src/backend_capability_matrix.cpp:68-336) with a HIP entry, so that the same functions,
// patched with a HIP case, slot into the reference tree.  In a standalone build the CPU backend
// (SolvePDE) is a weak reference: when it is not linked, CPU is reported unavailable and no
// request is routed to it.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <string>
#include <vector>

#include "backend.h"
#include "backend_capability_matrix.h"
#include "hip_solver.h"
#include "solver.h"

extern SolveOutput SolvePDE(const SolveInput& input, const ProgressCallback& progress)
    __attribute__((weak));

namespace {

BackendStatus MakeStatus(BackendKind kind, const std::string& name, bool available,
                         const std::string& note) {
  BackendStatus s;
  s.kind = kind;
  s.name = name;
  s.available = available;
  s.note = note;
  return s;
}

const BackendStatus* FindStatus(const std::vector<BackendStatus>& v, BackendKind k) {
  for (const auto& s : v)
    if (s.kind == k) return &s;
  return nullptr;
}

bool IsDirichlet(const BoundaryCondition& bc) { return bc.kind == BCKind::Dirichlet; }

bool AllDirichlet(const SolveInput& in) {
  auto side = [](const BoundaryCondition& bc, const std::vector<PiecewiseBoundaryCondition>& pw) {
    if (!pw.empty()) {
      for (const auto& p : pw) {
        if (!IsDirichlet(p.default_bc)) return false;
        for (const auto& seg : p.segments)
          if (!IsDirichlet(seg.second)) return false;
      }
      return true;
    }
    return IsDirichlet(bc);
  };
  return side(in.bc.left, in.bc.left_piecewise) && side(in.bc.right, in.bc.right_piecewise) &&
         side(in.bc.bottom, in.bc.bottom_piecewise) && side(in.bc.top, in.bc.top_piecewise) &&
         side(in.bc.front, in.bc.front_piecewise) && side(in.bc.back, in.bc.back_piecewise);
}

bool HasPiecewiseBC(const SolveInput& in) {
  return !in.bc.left_piecewise.empty() || !in.bc.right_piecewise.empty() ||
         !in.bc.bottom_piecewise.empty() || !in.bc.top_piecewise.empty() ||
         !in.bc.front_piecewise.empty() || !in.bc.back_piecewise.empty();
}

bool IsLinear(const SolveInput& in) { return in.nonlinear.empty() && in.nonlinear_derivatives.empty(); }

std::string Lower(std::string s) {
  for (char& c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
  return s;
}

}  // namespace

BackendKind ParseBackendKind(const std::string& text) {
  if (text == "cpu") return BackendKind::CPU;
  if (text == "cuda") return BackendKind::CUDA;
  if (text == "metal") return BackendKind::Metal;
  if (text == "tpu") return BackendKind::TPU;
  if (text == "hip") return BackendKind::HIP;
  return BackendKind::Auto;
}

std::string BackendKindName(BackendKind kind) {
  switch (kind) {
    case BackendKind::CPU: return "CPU";
    case BackendKind::CUDA: return "CUDA";
    case BackendKind::Metal: return "Metal";
    case BackendKind::TPU: return "TPU";
    case BackendKind::HIP: return "HIP";
    default: return "Auto";
  }
}

bool IsBackendToken(const std::string& text) {
  const std::string t = Lower(text);
  return t == "auto" || t == "cpu" || t == "cuda" || t == "metal" || t == "tpu" || t == "hip";
}

std::string BackendToken(BackendKind kind) {
  switch (kind) {
    case BackendKind::CPU: return "cpu";
    case BackendKind::CUDA: return "cuda";
    case BackendKind::Metal: return "metal";
    case BackendKind::TPU: return "tpu";
    case BackendKind::HIP: return "hip";
    default: return "auto";
  }
}

std::vector<BackendStatus> DetectBackends() {
  std::vector<BackendStatus> v;
  v.push_back(SolvePDE ? MakeStatus(BackendKind::CPU, "CPU", true, "available")
                       : MakeStatus(BackendKind::CPU, "CPU", false, "CPU backend not linked in this build"));
  v.push_back(MakeStatus(BackendKind::CUDA, "CUDA", false, "not compiled with CUDA support"));
  v.push_back(MakeStatus(BackendKind::Metal, "Metal", false, "not compiled with Metal support"));
  v.push_back(MakeStatus(BackendKind::TPU, "TPU", false, "not compiled with TPU support"));
  std::string note;
  const bool hip = HipIsAvailable(&note);
  v.push_back(MakeStatus(BackendKind::HIP, "HIP", hip, note.empty() ? "compiled with HIP support" : note));
  return v;
}

BackendCapabilities GetBackendCapabilities(BackendKind kind) {
  BackendCapabilities caps;
  const std::vector<SolveMethod> all = {SolveMethod::Jacobi, SolveMethod::GaussSeidel, SolveMethod::SOR,
                                        SolveMethod::CG, SolveMethod::BiCGStab, SolveMethod::GMRES,
                                        SolveMethod::MultigridVcycle};
  switch (kind) {
    case BackendKind::CPU:
      caps.supported_methods = all;
      caps.supports_3d = caps.supports_spatial_rhs = caps.supports_nonlinear = true;
      caps.supports_nonlinear_derivatives = caps.supports_conservation_fv = true;
      caps.supports_integrals = caps.supports_shapes = caps.supports_time_dependent = true;
      caps.supports_piecewise_bc = caps.supports_coupled_fields = true;
      caps.supports_explicit_coupling = caps.supports_implicit_coupling = true;
      return caps;
    case BackendKind::CUDA:
    case BackendKind::Metal:
      caps.supported_methods = all;
      caps.supports_spatial_rhs = true;
      caps.supports_nonlinear = kind == BackendKind::CUDA;
      return caps;
    case BackendKind::TPU:
      caps.supported_methods = {SolveMethod::Jacobi};
      return caps;
    case BackendKind::HIP:
      // 2D and 3D; Krylov and multigrid with Neumann/Robin rows and variable coefficients;
      // explicit 2D time stepping.  No shapes, integrals, piecewise BCs or coupled fields.
      caps.supported_methods = all;
      caps.supports_3d = true;
      caps.supports_spatial_rhs = true;
      caps.supports_nonlinear = true;
      caps.supports_time_dependent = true;
      return caps;
    default:
      return caps;
  }
}

bool BackendSupportsMethod(BackendKind kind, SolveMethod method) {
  return GetBackendCapabilities(kind).SupportsMethod(method);
}

bool BackendSupportsInput(BackendKind kind, const SolveInput& in, std::string* reason) {
  const BackendCapabilities caps = GetBackendCapabilities(kind);
  auto no = [&](const char* r) { if (reason) *reason = r; return false; };
  if (in.domain.nz > 1 && !caps.supports_3d) return no("3D domains are not supported");
  if (!in.pde.rhs_latex.empty() && !caps.supports_spatial_rhs) return no("spatial RHS expressions are not supported");
  if (HasVariableCoefficients(in.pde) && kind == BackendKind::TPU)
    return no("variable coefficients are not supported on this backend");
  if ((HasMixedDerivatives(in.pde) || HasHigherOrderDerivatives(in.pde)) && kind != BackendKind::CPU)
    return no("mixed or higher-order derivatives are not supported on this backend");
  if (in.domain.coord_system != CoordinateSystem::Cartesian && kind == BackendKind::HIP)
    return no("the HIP backend supports Cartesian coordinates only");
  if (!IsLinear(in) && !caps.supports_nonlinear) return no("nonlinear terms are not supported");
  if (!in.integrals.empty() && !caps.supports_integrals) return no("integral terms are not supported");
  if (HasImplicitShape(in) && !caps.supports_shapes) return no("implicit domain shapes are not supported");
  const bool td = in.time.enabled || std::abs(in.pde.ut) > 1e-12 || std::abs(in.pde.utt) > 1e-12;
  if (td && !caps.supports_time_dependent) return no("time-dependent PDEs are not supported");
  if (in.domain.nz > 1 && td) return no("time-dependent 3D solves are not supported");
  if (HasPiecewiseBC(in) && !caps.supports_piecewise_bc) return no("piecewise boundary conditions are not supported");
  if (!in.nonlinear_derivatives.empty() && !caps.supports_nonlinear_derivatives)
    return no("nonlinear derivative terms are not supported on this backend");
  return true;
}

bool BackendSupportsMethodForInput(BackendKind kind, SolveMethod method, const SolveInput& in,
                                   std::string* reason) {
  auto no = [&](const char* r) { if (reason) *reason = r; return false; };
  if (!BackendSupportsMethod(kind, method)) return no("selected method is not supported on this backend");
  if (!BackendSupportsInput(kind, in, reason)) return false;
  const bool krylov = method == SolveMethod::CG || method == SolveMethod::BiCGStab || method == SolveMethod::GMRES;
  const bool mg = method == SolveMethod::MultigridVcycle;
  const bool relax = method == SolveMethod::Jacobi || method == SolveMethod::GaussSeidel || method == SolveMethod::SOR;
  const bool var = HasVariableCoefficients(in.pde);
  const bool hip = kind == BackendKind::HIP;
  if ((krylov || mg) && !IsLinear(in)) return no("linear PDE required for this method");
  if ((krylov || mg) && HasImplicitShape(in)) return no("selected method does not support implicit domain shapes");
  // HIP lifts the 2D-only and Dirichlet-only gates of Krylov and multigrid (folded
  // Neumann/Robin rows); the other backends keep them (backend_capability_matrix.cpp:273-284).
  if (!hip) {
    if ((krylov || mg) && in.domain.nz > 1) return no("selected method supports 2D domains only");
    if (krylov && !AllDirichlet(in)) return no("Krylov methods require Dirichlet boundaries on all sides");
    if (mg && !AllDirichlet(in)) return no("Multigrid requires Dirichlet boundaries on all sides");
  }
  if (var && kind != BackendKind::CPU) {
    if (kind == BackendKind::CUDA || kind == BackendKind::Metal) {
      if (!relax) return no("variable coefficients are only supported for relaxation methods on GPU");
      if (!in.pde.az_latex.empty() || !in.pde.dz_latex.empty())
        return no("variable z coefficients are not supported on this backend");
    } else if (hip) {
      if (mg) return no("multigrid requires constant coefficients");
    } else {
      return no("variable coefficients are not supported on this backend");
    }
  }
  if (method == SolveMethod::CG &&
      (var || std::abs(in.pde.c) > 1e-12 || std::abs(in.pde.d) > 1e-12 || std::abs(in.pde.dz) > 1e-12 ||
       HasMixedDerivatives(in.pde) || HasHigherOrderDerivatives(in.pde)))
    return no("CG requires symmetric operators with constant coefficients (no convection/mixed/higher-order terms)");
  if (mg) {
    if (var) return no("multigrid requires constant coefficients");
    if (!in.integrals.empty()) return no("multigrid does not support integral terms");
    if (std::abs(in.pde.c) > 1e-12 || std::abs(in.pde.d) > 1e-12 || std::abs(in.pde.dz) > 1e-12 ||
        HasMixedDerivatives(in.pde) || HasHigherOrderDerivatives(in.pde))
      return no("multigrid requires no convection, mixed, or higher-order derivatives");
    if (in.domain.nx < 5 || in.domain.ny < 5 || (in.domain.nz > 1 && in.domain.nz < 5))
      return no("multigrid requires grid dimensions of at least 5");
    if (in.domain.nx % 2 == 0 || in.domain.ny % 2 == 0 || (in.domain.nz > 1 && in.domain.nz % 2 == 0))
      return no("multigrid requires odd grid dimensions");
  }
  return true;
}

SolveOutput SolveWithBackend(const SolveInput& input, BackendKind requested, BackendKind* selected,
                             std::string* selection_note, const ProgressCallback& progress) {
  if (std::abs(input.pde.ut) > 1e-12 || std::abs(input.pde.utt) > 1e-12)
    return {"time-dependent PDEs require the time-series solver", {}};
  const auto statuses = DetectBackends();
  std::string extra;
  BackendKind choice = requested;
  auto supports = [&](BackendKind k, std::string* r) {
    return BackendSupportsMethodForInput(k, input.solver.method, input, r);
  };
  if (choice == BackendKind::Auto) {
    // GUI callers (progress set) keep the reference's CPU-first order; others try HIP first.
    const bool interactive = static_cast<bool>(progress);
    const BackendKind inter[] = {BackendKind::CPU, BackendKind::HIP, BackendKind::CUDA, BackendKind::Metal, BackendKind::TPU};
    const BackendKind dflt[] = {BackendKind::HIP, BackendKind::CUDA, BackendKind::Metal, BackendKind::TPU, BackendKind::CPU};
    const BackendKind* order = interactive ? inter : dflt;
    std::string last;
    for (int i = 0; i < 5; ++i) {
      const BackendStatus* st = FindStatus(statuses, order[i]);
      if (!st || !st->available) continue;
      std::string r;
      if (!supports(order[i], &r)) { last = r; continue; }
      choice = order[i];
      if (interactive && choice == BackendKind::CPU) extra = "auto: preferring CPU for GUI responsiveness";
      break;
    }
    if (choice == BackendKind::Auto)
      return {"no available backend supports the selected solve: " +
                  (last.empty() ? std::string("no compatible backend available") : last), {}};
  }
  std::string reason;
  if (!supports(choice, &reason)) {
    if (choice == BackendKind::CPU) return {"requested solve not supported on CPU: " + reason, {}};
    if (!extra.empty()) extra += "; ";
    extra += reason.empty() ? "requested backend unsupported" : reason;
    choice = BackendKind::CPU;
    reason.clear();
    if (!supports(choice, &reason)) return {"requested solve not supported on any backend: " + reason, {}};
  }
  const BackendStatus* st = FindStatus(statuses, choice);
  if (!st || !st->available) return {"requested backend unavailable: " + BackendKindName(choice), {}};
  if (selected) *selected = choice;
  if (selection_note) {
    *selection_note = st->note;
    if (!extra.empty()) *selection_note = selection_note->empty() ? extra : *selection_note + "; " + extra;
  }
  SolveOutput out;
  switch (choice) {
    case BackendKind::HIP:
      if (progress) progress("solve", 0.0);
      out = SolvePDEHip(input);   // residual norms come from the device
      if (progress) progress("solve", 1.0);
      break;
    case BackendKind::CPU:
      out = SolvePDE(input, progress);
      break;
    default:
      out = {BackendKindName(choice) + " backend not compiled", {}};
  }
  return out;
}
