// integration_driver.cpp — links the reference's backend registry files with the integration
// patch applied (backend_detect, solver_tokens, backend_capability_matrix, backend_solve, the
// GUI name map) against the HIP plug-in built with the reference's own headers.  Runs on a
// machine without a GPU: the HIP row, tokens and gates are checked, a HIP request reports the
// backend unavailable, and an automatic selection falls back to the reference CPU solver.
#include <cmath>
#include <iostream>
#include <string>

#include "backend.h"
#include "backend_capability_matrix.h"
#include "hip_solver.h"
#include "solver_tokens.h"

static int g_fail = 0;
static void check(bool ok, const std::string& what) {
  std::cout << (ok ? "[OK] " : "[FAIL] ") << what << "\n";
  if (!ok) ++g_fail;
}

int main() {
  check(ParseBackendKind("hip") == BackendKind::HIP, "ParseBackendKind(\"hip\")");
  check(BackendKindName(BackendKind::HIP) == "HIP", "BackendKindName(HIP)");
  check(IsBackendToken("hip") && IsBackendToken("HIP"), "IsBackendToken(\"hip\")");
  check(BackendToken(BackendKind::HIP) == "hip", "BackendToken(HIP)");
  const BackendCapabilities caps = GetBackendCapabilities(BackendKind::HIP);
  check(caps.supports_3d && caps.supports_time_dependent && !caps.supports_shapes &&
            caps.SupportsMethod(SolveMethod::MultigridVcycle),
        "HIP capability row");
  bool listed = false;
  for (const BackendStatus& s : DetectBackends()) listed |= s.kind == BackendKind::HIP;
  check(listed, "DetectBackends lists HIP");

  // 3D Robin-faced Krylov and multigrid: HIP accepts, CUDA keeps the reference's gates
  SolveInput in;
  in.domain.nx = in.domain.ny = in.domain.nz = 17;
  in.pde.a = in.pde.b = in.pde.az = -1.0;
  in.bc.front.kind = BCKind::Robin;
  std::string why;
  check(BackendSupportsMethodForInput(BackendKind::HIP, SolveMethod::CG, in, &why), "HIP: 3D Robin CG allowed");
  check(BackendSupportsMethodForInput(BackendKind::HIP, SolveMethod::MultigridVcycle, in, &why),
        "HIP: 3D Robin multigrid allowed");
  check(!BackendSupportsMethodForInput(BackendKind::CUDA, SolveMethod::CG, in, &why), "CUDA: 3D CG rejected");
  in.pde.dz = 0.5;
  check(!BackendSupportsMethodForInput(BackendKind::HIP, SolveMethod::CG, in, &why), "HIP: z convection rejects CG");
  in.pde.dz = 0.0;
  in.domain.nz = 16;
  check(!BackendSupportsMethodForInput(BackendKind::HIP, SolveMethod::MultigridVcycle, in, &why),
        "HIP: even nz rejects multigrid");
  in.pde.a_latex = "1+x";
  in.domain.nz = 17;
  check(BackendSupportsMethodForInput(BackendKind::HIP, SolveMethod::BiCGStab, in, &why),
        "HIP: variable-coefficient BiCGStab allowed");
  check(!BackendSupportsMethodForInput(BackendKind::HIP, SolveMethod::MultigridVcycle, in, &why),
        "HIP: variable-coefficient multigrid rejected");

  // a small 2D Poisson solve: HIP requested but unavailable here; Auto falls back to the CPU
  SolveInput p;
  p.domain.nx = p.domain.ny = 17;
  p.pde.a = p.pde.b = -1.0;
  p.pde.f = 1.0;
  p.solver.method = SolveMethod::CG;
  p.solver.tol = 1e-10;
  std::string note;
  const bool hip_up = HipIsAvailable(&note);
  BackendKind sel = BackendKind::Auto;
  std::string sel_note;
  SolveOutput out = SolveWithBackend(p, BackendKind::HIP, &sel, &sel_note);
  if (hip_up) check(out.error.empty() && sel == BackendKind::HIP, "HIP solve through SolveWithBackend");
  else check(out.error.find("unavailable") != std::string::npos, "HIP request without a device: " + out.error);
  out = SolveWithBackend(p, BackendKind::Auto, &sel, &sel_note);
  check(out.error.empty() && out.grid.size() == 17u * 17u &&
            (sel == BackendKind::CPU || (hip_up && sel == BackendKind::HIP)),
        "Auto selection solves (" + BackendKindName(sel) + ")");
  std::cout << (g_fail ? "FAILED\n" : "ALL OK\n");
  return g_fail ? 1 : 0;
}
