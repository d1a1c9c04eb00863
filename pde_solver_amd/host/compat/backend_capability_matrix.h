// backend_capability_matrix.h — per-backend feature gates of the reference
// (include/backend_capability_matrix.h:12-47), used with the HIP row.
#ifndef BACKEND_CAPABILITY_MATRIX_H
#define BACKEND_CAPABILITY_MATRIX_H

#include <algorithm>
#include <string>
#include <vector>

#include "backend.h"
#include "pde_types.h"

struct BackendCapabilities {
  std::vector<SolveMethod> supported_methods;
  bool supports_3d = false;
  bool supports_spatial_rhs = false;
  bool supports_nonlinear = false;
  bool supports_nonlinear_derivatives = false;
  bool supports_conservation_fv = false;
  bool supports_integrals = false;
  bool supports_shapes = false;
  bool supports_time_dependent = false;
  bool supports_piecewise_bc = false;
  bool supports_coupled_fields = false;
  bool supports_explicit_coupling = false;
  bool supports_implicit_coupling = false;
  bool SupportsMethod(SolveMethod m) const {
    return std::find(supported_methods.begin(), supported_methods.end(), m) != supported_methods.end();
  }
};

BackendCapabilities GetBackendCapabilities(BackendKind kind);
bool BackendSupportsMethod(BackendKind kind, SolveMethod method);
bool BackendSupportsInput(BackendKind kind, const SolveInput& input, std::string* reason);
bool BackendSupportsMethodForInput(BackendKind kind, SolveMethod method, const SolveInput& input,
                                   std::string* reason);

#endif
