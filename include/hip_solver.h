/*
 * hip_solver.h — C++ plug-in entry points of the MI355X HIP backend, shaped like the
 * reference's CUDA plug-in (include/cuda_solver.h:8-9) and Metal time-series entry
 * (src/backends/metal/MetalSolve.h:21-23).  Implemented in pde_solver_amd/host/hip_solver.cpp
 * on top of the C ABI in pdehip.h.
 */
#ifndef HIP_SOLVER_H
#define HIP_SOLVER_H

#include <string>
#include <vector>

#include "pde_types.h"
#include "progress.h"
#include "solver.h"

// Replaces CudaIsAvailable (include/cuda_solver.h:8): true when a gfx9 (CDNA) device is usable.
bool HipIsAvailable(std::string* note);

// Backend options with no SolverConfig member in the reference (include/pde_types.h:282-307),
// so that this file and hip_solver.cpp compile unchanged against the reference's own header.
// Process-wide; each solve reads them once at its start.  Defaults are the reference's
// semantics: FP64, absolute tolerance, the CPU's axis Robin convention, no preconditioner,
// MultigridVcycle as iterated V-cycles.
struct HipSolverOptions {
  bool mixed_precision = false;  // FP32 (FP16-stored coarse levels) inner iterations, FP64 refinement
  bool mg_pcg = false;           // MultigridVcycle runs MG-preconditioned CG
  bool mg_single_cycle = false;  // MultigridVcycle runs exactly one V-cycle (CPU solver.cpp:582)
  bool relative_tol = false;     // SolverConfig.tol relative to ||b||_2 instead of absolute
  bool outward_robin = false;    // alpha u + beta du/dn = gamma with the outward normal
                                 // (CUDA/Metal/residual.cpp) instead of the CPU's axis form
  bool jacobi_precond = false;   // Jacobi right preconditioner for BiCGStab / GMRES
};
void HipSetSolverOptions(const HipSolverOptions& options);
HipSolverOptions HipGetSolverOptions();

// Replaces SolvePDECuda (include/cuda_solver.h:9): steady solve on device 0.  Errors are
// reported through SolveOutput.error with the reference's messages where it has them.
SolveOutput SolvePDEHip(const SolveInput& input);

// For the sequential batch runner (src/main.cpp:2071-2097, one SolveWithBackend per entry):
// solves every input in one call and returns one SolveOutput per input, in order.  Runs of
// consecutive inputs with the same solver options go to pdehip_solve_batch together, so sets
// the lock-step engine takes (2D constant-coefficient MG-PCG or BiCGStab) are solved as one
// batch; any other input is solved like SolvePDEHip.
std::vector<SolveOutput> SolvePDEBatchHip(const std::vector<SolveInput>& inputs);

// Legacy-VTK STRUCTURED_POINTS file of a solution grid (the layout the reference's
// WriteVtkStructuredPoints writes, src/vtk_writer.cpp:28-60: DIMENSIONS / ORIGIN / SPACING,
// point data x fastest) with `digits` significant digits per value.  17 digits round-trip FP64
// exactly; the reference's writer keeps 6 decimals of a float.  Returns false and sets *error
// when the grid does not match the domain or the file cannot be written.
bool HipWriteVtk(const std::string& path, const Domain& domain, const std::vector<double>& grid,
                 int digits, std::string* error);

// Replaces SolvePDETimeSeriesMetal (src/backends/metal/MetalSolve.h:21-23) and the CPU
// SolvePDETimeSeries (include/solver.h:13-14) for 2D explicit time stepping.
SolveOutput SolvePDETimeSeriesHip(const SolveInput& input, const FrameCallback& on_frame,
                                  const ProgressCallback& progress = ProgressCallback());

#endif
