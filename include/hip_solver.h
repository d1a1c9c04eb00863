/*
 * hip_solver.h — C++ plug-in entry points of the MI355X HIP backend, shaped like the
 * reference's CUDA plug-in (include/cuda_solver.h:8-9) and Metal time-series entry
 * (src/backends/metal/MetalSolve.h:21-23).  Implemented in pde_solver_amd/host/hip_solver.cpp
 * on top of the C ABI in pdehip.h.
 */
#ifndef HIP_SOLVER_H
#define HIP_SOLVER_H

#include <string>

#include "pde_types.h"
#include "progress.h"
#include "solver.h"

// Replaces CudaIsAvailable (include/cuda_solver.h:8): true when a gfx9 (CDNA) device is usable.
bool HipIsAvailable(std::string* note);

// Replaces SolvePDECuda (include/cuda_solver.h:9): steady solve on device 0.  Errors are
// reported through SolveOutput.error with the reference's messages where it has them.
SolveOutput SolvePDEHip(const SolveInput& input);

// Replaces SolvePDETimeSeriesMetal (src/backends/metal/MetalSolve.h:21-23) and the CPU
// SolvePDETimeSeries (include/solver.h:13-14) for 2D explicit time stepping.
SolveOutput SolvePDETimeSeriesHip(const SolveInput& input, const FrameCallback& on_frame,
                                  const ProgressCallback& progress = ProgressCallback());

#endif
