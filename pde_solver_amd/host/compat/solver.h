// solver.h — FrameCallback and the CPU entry points as declared by the reference
// (include/solver.h:9-14).  In a standalone build of this repository the CPU backend is not
// linked (SolvePDE / SolvePDETimeSeries are weak references resolved in-tree).
#ifndef SOLVER_H
#define SOLVER_H

#include <functional>
#include <vector>

#include "pde_types.h"
#include "progress.h"

using FrameCallback = std::function<bool(int frame, double time, const std::vector<double>& grid,
                                         const std::vector<double>* velocity)>;

SolveOutput SolvePDE(const SolveInput& input, const ProgressCallback& progress = ProgressCallback());
SolveOutput SolvePDETimeSeries(const SolveInput& input, const FrameCallback& on_frame,
                               const ProgressCallback& progress = ProgressCallback());

#endif
