/*
 * pdehip.h — C ABI of the MI355X (gfx950) HIP solver backend for pde_solver.
 *
 * Plain C: no C++ types, no torch types, caller-owned host buffers.  Every
 * array is a full vertex-centred grid including boundary nodes, row-major with
 * x fastest: idx = (k*ny + j)*nx + i  (reference: src/backends/cpu/cpu_utils.h:13-19).
 *
 * Sign convention follows the reference (src/backends/cpu/solvers/relaxation.cpp:281-283):
 *     a u_xx + b u_yy + az u_zz + c u_x + d u_y + dz u_z + e u + f = 0
 * so the assembled system is A u = b with b = -f on interior rows and the
 * boundary datum on boundary rows (reference src/backends/cpu/solver.cpp:222-228).
 *
 * Entry points and the reference interface each one replaces:
 *   pdehip_is_available      <- CudaIsAvailable          include/cuda_solver.h:8
 *   pdehip_solve             <- SolvePDECuda              include/cuda_solver.h:9,
 *                               (CudaSolve.cu:31-65 dispatch by SolveMethod)
 *   pdehip_time_series       <- SolvePDETimeSeriesMetal   src/backends/metal/MetalSolve.h:21-23
 *                               and SolvePDETimeSeries    include/solver.h:13-14
 *   pdehip_residual_norms    <- ComputeResidualNorms      include/residual.h / src/residual.cpp:153
 *   pdehip_apply             <- LinearOperator2D::Apply   src/backends/cpu/cpu_utils.cpp:324
 *                               and CUDA AxKernel          src/backends/cuda/solvers/krylov.cu:12-35
 *   pdehip_plan_*            <- no reference counterpart: resident operator/hierarchy
 *                               for repeated solves (the reference mallocs and frees
 *                               per solve, src/backends/cuda/solvers/relaxation.cu:272-316)
 *   pdehip_solve_batch       <- the sequential batch runner src/main.cpp:2071-2097
 *   pdehip_server_*          <- no reference counterpart (request loop, continuous batching)
 *   pdehip_residual_norms_noalloc <- no reference counterpart (allocation-free variant of
 *                               pdehip_residual_norms for callers that cannot allocate)
 *
 * Status codes map onto the reference's SolveOutput.error idiom
 * (src/backends/cuda/CudaSolve.cu:33-45); pdehip_last_error() returns the
 * message the C++ wrapper copies into SolveOutput.error.
 */
#ifndef PDEHIP_H
#define PDEHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PDEHIP_ABI_VERSION 2

/* ---------------------------------------------------------------- status */
enum pdehip_status {
  PDEHIP_OK = 0,
  PDEHIP_ERR_INVALID = 1,        /* malformed problem / options */
  PDEHIP_ERR_UNSUPPORTED = 2,    /* feature outside this backend (maps to SolveStatus::UnsupportedFeature) */
  PDEHIP_ERR_NOT_CONVERGED = 3,  /* "failed to converge" (reference solver.cpp:544-546) */
  PDEHIP_ERR_BREAKDOWN = 4,      /* Krylov breakdown |denom| < 1e-30 (reference krylov.cpp:46,116,132,151) */
  PDEHIP_ERR_DEGENERATE = 5,     /* "degenerate PDE center coefficient" (relaxation.cpp:116-118) */
  PDEHIP_ERR_DEVICE = 6,         /* HIP runtime error */
  PDEHIP_ERR_OOM = 7,            /* allocation beyond device memory or the HBM budget */
  PDEHIP_ERR_CANCELLED = 8,      /* cancel flag observed (pde_types.h:402) */
  PDEHIP_ERR_ADMISSION = 9,      /* request footprint larger than the whole HBM budget */
  PDEHIP_PENDING = 10,           /* serving: request not finished yet */
  PDEHIP_ERR_TIMEOUT = 11        /* serving: per-request deadline passed */
};

/* ---------------------------------------------------------------- problem */
enum pdehip_bc_kind { PDEHIP_BC_DIRICHLET = 0, PDEHIP_BC_NEUMANN = 1, PDEHIP_BC_ROBIN = 2 };

/* Face order follows BoundarySet (include/pde_types.h:214-220). */
enum pdehip_face {
  PDEHIP_LEFT = 0, PDEHIP_RIGHT = 1, PDEHIP_BOTTOM = 2,
  PDEHIP_TOP = 3, PDEHIP_FRONT = 4, PDEHIP_BACK = 5
};

/* Robin/Neumann closure convention (one-sided first-order difference in both).
 *  AXIS    : alpha*u - beta*du/dx_axis = gamma and du/dx_axis = g on both faces of an axis
 *            (reference CPU backend, src/backends/cpu/cpu_utils.cpp:645-673,790-895).
 *  OUTWARD : alpha*u + beta*du/dn = gamma and du/dn = g with the outward normal n.
 *            Its Robin closure is the CUDA/Metal/residual one (denominator alpha+beta/h on every
 *            face: src/backends/cuda/utils/cuda_utils.cu:119-121,165-167, src/residual.cpp:339,359);
 *            for homogeneous Neumann (g = 0) the two conventions coincide. */
enum pdehip_bc_convention { PDEHIP_BCCONV_AXIS = 0, PDEHIP_BCCONV_OUTWARD = 1 };

/* BoundaryCondition::Expression (include/pde_types.h:194-200): value = c + cx*x + cy*y + cz*z,
 * or per-node samples on the face when `samples` is non-NULL.  Face node order:
 *   left/right  : k*ny + j      bottom/top : k*nx + i      front/back : j*nx + i            */
typedef struct pdehip_expr {
  double c, cx, cy, cz;
  const double* samples;
} pdehip_expr;

typedef struct pdehip_face_bc {
  int kind;             /* enum pdehip_bc_kind */
  pdehip_expr value;    /* Dirichlet value or Neumann flux g */
  pdehip_expr alpha, beta, gamma;
} pdehip_face_bc;

/* Operator kinds.
 *  CONST     : reference LinearOperator2D constant-coefficient form (cpu_utils.cpp:293-297),
 *              extended to 3D like SolvePDE3D (solver.cpp:723-729).
 *  POINTWISE : reference variable-coefficient form a(x)u_xx + ... (relaxation.cpp:187-201);
 *              coefficient fields sampled per node (coef[] arrays).
 *  FLUX      : conservative a*div(kappa grad u) + e*u, kappa per node, face value = harmonic
 *              mean of the two nodes (no reference counterpart; needed for -div(k grad u)). */
enum pdehip_op_kind { PDEHIP_OP_CONST = 0, PDEHIP_OP_POINTWISE = 1, PDEHIP_OP_FLUX = 2 };

/* First-derivative discretisation: centred (reference cpu_utils.cpp:295) or first-order upwind. */
enum pdehip_convection { PDEHIP_CONV_CENTERED = 0, PDEHIP_CONV_UPWIND = 1 };

/* NonlinearTerm (include/pde_types.h:88-100). */
enum pdehip_nonlinear_kind {
  PDEHIP_NL_POWER = 0, PDEHIP_NL_SIN = 1, PDEHIP_NL_COS = 2, PDEHIP_NL_EXP = 3, PDEHIP_NL_ABS = 4
};
typedef struct pdehip_nonlinear { int kind; double coeff; int power; } pdehip_nonlinear;
#define PDEHIP_MAX_NONLINEAR 4

enum pdehip_coef_index { PDEHIP_COEF_A = 0, PDEHIP_COEF_B, PDEHIP_COEF_AZ, PDEHIP_COEF_C,
                         PDEHIP_COEF_D, PDEHIP_COEF_DZ, PDEHIP_COEF_E, PDEHIP_NCOEF };

typedef struct pdehip_problem {
  int nx, ny, nz;                          /* nz <= 1 means 2D (Domain, pde_types.h:148-159) */
  double xmin, xmax, ymin, ymax, zmin, zmax;
  int op_kind;                             /* enum pdehip_op_kind */
  double a, b, az, c, d, dz, e;            /* PDECoefficients names (pde_types.h:10-18) */
  const double* coef[PDEHIP_NCOEF];        /* POINTWISE fields; NULL entry = constant member */
  const double* kappa;                     /* FLUX conductivity per node */
  int convection;                          /* enum pdehip_convection */
  double f;                                /* constant source term (pde_types.h:18) */
  const double* f_field;                   /* per-node source, overrides f (rhs_latex sampled) */
  pdehip_face_bc bc[6];                    /* enum pdehip_face order */
  int bc_convention;                       /* enum pdehip_bc_convention */
  int n_nonlinear;                         /* relaxation / time series only */
  pdehip_nonlinear nonlinear[PDEHIP_MAX_NONLINEAR];
  double ut, utt;                          /* time derivative coefficients */
  const double* initial;                   /* optional initial grid (checkpoint restart) */
  const double* initial_velocity;          /* optional u_tt velocity */
} pdehip_problem;

/* ---------------------------------------------------------------- options */
/* SolveMethod (include/pde_types.h:230-238) plus MG-preconditioned CG. */
enum pdehip_method {
  PDEHIP_JACOBI = 0, PDEHIP_GAUSS_SEIDEL = 1, PDEHIP_SOR = 2, PDEHIP_CG = 3,
  PDEHIP_BICGSTAB = 4, PDEHIP_GMRES = 5, PDEHIP_MG = 6, PDEHIP_MGPCG = 7
};
enum pdehip_precision { PDEHIP_FP64 = 0, PDEHIP_MIXED = 1 };
enum pdehip_precond { PDEHIP_PC_NONE = 0, PDEHIP_PC_JACOBI = 1, PDEHIP_PC_MG = 2 };
/* ABS: ||r||_2 < tol, the reference rule (krylov.cpp:39,66).  REL: ||r||_2 < tol*||b||_2. */
enum pdehip_tol_mode { PDEHIP_TOL_ABS = 0, PDEHIP_TOL_REL = 1 };
/* How BC rows enter Krylov/MG operators.
 *  REFERENCE : identity boundary rows, interior rows couple to boundary nodes; Dirichlet only
 *              (reference LinearOperator2D::Apply boundary handling, cpu_utils.cpp:368-381,497-512).
 *  FOLDED    : Neumann/Robin closures folded into the adjacent interior row; the operator on
 *              interior nodes is symmetric for symmetric PDEs, so CG/MG-PCG apply. */
enum pdehip_bc_rows { PDEHIP_ROWS_REFERENCE = 0, PDEHIP_ROWS_FOLDED = 1 };

typedef struct pdehip_options {
  int method;             /* enum pdehip_method */
  int max_iter;           /* SolverConfig.max_iter (default 10000) */
  double tol;             /* SolverConfig.tol (default 1e-6) */
  int tol_mode;           /* enum pdehip_tol_mode */
  double sor_omega;       /* default 1.5 */
  int gmres_restart;      /* default 30 */
  int mg_pre, mg_post, mg_coarse_iters, mg_max_levels;   /* defaults 2/2/50/10 */
  int mg_single_cycle;    /* 1: exactly one V-cycle, as the reference CPU 'mg' (solver.cpp:582) */
  int precond;            /* enum pdehip_precond (BiCGStab/GMRES/CG) */
  int precision;          /* enum pdehip_precision */
  int bc_rows;            /* enum pdehip_bc_rows */
  double inner_tol;       /* mixed MG-PCG (single and batched): residual-replacement threshold;
                             the FP32 residual is replaced by the FP64 b - A x once it has
                             fallen by this factor since the last replacement (0 = 1e-2) */
  int check_every;        /* iterations between host convergence checks (0 = default) */
  const volatile int* cancel;  /* optional cancel flag polled between iterations (pde_types.h:402) */
} pdehip_options;

typedef struct pdehip_report {
  int status;             /* enum pdehip_status */
  int iterations;         /* outer iterations (Krylov / relaxation sweeps / V-cycles) */
  int inner_iterations;   /* mixed: total FP32 inner iterations */
  int refinements;        /* mixed: FP64 refinement / residual-replacement steps; -1 when a
                             diverging FP32 BiCGStab was completed by the FP64 iteration */
  int converged;
  double final_metric;    /* the quantity the stop rule tested (||r||_2 or max|du|) */
  double rel_residual;    /* ||b - A u||_2 / ||b||_2 in FP64 (FOLDED/REFERENCE rows) */
  double residual_l2;     /* ComputeResidualNorms definition (residual.cpp:153) */
  double residual_linf;
  double solve_ms;        /* device time of the solve (HIP events) */
} pdehip_report;

/* Defaults matching SolverConfig (include/pde_types.h:282-307). */
void pdehip_default_options(pdehip_options* opt);
void pdehip_default_problem(pdehip_problem* prob);

/* ---------------------------------------------------------------- context */
typedef struct pdehip_ctx pdehip_ctx;
typedef struct pdehip_plan pdehip_plan;

int pdehip_abi_version(void);
/* Returns 1 when a gfx9xx device is usable; fills `note` like CudaIsAvailable (CudaSolve.cu:9-29). */
int pdehip_is_available(char* note, int note_len);
/* hbm_budget_bytes: admission/allocation cap for this context (0 = 80% of free device memory). */
int pdehip_create(int device, size_t hbm_budget_bytes, pdehip_ctx** out);
/* Plans, batches and servers created on ctx must be destroyed before it. */
void pdehip_destroy(pdehip_ctx* ctx);
const char* pdehip_last_error(const pdehip_ctx* ctx);

/* ---------------------------------------------------------------- steady solves */
int pdehip_solve(pdehip_ctx* ctx, const pdehip_problem* prob, const pdehip_options* opt,
                 double* u_out, pdehip_report* report);
/* y = A x with the REFERENCE row convention (LinearOperator2D::Apply; 3D by SolvePDE3D's stencil). */
int pdehip_apply(pdehip_ctx* ctx, const pdehip_problem* prob, const double* x, double* y);
/* ComputeResidualNorms on the device (residual.cpp:153-583, Cartesian, no shapes/integrals). */
int pdehip_residual_norms(pdehip_ctx* ctx, const pdehip_problem* prob, const double* u,
                          double* l2, double* linf);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): pdehip_residual_norms
 * on device-resident inputs.  u_dev and every array member of *prob (f_field, kappa, face
 * samples) are DEVICE pointers; l2 and linf are host outputs.  The call allocates nothing: its
 * partial sums, flags and sampled boundary data live in a static scratch array of the library
 * (2.16 MB per device, held until the process ends).  Calls that overlap on one device share
 * that array and race.  CONST and FLUX operators only (PDEHIP_ERR_UNSUPPORTED for POINTWISE);
 * a 2D grid side above 16385 nodes is rejected (PDEHIP_ERR_INVALID).  Results equal
 * pdehip_residual_norms on the same data. */
int pdehip_residual_norms_noalloc(pdehip_ctx* ctx, const pdehip_problem* prob, const double* u_dev,
                                  double* l2, double* linf);

/* ---------------------------------------------------------------- resident plans */
int pdehip_plan_create(pdehip_ctx* ctx, const pdehip_problem* prob, const pdehip_options* opt,
                       pdehip_plan** out);
/* Replace the source field (host array of nx*ny*nz, or NULL for the constant f). */
int pdehip_plan_set_source(pdehip_plan* plan, const double* f_field);
/* Solve from a zero initial guess; the solution stays resident on the device. */
int pdehip_plan_solve(pdehip_plan* plan, pdehip_report* report);
int pdehip_plan_download(pdehip_plan* plan, double* u_out);
/* Bytes of device memory the plan holds. */
size_t pdehip_plan_bytes(const pdehip_plan* plan);
/* Time `reps` back-to-back launches of one kernel of the plan with HIP events on the plan's
 * stream.  kernel: 0 = fine-level operator apply, 1 = fine-level red-black sweep as two in-place
 * colour launches, 2 = fine-level residual, 3 = stream copy of one fine vector (bandwidth
 * probe), 4 = fine-level red-black sweep as one fused out-of-place launch.
 * Returns the mean milliseconds per sweep/launch and the algorithmic bytes per sweep/launch. */
int pdehip_plan_time_kernel(pdehip_plan* plan, int kernel, int precision, int reps,
                            double* avg_ms, double* algorithmic_bytes);
void pdehip_plan_destroy(pdehip_plan* plan);

/* ---------------------------------------------------------------- launch tuning */
/* Process-wide launch-shape knobs (no counterpart in the reference; used by the autotuner and
 * for A/B timing).  Keys: "smooth.fused" (0/1: multigrid smooths with the fused out-of-place
 * sweep), "smooth.rows" (2 or 4: rows per thread of the 3D sweep tile), "smooth.zchunk" (z
 * planes per 3D workgroup, 4..512, or 0 to choose from occupancy), "smooth.ychunk" (2D rows
 * per wavefront, 16..4096), "stencil.plane" (0/1: 3D apply and residual as plane-march
 * kernels), "smooth.pair" (0/1: 3D FP32/FP16 sweeps with column-pair tiles, no colour
 * divergence), "smooth.ring" (0/1, default 1: 3D FP32/FP16 sweeps by the ring-buffered,
 * 6-step unrolled kernel), "smooth.reg" (0/1, default 1: of the ring sweeps, the one with
 * in-plane neighbours in registers, four rows per wave), "smooth.dot" (0/1, default 1: the last
 * level-0 sweep of a mixed V-cycle also leaves the partials of r.z), "stencil.ring" (0/1,
 * default 1: the same as "smooth.ring" for the 3D FP32/FP16
 * apply and residual), "smooth.box" (0..1025, default 65: coarse 3D levels of mixed-precision
 * V-cycles whose sides are all at most this smooth on LDS boxes, the prolongation folded into
 * the first post-smoothing launch; 0: off), "smooth.box2" (0..1025, default 33: of those, the
 * levels whose sides are all at most this take two sweeps per launch, the others one),
 * "smooth.box_z" (4 or 8, default 4: interior planes per box), "smooth.box_rr" (0..4097,
 * default 65: mixed 3D levels whose sides are all at most this form their residual and its
 * restriction in one launch; 0: off), "smooth.box_small" (0..1025, default 33: box levels whose
 * sides are all at most this use small boxes, 8 x "smooth.box_sxy" x 4 for the sweeps (with
 * 4-plane boxes) and 4 x 4 x 4 coarse nodes for the residual-restriction; 0: off),
 * "smooth.box_sxy" (4 or 8, default 4).  Results do not depend on any of these.
 * "time.fused" (0/1, default 1): RK4 time series take one launch per step (the four stages and
 * the closures, temporally blocked in LDS) instead of one per stage; the same bits either way.
 * "mixed.fp16" (0/1, default 1): mixed-precision plans created afterwards store their coarse
 * multigrid levels in FP16 (1) or FP32 (0); this one changes the preconditioner's rounding.
 * "mixed.replace" (0/1, default 1): mixed MG-PCG on 3D grids replaces its FP32 residual in one
 * launch (fold the FP32 correction into the FP64 solution, FP64 residual, scaled FP32 copy) or
 * in three; the same bits either way.  "mixed.xpay" (0/1, default 1): mixed MG-PCG on 3D grids
 * forms the new search direction inside the operator apply (one launch instead of two); the
 * same bits either way.  "restrict.march" (0/1, default 1): 3D multigrid restrictions march the
 * coarse planes (each fine plane loaded once, weighted in the plane and carried to the next
 * coarse plane) instead of gathering 27 fine values per coarse node; the same bits either way.
 * Not thread-safe against running solves. */
int pdehip_tune_set(const char* key, int value);
int pdehip_tune_get(const char* key, int* value);

/* ---------------------------------------------------------------- time series */
/* TimeIntegrator (include/pde_types.h:310-319); explicit ones only. */
enum pdehip_integrator {
  PDEHIP_TI_FORWARD_EULER = 0, PDEHIP_TI_RK2 = 1, PDEHIP_TI_RK4 = 2,
  PDEHIP_TI_SSPRK2 = 3, PDEHIP_TI_SSPRK3 = 4
};
typedef struct pdehip_time_config {
  int integrator;         /* enum pdehip_integrator */
  double t_start, dt;
  int frames;             /* TimeConfig.frames: frame 0 is the initial state */
  int precision;          /* PDEHIP_FP64 or PDEHIP_MIXED (= FP32 state) */
  int buffer_frames;      /* frames held on the device before one batched copy (TimeConfig.buffer_mb) */
} pdehip_time_config;
/* Frame callback (FrameCallback, include/solver.h:9-11); return 0 to stop. */
typedef int (*pdehip_frame_fn)(void* user, int frame, double t, const double* grid,
                               const double* velocity);
int pdehip_time_series(pdehip_ctx* ctx, const pdehip_problem* prob, const pdehip_time_config* tc,
                       pdehip_frame_fn on_frame, void* user, double* u_out, pdehip_report* report);

/* ---------------------------------------------------------------- batches */
/* Solve `count` independent problems.  2D constant-coefficient linear sets solved by MG-PCG
 * (no convection) or BiCGStab (no preconditioner or Jacobi) go through the lock-step batched
 * engine below; other sets are solved one resident plan after the other.  u_out[i] receives
 * problem i.  reports may be NULL. */
typedef struct pdehip_batch_config {
  int precision;          /* PDEHIP_FP64 or PDEHIP_MIXED (pdehip_solve_batch; the resident
                             batch takes the precision from its options) */
  int use_graph;          /* 1: the launch sequence of two PCG (eight BiCGStab) iterations is
                             captured once in a hipGraph and replayed */
  int autotune;           /* 1: per-level work-item shapes picked by timing candidates on the
                             batch's own buffers (cached per context by the batch's sizes) */
  int tile_x;             /* nodes per row block of the residual/prolongation/restriction
                             kernels, or of the BiCGStab vector kernels (whole rows; 0 = default
                             256/128, BiCGStab 4096) */
  int tile_y;             /* smoother / operator-apply strip height in rows (0 = by level size,
                             BiCGStab 16) */
  int waves_per_cu;       /* reserved, 0 */
} pdehip_batch_config;
int pdehip_solve_batch(pdehip_ctx* ctx, int count, const pdehip_problem* probs,
                       const pdehip_options* opt, const pdehip_batch_config* cfg,
                       double* const* u_out, pdehip_report* reports);
/* Resident batch: operators, hierarchies and Krylov state of `count` problems kept in HBM and
 * solved together in lock step (2D constant-coefficient MG-PCG, or BiCGStab with the reference's
 * recomputed residual, FP64 or mixed precision): each operation of an iteration is one launch
 * over every unfinished problem.  No reference counterpart (the reference runs a batch as
 * sequential solves, src/main.cpp:2071-2097).  Mixed BiCGStab keeps the Krylov vectors and the
 * applies in FP32 and the solution and every iteration's residual b - A x in FP64; problems it
 * leaves unconverged start again from the initial guess in FP64 (report.refinements = -1).
 * A BiCGStab breakdown restarts the problem from its true residual (at most 8 times) instead of
 * ending it.  options.max_iter is rounded up to whole polling chunks (2 MG-PCG / 8 BiCGStab
 * iterations).
 * pdehip_batch_solve fills count reports and returns PDEHIP_OK when every problem converged,
 * else the status of the first one that did not; download copies problem `index`'s solution. */
typedef struct pdehip_batch pdehip_batch;
int pdehip_batch_create(pdehip_ctx* ctx, int count, const pdehip_problem* probs,
                        const pdehip_options* opt, const pdehip_batch_config* cfg,
                        pdehip_batch** out);
int pdehip_batch_solve(pdehip_batch* batch, pdehip_report* reports);
int pdehip_batch_download(pdehip_batch* batch, int index, double* u_out);
size_t pdehip_batch_bytes(const pdehip_batch* batch);
/* Measurement: mean duration of one batched level-0 launch over every problem, timed with HIP
 * events on the batch stream.  kernel: 0 smoother red+black sweep, 1 residual, 2 Krylov
 * apply + dot, 3 FP64 residual (a BiCGStab batch: 2 and 3 only).  algorithmic_bytes (may be NULL): compulsory bytes per launch.
 * Overwrites the batch's work vectors and scalars; the next pdehip_batch_solve starts afresh. */
int pdehip_batch_time_kernel(pdehip_batch* batch, int kernel, int reps, double* avg_ms,
                             double* algorithmic_bytes);
void pdehip_batch_destroy(pdehip_batch* batch);

/* ---------------------------------------------------------------- serving */
typedef struct pdehip_server pdehip_server;
/* The server runs two kinds of device lanes, each on its own HIP stream (default priority):
 *  - plan workers (`streams`): one resident plan per request (any problem the backend solves);
 *  - batch lanes (`batch_lanes`): requests the lock-step batched engine takes (2D CONST operators,
 *    MG-PCG without convection with even smoothing counts, or FP64 BiCGStab) are gathered, as
 *    many as are queued with the same options (up to `batch_max`), and solved together: one
 *    launch per operation over every member.  While one lane's batch runs, arrivals start on
 *    the next free lane.  Batch members keep fixed work-item shapes, so a request's result is
 *    bitwise the same alone and in a mix.
 * Admission: a request starts when the bytes reserved by running requests plus its own fit
 * the budget; its bytes are the measured size of an earlier plan of the same shape and
 * options, else an estimate.  Deadlines (timeout_ms) stop running solves on the device: plan
 * workers through the solve's cancel flag, batch lanes by marking the member done between
 * polling chunks; such requests end with PDEHIP_ERR_TIMEOUT. */
typedef struct pdehip_server_config {
  size_t hbm_budget_bytes;  /* admission budget (0 = context budget) */
  int max_inflight;         /* requests running on the device at once (0 = default 256) */
  int streams;              /* plan-worker streams (0 = default 4) */
  int cache_operators;      /* keep operators/hierarchies resident across requests */
  double timeout_ms;        /* per-request deadline from submission, 0 = none */
  int batch_lanes;          /* batch-lane streams (0 = default 2; -1 = none: every request
                               runs on a plan worker) */
  int batch_max;            /* most requests in one lane batch (0 = default 64) */
} pdehip_server_config;
typedef struct pdehip_server_stats {
  uint64_t submitted, completed, rejected, cache_hits;
  size_t bytes_in_use, peak_bytes;
  uint64_t batched;         /* requests solved in a batch lane */
  uint64_t batches;         /* lane batches run */
  uint64_t timed_out;       /* requests ended by their deadline */
} pdehip_server_stats;
int pdehip_server_create(pdehip_ctx* ctx, const pdehip_server_config* cfg, pdehip_server** out);
/* The problem's host arrays and u_out must stay valid until the request completes. */
int pdehip_server_submit(pdehip_server* srv, const pdehip_problem* prob,
                         const pdehip_options* opt, double* u_out, uint64_t* ticket);
/* Returns PDEHIP_PENDING while the request runs, else its final status (report filled). */
int pdehip_server_poll(pdehip_server* srv, uint64_t ticket, pdehip_report* report);
int pdehip_server_wait(pdehip_server* srv, uint64_t ticket, pdehip_report* report);
int pdehip_server_stats_get(pdehip_server* srv, pdehip_server_stats* stats);
void pdehip_server_destroy(pdehip_server* srv);

#ifdef __cplusplus
}
#endif
#endif /* PDEHIP_H */
