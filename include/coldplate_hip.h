/*
 * coldplate_hip.h -- C ABI of libcoldplate_hip.so, the MI355X (gfx950) backend
 * for the coldplate-solver SIMPLE hot path and its 3-D thermal solve.
 *
 * Conventions (all entry points):
 *   - plain C types, host pointers, sizes in elements; no exceptions cross
 *     the ABI; every function returns an int status (CPS_OK = 0) and the
 *     message of the last failure on this thread is in cps_last_error();
 *   - arrays are row-major.  Grid arrays use the reference's padded staggered
 *     shapes with M = ny + 1, N = nx + 1 (SIMPLE.cpp:53-54):
 *       u: (M+1) x N, v: M x (N+1), p: (M+1) x (N+1), gamma/alpha: M x N;
 *   - one context owns one device and one HIP stream; calls on one context
 *     are serialised by the caller (the reference's solvers are not
 *     re-entrant either: pressure_solver.cpp:502-527 keeps static state).
 */
#ifndef COLDPLATE_HIP_H
#define COLDPLATE_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define CPS_OK 0
#define CPS_ERR_ARG 1   /* invalid argument (sizes, null pointers, options) */
#define CPS_ERR_HIP 2   /* a HIP runtime call failed */
#define CPS_ERR_STATE 3 /* object used in the wrong state */
#define CPS_ERR_DIVERGED 4 /* NaN/Inf detected in a residual */

typedef struct cps_ctx cps_ctx;
typedef struct cps_simple cps_simple;
typedef struct cps_thermal cps_thermal;

/* Outcome of an iterative solve (thermal, adjoints). */
typedef struct cps_solve_stats {
  int iterations;            /* Krylov iterations */
  int refinements;           /* restarts from the true residual */
  double rel_residual;       /* ||b - A x|| / ||b|| in FP64 at exit */
  int converged;             /* rel_residual <= rtol */
} cps_solve_stats;

/* Message of the last failed call on the calling thread ("" if none). */
const char* cps_last_error(void);
/* ABI version, bumped on any incompatible change. */
int cps_abi_version(void);

/* Context: one device + one stream.  Replaces the process-global state that
 * amgcl_cuda::initialize / cleanup manage (cuda_solver.hpp:17-21). */
int cps_create(int device, cps_ctx** out);
void cps_destroy(cps_ctx* ctx);
int cps_synchronize(cps_ctx* ctx);

/* ---------------------------------------------------------------------------
 * CSR drop-ins.  Same contract as the reference's GPU backend ABI
 * (cuda_solver.hpp:24-57): host CSR in, solution in/out (initial guess in),
 * iteration count out.
 *
 * cps_pressure_solve_csr replaces amgcl_cuda::solve (cuda_solver.hpp:24-33);
 * it runs the reference CPU Jacobi-PCG semantics of solvePCG_CSR
 * (pressure_solver.cpp:121-184): stop when ||r||_2 < tolerance (absolute) or
 * |p.Ap| < 1e-20; *iterations = loop index at exit.
 *
 * cps_momentum_jacobi_csr replaces amgcl_cuda::solveMomentumU/V
 * (cuda_solver.hpp:36-57) with the CPU convergence rule of
 * solvers::cpu::solveMomentumJacobi (jacobi_solver.cpp:20-75): stop after the
 * sweep whose pre-update residual satisfies sum r^2 < tol^2 n.
 * ------------------------------------------------------------------------- */
int cps_pressure_solve_csr(cps_ctx* ctx, int n, const int* row_ptr, const int* col_idx, const float* values,
                           const float* rhs, float* solution, float tolerance, int max_iterations,
                           int* iterations);
int cps_momentum_jacobi_csr(cps_ctx* ctx, int n, const int* row_ptr, const int* col_idx, const float* values,
                            const float* rhs, float* solution, float tolerance, int max_iterations,
                            int* iterations);
/* cps_pressure_solve_csr without device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (tight
 * per-iteration loops such as the reference's own call site,
 * pressure_solver.cpp:628, serving loops, harnesses that forbid hipMalloc).
 * Same arguments, semantics and results as cps_pressure_solve_csr; its
 * buffers (the uploaded CSR, the PCG vectors, partials and state) live in one
 * static device array per GPU (about 40 MB, held until the process ends),
 * whose address cps_create looks up.  Limits: n <= 524288 rows and
 * row_ptr[n] <= 2621440 nonzeros, else CPS_ERR_ARG.  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (from several
 * threads or contexts) race. */
int cps_pressure_solve_csr_noalloc(cps_ctx* ctx, int n, const int* row_ptr, const int* col_idx,
                                   const float* values, const float* rhs, float* solution, float tolerance,
                                   int max_iterations, int* iterations);

/* ---------------------------------------------------------------------------
 * Structured SIMPLE solver (device-resident fields).  Knobs carry the
 * reference names of SIMPLE.h; defaults come from cps_simple_config_default.
 * ------------------------------------------------------------------------- */
#define CPS_PRESSURE_SOR 0    /* pressureSolverType 0: red-black SOR, max-change stop */
#define CPS_PRESSURE_PCG 1    /* pressureSolverType 1: Jacobi-PCG, absolute tol */
#define CPS_PRESSURE_AMGCL 2  /* pressureSolverType 2: AMGCL CG, stop ||r|| <= p_tol ||b||; no AMGCL here, so it
                                 falls back (as SIMPLE.cpp:147-186 does) to the GMG PCG with that stop rule */
#define CPS_PRESSURE_LDLT 3   /* pressureSolverType 3: direct LDLT; falls back to the GMG PCG solved to 1e-12 ||b|| */
#define CPS_PRESSURE_AMGCL_CUDA 4 /* pressureSolverType 4 (the reference default, SIMPLE.h:145): as type 2 */
#define CPS_PRESSURE_GMG 5    /* new type: geometric-multigrid PCG on the device; the grid's pressure DOFs may
                                 span at most 2048 tiles of 64 x 16 (e.g. 2048 x 1024 cells) */

#define CPS_PREC_FP64 64      /* every kernel in FP64 */
#define CPS_PREC_MIXED 32     /* FP32 inner Jacobi / MG-PCG, FP64 state, residuals, refinement */

typedef struct cps_simple_config {
  int nx, ny;                /* physical cells; M = ny+1, N = nx+1 */
  double hx, hy;             /* cell size [m] */
  double rho, eta;           /* SIMPLE.h:374-375 */
  double u_in;               /* targetVel [m/s] (uniform inlet, or peak of inlet profile) */
  int convection_scheme;     /* 0 FOU, 1 SOU (SIMPLE.h:193) */
  int enable_25d;            /* SIMPLE.h:205 */
  double ht;                 /* Ht_channel [m] */
  double conv_factor_25d;    /* SIMPLE.h:206 */
  double sink_mult_25d;      /* SIMPLE.h:208 */
  double brinkman_kmin;      /* SIMPLE.h:242 */
  double brinkman_q;         /* SIMPLE.h:244 */
  double uv_alpha, p_alpha;  /* SIMPLE.h:397-398 */
  double momentum_tol;       /* SIMPLE.h:176 */
  int max_momentum_iter;     /* SIMPLE.h:177 */
  double p_tol;              /* SIMPLE.h:153 (absolute 2-norm) */
  int max_pressure_iter;     /* SIMPLE.h:152 */
  int pressure_solver;       /* CPS_PRESSURE_* */
  int precision;             /* CPS_PREC_* */
  double p_rtol;             /* GMG-PCG: also stop when ||r|| <= p_rtol ||r0|| (0 = off) */
  int max_refine;            /* mixed precision: max FP64 refinement sweeps per pressure solve */
  double epsilon;            /* SIMPLE.h:68 */
  int max_iterations;        /* SIMPLE.h:69 */
  double dp_fluid_tol;       /* pure-fluid test of sampleBoundaryPatch (postprocessing.cpp:268) */
  /* outer loop of cps_simple_run (SIMPLE.cpp:310-491) */
  double x_core_in, x_core_out;  /* xPlaneCoreInlet / Outlet [m] (SIMPLE.h:88-89) */
  int use_dp_convergence;    /* usePressureDropConvergence (SIMPLE.h:98) */
  int dp_window;             /* dpConvergenceWindow (SIMPLE.h:99), >= 1 */
  double dp_percent;         /* dpConvergencePercent (SIMPLE.h:100) */
  int use_dp_slope_gate;     /* usePressureDropSlopeGate (SIMPLE.h:101) */
  int dp_slope_window;       /* dpSlopeWindowIters (SIMPLE.h:102) */
  double dp_slope_max_delta; /* dpSlopeMaxDelta [Pa] (SIMPLE.h:103) */
  int inlet_ramp;            /* enableInletRamp (SIMPLE.h:385) */
  int ramp_steps;            /* rampSteps (SIMPLE.h:386); also gates the dP window */
} cps_simple_config;

/* Per-iteration statistics (iterations.cpp:499-533, 628-658). */
typedef struct cps_step_stats {
  double error;              /* max(residU, residV, residMass), raw RMS */
  double resid_u, resid_v, resid_mass;
  double resid_u_max, resid_v_max, resid_mass_max;
  int u_iters, v_iters, p_iters;
  int iteration;             /* SIMPLE iterations done so far on this object */
  /* pressure drops after the last iteration (SIMPLE.cpp:421-439): core =
   * samplePlaneAtX(x_core_in) - samplePlaneAtX(x_core_out), full = inlet -
   * outlet sampleBoundaryPatch; total = mass-flux-weighted total pressure */
  double dp_core_total, dp_core_static, dp_full_total, dp_full_static;
  int dp_converged;          /* cps_simple_run stopped on the dP window */
  int pad_;
} cps_step_stats;

/* Boundary patch metrics (postprocessing.cpp:245-327). */
typedef struct cps_patch {
  double area, mass_flux, avg_static, avg_dynamic, avg_total;
  int valid;
} cps_patch;

#define CPS_FIELD_U 0
#define CPS_FIELD_V 1
#define CPS_FIELD_P 2
#define CPS_FIELD_USTAR 3
#define CPS_FIELD_VSTAR 4
#define CPS_FIELD_PSTAR 5
#define CPS_FIELD_DE 6
#define CPS_FIELD_DN 7
#define CPS_FIELD_GAMMA 8
#define CPS_FIELD_ALPHA 9

void cps_simple_config_default(cps_simple_config* cfg);

/* gamma: ny x nx physical density (1 = fluid, clamped to [0,1]; padded like
 * SIMPLE.cpp:720-729).  inlet_u: M+1 per-u-row inlet speeds (before the
 * gamma > 0.5 inlet mask of boundaries.cpp:82-100) or NULL for the uniform
 * reference inlet u_in.  Replaces SIMPLE::SIMPLE() (SIMPLE.cpp:18-29). */
int cps_simple_create(cps_ctx* ctx, const cps_simple_config* cfg, const double* gamma, const double* inlet_u,
                      cps_simple** out);
void cps_simple_destroy(cps_simple* s);
/* n_steps SIMPLE iterations (SIMPLE::calculateStep, iterations.cpp:82-659);
 * stats of the last one in *last (may be NULL). */
int cps_simple_step(cps_simple* s, int n_steps, cps_step_stats* last);
/* SIMPLE::runIterations (SIMPLE.cpp:310-491): iterate while iteration <
 * max_iterations, error > epsilon and the core dP window has not settled.
 * With inlet_ramp the inlet speed (and the velocity clamp) is scaled by
 * (k+1)/ramp_steps on iteration k < ramp_steps.  From iteration ramp_steps
 * on, each core dP enters a window of dp_window values; the run stops once
 * the window is full, its standard deviation and the newest value's distance
 * to its mean are both below dp_percent % of max(|mean|, 1) and, with
 * use_dp_slope_gate, the least-squares slope of the last dp_slope_window
 * core dPs is at most dp_slope_max_delta / dp_slope_window.  The
 * pseudo-time controller (enablePseudoTimeStepping, off by default) is not
 * built.  Iteration counts restart at 0 on every call, as in the reference. */
int cps_simple_run(cps_simple* s, cps_step_stats* last);
/* One row of the reference's per-iteration logs (writeIterationLogs,
 * output.cpp:330-360): raw RMS residuals and the four pressure drops. */
typedef struct cps_iter_log {
  double resid_mass, resid_u, resid_v;
  double dp_core_total, dp_full_total, dp_core_static, dp_full_static;
  int p_iters, pad_;
} cps_iter_log;
/* cps_simple_run that also fills log[k] for iterations k < log_cap (log may
 * be NULL); *n_iters (may be NULL) receives the iterations done. */
int cps_simple_run_logged(cps_simple* s, cps_step_stats* last, cps_iter_log* log, int log_cap, int* n_iters);
/* cps_simple_run_logged with the host's polling interval exposed.  The loop's
 * services (the log rows, the dP window, the slope gate and the stop test)
 * run on the device after every iteration; the host enqueues poll_every
 * iterations between reads of the device's stop flag, so a run of k
 * iterations synchronises ceil(k / poll_every) times (+1 to copy the log
 * out), reported in *host_syncs (may be NULL).  A chunk that passes the stop
 * iteration is rolled back to its start and replayed to the stop, which
 * reproduces the per-iteration loop bit for bit.  cps_simple_run and
 * cps_simple_run_logged poll every 32 iterations. */
int cps_simple_run_polled(cps_simple* s, int poll_every, cps_step_stats* last, cps_iter_log* log, int log_cap,
                          int* n_iters, int* host_syncs);
int cps_simple_get_field(cps_simple* s, int field, double* out);
int cps_simple_set_field(cps_simple* s, int field, const double* in);
/* inlet = 1: x = 0 patch, 0: x = L patch. */
int cps_simple_boundary_patch(cps_simple* s, int inlet, cps_patch* out);
/* samplePlaneAtX (postprocessing.cpp:102-243): p, u and v interpolated
 * linearly between the two cell columns whose centres bracket x (clamped to
 * the first / last centre), over rows where both cells are pure fluid
 * (gamma >= 1 - dp_fluid_tol). */
int cps_simple_plane(cps_simple* s, double x, cps_patch* out);

/* Discrete adjoint of the converged state (no reference counterpart: the
 * reference optimiser's sensitivities ignore the flow, mma_optimizer.py:
 * 442-567).  Linearises the momentum and continuity equations the SIMPLE
 * iteration converges to (momentum_solver.cpp:195-314, convection.cpp:53-186,
 * pressure_solver.cpp:470-496, boundaries.cpp:50-146) at the current fields
 * and solves the transposed system (FGMRES, SIMPLE block preconditioner) to
 * ||g - J^T lambda|| <= rtol ||g||.  dp: the full static pressure drop of
 * sampleBoundaryPatch (postprocessing.cpp:245-327, dp_fluid_tol rows);
 * dJ_dgamma (ny x nx, may be NULL): its derivative through the Brinkman term
 * alpha(gamma) (SIMPLE.cpp:751-777), evaluated at gamma clamped to [0, 1].
 * The result is exact for the discrete equations when the state is a fixed
 * point (cps_simple_state_residual). */
int cps_simple_gradient_dp(cps_simple* s, double rtol, int max_iter, double* dp, double* dJ_dgamma,
                           cps_solve_stats* stats);
/* The same adjoint for a caller's objective J(u, v, p) of the converged
 * state (no reference counterpart; the coupled optimiser objective of
 * mma_optimizer.py:477-487 reaches the flow through the thermal solve's
 * velocity): dJ_du, dJ_dv, dJ_dp are dJ/d(field) in the field shapes above
 * (NULL = 0).  Entries the boundary rules fix (inlet, walls, pinned pressure
 * columns) drop out and ghost copies count for the DOF they copy
 * (boundaries.cpp:50-146); dJ_dgamma (ny x nx) = -lambda^T dR/dgamma with
 * J^T lambda = dJ/dw solved to rtol. */
int cps_simple_gradient_vjp(cps_simple* s, const double* dJ_du, const double* dJ_dv, const double* dJ_dp,
                            double rtol, int max_iter, double* dJ_dgamma, cps_solve_stats* stats);
/* max |R| over the U-momentum, V-momentum and continuity equations at the
 * current fields (out3[3]). */
int cps_simple_state_residual(cps_simple* s, double* out3);
/* y = J x (transpose = 0) or J^T x (1) with J the Jacobian of those
 * equations at the current fields; x, y are [u | v | p] in the field shapes
 * above, concatenated; entries that are not DOFs pass through (identity). */
int cps_simple_jacobian_apply(cps_simple* s, int transpose, const double* x, double* y);
/* Preconditioner of the adjoint solve behind cps_simple_gradient_dp
 * (no reference counterpart).  All forms solve the same system and give the
 * same gradient; they differ in iteration count and cost per iteration:
 *   CPS_ADJ_PRECOND_SIMPLE    momentum by 1/aP0, Schur complement by the
 *                             pressure-correction operator;
 *   CPS_ADJ_PRECOND_LSC_LINE  momentum by pentadiagonal x-line solves plus
 *                             line_sweeps damped x-line Jacobi sweeps, Schur
 *                             complement by the least-squares commutator;
 *   CPS_ADJ_PRECOND_LSC_ADI   as LSC_LINE with each sweep an x-line pass then
 *                             a y-line pass (default: line_sweeps 2,
 *                             omega 0.8, schur_iters 2).
 * line_sweeps in [0, 64], omega (sweep damping) in (0, 1]; schur_iters
 * (multigrid-PCG iterations per Schur-complement solve, rtol 1e-3) in
 * [1, 1000]. */
#define CPS_ADJ_PRECOND_SIMPLE 0
#define CPS_ADJ_PRECOND_LSC_LINE 1
#define CPS_ADJ_PRECOND_LSC_ADI 2
int cps_simple_adjoint_options(cps_simple* s, int preconditioner, int line_sweeps, double omega,
                               int schur_iters);

/* ---------------------------------------------------------------------------
 * Batches of fields (no reference counterpart: the reference solves one
 * design per process, run_solvers.py).  Each field keeps its own solver
 * state.  Fields with the same grid and step settings form a size class,
 * whose buffers share one allocation at a fixed stride; two plans step them:
 *   CPS_BATCH_CLASS      one launch per kernel of the SIMPLE chain covers
 *                        every field of a chain of consecutive fields of a
 *                        class (gridDim.z = fields; a class is split into
 *                        `streams` chains); one chain iteration (or `unroll`
 *                        of them) is one hipGraph, the chains' graphs on
 *                        streams of their own;
 *   CPS_BATCH_PER_FIELD  each field's own chain; fields spread over
 *                        `streams` streams, one graph per stream.
 * A field's results are the same, bit for bit, under either plan and equal
 * to stepping it alone.  Fields need capture-safe caps: max_momentum_iter
 * and max_pressure_iter <= 32, max_refine <= 4.  Creation runs one eager
 * iteration per field.
 * ------------------------------------------------------------------------- */
#define CPS_BATCH_CLASS 0
#define CPS_BATCH_PER_FIELD 1
typedef struct cps_batch cps_batch;
/* cfgs[n_fields]; gammas[k]: ny x nx of field k; inlets: NULL or n_fields
 * pointers (NULL entries = uniform inlet). */
int cps_batch_create(cps_ctx* ctx, int n_fields, const cps_simple_config* cfgs, const double* const* gammas,
                     const double* const* inlets, cps_batch** out);
void cps_batch_destroy(cps_batch* b);
/* Captures the batch iteration.  streams > 0: the per-field plan on that
 * many streams; 0: the autotuner times the class plan with 1 .. 4 chains
 * per class and unroll 1, 4, 8 and the per-field plan on 1 .. 32 streams
 * (max(reps x unroll, 16) iterations each, timed in two passes, the second
 * in reverse order, after fields younger than 128 iterations have reached
 * it; those iterations advance the fields) and keeps the fastest.  chosen: the streams the chosen plan uses (one per
 * chain for the class plan). */
int cps_batch_tune(cps_batch* b, int streams, int reps, int* chosen, double* us_per_step);
/* Captures the given plan without timing (a cached tuner choice): plan
 * CPS_BATCH_*; streams: chains per class (class plan; 0 = 1) or streams
 * (per-field plan); unroll in 1..64. */
int cps_batch_plan(cps_batch* b, int plan, int streams, int unroll);
/* The captured plan and its parameter (the `streams` argument of
 * cps_batch_plan: chains per class or streams), the streams its graphs use,
 * its unroll, the number of size classes and the tuner's time per batch
 * iteration (0 when not timed). */
int cps_batch_plan_info(cps_batch* b, int* plan, int* param, int* streams, int* unroll, int* classes,
                        double* us_per_step);
/* n_steps graph launches (asynchronous). */
int cps_batch_step(cps_batch* b, int n_steps);
int cps_batch_synchronize(cps_batch* b);
int cps_batch_field_stats(cps_batch* b, int field, cps_step_stats* out);
int cps_batch_get_field(cps_batch* b, int field, int which, double* out);
int cps_batch_boundary_patch(cps_batch* b, int field, int inlet, cps_patch* out);
/* cps_simple_gradient_dp for one field of the batch, on that field's stream.
 * Calls for different fields may run concurrently from different host
 * threads (while no batch iteration is in flight). */
int cps_batch_gradient_dp(cps_batch* b, int field, double rtol, int max_iter, double* dp, double* dJ_dgamma,
                          cps_solve_stats* stats);

/* Measurement hook (no reference counterpart): brackets one launch in 8 of
 * one kernel of the step with HIP events carried by the launch itself, on
 * the solver stream (sampling keeps the timing from slowing the run it
 * measures).  kernel: CPS_KT_*; -1 stops timing.  Selecting a kernel clears
 * the counters. */
#define CPS_KT_MG_SMOOTH0 0   /* fine-level multigrid Jacobi smoother (pressure) */
#define CPS_KT_MOM_JACOBI 1   /* fused U+V momentum Jacobi sweep; while selected, the FP32 path runs one sweep
                                 per launch (otherwise two per launch without host polling, same results) */
#define CPS_KT_PCG_STENCIL 2  /* fine-level PCG p-update + 5-point apply (classical form: caps above 32) */
int cps_simple_time_kernel(cps_simple* s, int kernel);
/* launches timed so far, their mean duration [us] and the algorithmic bytes
 * one launch moves (every operand read once, every result written once).
 * For CPS_KT_MG_SMOOTH0 only launches that did the work count: a fine down
 * kernel enqueued after its PCG converged exits on the device stop flag and
 * is left out of both figures. */
int cps_simple_kernel_timing(cps_simple* s, int* launches, double* avg_us, double* bytes_per_launch);

/* ---------------------------------------------------------------------------
 * 3-D conjugate heat transfer (ThermalSolver/thermal_solver.cpp).  Replaces
 * the assembly + AMGCL solve of ThermalSolver3D::solve (thermal_solver.cpp:
 * 422-553) and the post-processing of thermal_output.cpp:135-270.  Cells are
 * (i, j, k) with k the layer from the bottom; 3-D host arrays are
 * [k][j][i] with i fastest (thermal_solver.cpp:235), 2-D ones [j][i].
 * Layers are a list: channel[k] = 0 is solid (k_s), 1 is a channel layer
 * (gamma-interpolated k, advected with u_xy * pois(z)); nz_solid base layers
 * followed by nz_fluid channel layers is the reference stack.
 * ------------------------------------------------------------------------- */
typedef struct cps_thermal_config {
  int nx, ny, nz;
  double dx, dy;             /* cell size [m] (Params::dx, dy) */
  double ht;                 /* channel height H_t for pois(z) (thermal_solver.cpp:323-327) */
  double k_s, k_f, rho, cp, t_in;  /* thermal_solver.cpp:160-164 */
  int k_mode;                /* 0 linear, 1 RAMP, 2 harmonic (:182) */
  double qk;                 /* RAMP q_k (:183) */
  int precision;             /* CPS_PREC_FP64 or CPS_PREC_MIXED (FP32 multigrid inside FP64 BiCGStab) */
  double rtol;               /* stop at ||b - A T|| <= rtol ||b|| (:167) */
  int max_iter;              /* Krylov iterations per solve, all refinements together (:168) */
  int max_refine;            /* restarts from the true FP64 residual */
} cps_thermal_config;

/* ThermalMetrics (thermal_output.cpp:16-44, without the gradient fields),
 * the objective quantities of compute_optimization_results
 * (thermal_solver.cpp:361-381) and the energy balance (:592-620). */
typedef struct cps_thermal_metrics {
  double t_avg_base, t_avg_fluid, t_avg_solid, t_avg_global;
  double t_std_base, t_std_fluid, t_std_global;
  double t_range_base, t_range_fluid, t_range_global;
  double t_uniformity_base, t_uniformity_global;
  double tb_max, tb_pnorm;   /* over layer k = 0 */
  double q_in, q_out, t_out, imbalance_pct;
  /* |grad T| statistics (thermal_output.cpp:270-321): central differences
   * inside, one-sided at the faces (:63-108); volume-weighted means; the
   * maximum's cell is the first in (k, j, i) order; -1e30 when a region has
   * no cells; an axis of one cell contributes 0 */
  double grad_max, grad_avg, grad_max_base, grad_avg_base;
  int grad_max_i, grad_max_j, grad_max_k, pad_;
} cps_thermal_metrics;

#define CPS_THERMAL_OBJ_T_AVG_BASE 0  /* volume mean of T over the layers below the channel */
#define CPS_THERMAL_OBJ_PNORM_L0 1    /* (mean over layer 0 of T^p)^(1/p) */

void cps_thermal_config_default(cps_thermal_config* cfg);
/* dz, channel: nz entries; gamma, u, v (cell-centred velocity), q0 (bottom
 * heat flux [W/m^2]): ny x nx.  Assembles the operator. */
int cps_thermal_create(cps_ctx* ctx, const cps_thermal_config* cfg, const double* dz, const int* channel,
                       const double* gamma, const double* u, const double* v, const double* q0,
                       cps_thermal** out);
void cps_thermal_destroy(cps_thermal* t);
/* New gamma / u / v / q0 (NULL keeps the current array); reassembles. */
int cps_thermal_update(cps_thermal* t, const double* gamma, const double* u, const double* v, const double* q0);
/* Solves A T = b from T0 (NULL: T = t_in everywhere, thermal_solver.cpp:428). */
int cps_thermal_solve(cps_thermal* t, const double* T0, cps_solve_stats* stats);
int cps_thermal_get_temperature(cps_thermal* t, double* T);
int cps_thermal_get_metrics(cps_thermal* t, double pnorm_exp, cps_thermal_metrics* out);
/* A^T lambda = rhs (3-D arrays). */
int cps_thermal_solve_adjoint(cps_thermal* t, const double* rhs, double* lambda, cps_solve_stats* stats);
/* Objective J (CPS_THERMAL_OBJ_*) of the last solve and dJ/dgamma (ny x nx)
 * by one adjoint solve.  The reference's thermal adjoint is a stub
 * (thermal_adjoint.cpp:16-60); the derivative follows its dk_dgamma
 * (thermal_solver.cpp:317-321) through every conduction term of the
 * assembly; the region masks (gamma > 0.01) are held fixed. */
int cps_thermal_gradient(cps_thermal* t, int objective, double pnorm_exp, double* dJ_dgamma, double* J,
                         cps_solve_stats* stats);
/* 2-D conduction surrogate adjoint of AdjointMethod/thermal_adjoint.py:8-159
 * (solve_thermal_adjoint) on the device.  k, P, T_in, gamma, T_out, lambda,
 * sens, dJ_dgamma: ny x nx, row = y.  K: harmonic-mean faces k_eff / dx^2 with
 * dx = 1 / nx and 2 k / dx^2 per boundary side; T = K^-1 P, or T_in when
 * given (the external-T mode of run_coldplate_adjoint.py:112-117); lambda =
 * K^-T (1/n); sens = dQ/dk as in the script (:101-123).  With gamma,
 * dJ_dgamma = sens * dk/dgamma of the RAMP law (k_s, k_f, q_k) at the
 * clipped gamma (run_coldplate_adjoint.py:51-70, 136-150).  Solves run to
 * ||b - K x|| <= rtol ||b||.  Output pointers may be NULL. */
int cps_surrogate_adjoint(cps_ctx* ctx, int nx, int ny, const double* k, const double* P, const double* T_in,
                          double rtol, const double* gamma, double k_s, double k_f, double q_k, double* T_out,
                          double* lambda, double* sens, double* dJ_dgamma, cps_solve_stats* stats);
/* Drop-in for the AMGCL call itself (thermal_solver.cpp:534-552:
 * Solver solver(crs(n, n, ptr, col, val), prm); solver(b, T)): the caller's
 * assembled matrix in CSR (n = nx ny nz rows, row r = (k ny + j) nx + i as
 * thermal_solver.cpp:235 numbers them; duplicate entries add up) must have
 * the 7-point pattern on that grid, else CPS_ERR_ARG.  T: initial guess in
 * (the reference starts from T_in, :428), solution out; stop at
 * ||b - A T|| <= rtol ||b|| or max_iter Krylov iterations (AMGCL's tol /
 * maxiter, :547-548) with the device BiCGStab + geometric multigrid of
 * cps_thermal_solve in CPS_PREC_FP64 or CPS_PREC_MIXED.  stats (may be NULL)
 * take the iterations and the relative residual, the pair solver(b, T)
 * returns. */
int cps_thermal_solve_csr(cps_ctx* ctx, int nx, int ny, int nz, const int* row_ptr, const int* col_idx,
                          const double* values, const double* rhs, double* T, double rtol, int max_iter,
                          int precision, cps_solve_stats* stats);
/* -lambda^T d(A T - b) / d(gamma, u, v) at the last solve's T for a caller's
 * adjoint lambda (nz x ny x nx, from cps_thermal_solve_adjoint with the
 * caller's dJ/dT): the design sensitivity through k(gamma) as in
 * cps_thermal_gradient, and the sensitivities to the cell-centred channel
 * velocity (ny x nx each) through the upwind advection, inlet and outflow
 * terms (thermal_solver.cpp:469-507); the gamma > 0.01 masks are held fixed.
 * Outputs may be NULL.  No reference counterpart: it closes the chain rule
 * from a thermal objective through the velocity to the flow adjoint. */
int cps_thermal_sensitivities(cps_thermal* t, const double* lambda, double* dJ_dgamma, double* dJ_du,
                              double* dJ_dv);
/* Assembled operator as 7 stencil arrays [d, w, e, s, n, b, t] of nz*ny*nx
 * each (entry of A for the neighbour; 0 where none) and the right side b. */
int cps_thermal_get_operator(cps_thermal* t, double* coef7, double* rhs);

#ifdef __cplusplus
}
#endif

#endif /* COLDPLATE_HIP_H */
