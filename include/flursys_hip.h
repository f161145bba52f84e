/*
 * flursys_hip.h -- C ABI of the MI355X (gfx950) pressure-projection library.
 *
 * Plain C: opaque handles, POD configs, caller-owned host arrays, status codes.
 * No exceptions cross this boundary and no torch types appear in it.  One HIP
 * stream per context; a context and everything created from it may be moved
 * to another thread but must not be used from two threads at once (the
 * reference's solver state is `Send`, not `Sync`: src/runtime.rs:117-120).
 *
 * Each entry point names the reference interface it replaces (file:line in
 * the FLURSYS Rust crate).  Array layouts are the reference's (src/field.rs):
 *   2D  index i + nx*j        u (nx+1) x ny, v nx x (ny+1), p nx x ny
 *   3D  index i + nx*(j+ny*k) u (nx+1) x ny x nz, v nx x (ny+1) x nz,
 *                             w nx x ny x (nz+1), p nx x ny x nz
 */
#ifndef FLURSYS_HIP_H
#define FLURSYS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------- */
/* Status codes.  1..7 mirror LinearAlgebraError (src/linear/mod.rs:17-51);
 * FL_ERR_SOLVER carries the structured solvers' String errors (text through
 * fl_last_error), e.g. "PCG pressure solver encountered a singular search
 * direction" (src/solver/incompressible.rs:1129-1161). */
typedef enum {
    FL_OK = 0,
    FL_ERR_DIMENSION_MISMATCH = 1,
    FL_ERR_INVALID_CSR = 2,
    FL_ERR_INDEX_OUT_OF_BOUNDS = 3,
    FL_ERR_NON_FINITE = 4,
    FL_ERR_MISSING_DIAGONAL = 5,
    FL_ERR_ZERO_DIAGONAL = 6,
    FL_ERR_BREAKDOWN = 7,
    FL_ERR_SOLVER = 8,
    FL_ERR_CONFIG = 9,
    FL_ERR_INVALID_ARGUMENT = 10,
    FL_ERR_HIP = 11,
    FL_ERR_OUT_OF_MEMORY = 12,
    FL_ERR_IO = 13,                /* an output file could not be written */
    FL_ERR_MESH = 14,              /* MeshError of a mesh build (kind in FL_MESH_*) */
    FL_ERR_POISSON = 15,           /* PoissonError of an assembly (kind in FL_POISSON_*) */
    FL_ERR_SIMPLE = 16             /* SimpleError of the unstructured SIMPLE (kind in FL_SIMPLE_*) */
} fl_status;

/* Stop rule: the reference's absolute max|r| < tol (incompressible.rs:
 * 1111-1164) or ||r||_2 <= tol * ||r_0||_2 (linear/solver.rs:84-85, atol 0). */
enum { FL_STOP_ABS_INF = 0, FL_STOP_REL_L2 = 1 };
/* Initial guess: previous pressure (incompressible.rs:1061) or zero. */
enum { FL_GUESS_PREVIOUS = 0, FL_GUESS_ZERO = 1 };
/* Pressure solver.  PCG is the reference's Jacobi-PCG; on the 3D cavity it is
 * the same algorithm on the pinned 7-point operator (the reference 3D solver
 * only has SOR, incompressible_3d.rs:293-332).  SOR is PressureSolverKind::Sor
 * (2D incompressible.rs:979-1046, 3D incompressible_3d.rs:293-332) in
 * red-black order: the same update, relaxation factor, gauge and max-residual
 * stop test after every sweep, reaching the same fixed point as the
 * reference's lexicographic sweep (iterates and counts differ; FP64 only). */
enum { FL_SOLVER_PCG = 0, FL_SOLVER_SOR = 1 };
/* Arithmetic of the pressure solve: FP64 throughout, or an FP32 stencil CG
 * inside FP64 iterative refinement (FP64 residual, dots, alpha/beta, test). */
enum { FL_PRECISION_FP64 = 0, FL_PRECISION_MIXED = 1 };
/* Preconditioner of the inner solve.  JACOBI is the reference's.  MULTIGRID
 * (no reference counterpart) runs flexible CG preconditioned by one
 * cell-centred geometric-multigrid V-cycle (3+3 damped-Jacobi sweeps per
 * level, piecewise-constant transfers with a 2.0 coarse-correction factor);
 * when an inner solve exceeds its budget or stagnates, the solve falls back
 * to plain FP64 Jacobi-PCG from the current iterate (counted in the report).
 * LEARNED (no reference counterpart) is MULTIGRID with a U-Net attached
 * with fl_cavity3d_set_unet as the V-cycle's coarse solve (or, where the grid
 * has the network's resolution, as a warm start). */
enum { FL_PRECOND_JACOBI = 0, FL_PRECOND_MULTIGRID = 1, FL_PRECOND_LEARNED = 2 };

typedef struct fl_ctx fl_ctx;
typedef struct fl_unet fl_unet;

/* Creates a context bound to HIP device `device` with its own stream. */
fl_status fl_create(int device, fl_ctx **out);
/* Releases the context.  Solvers created from it keep it alive: the context
 * is freed when the last of them is destroyed. */
void fl_destroy(fl_ctx *ctx);
/* Text of the last error raised on this context ("" if none). */
const char *fl_last_error(const fl_ctx *ctx);
/* Library version string. */
const char *fl_version(void);
/* Measurement utility with no reference counterpart: times a plain device
 * copy of `bytes` (read + write) averaged over `iterations` launches.  It is
 * the calibration kernel for HBM byte counters and the achievable-bandwidth
 * figure that roofline fractions are read against. */
fl_status fl_probe_copy(fl_ctx *ctx, int64_t bytes, int32_t iterations, double *ms_per_copy);

/* ---------------------------------------------------------------------- */
/* Pressure-solve report: the reference returns (residual, iterations)
 * (incompressible.rs:972-977); rel-L2 figures are added for the RelL2 rule. */
typedef struct {
    int64_t iterations;        /* CG iterations (inner iterations when mixed) */
    int64_t refinements;       /* outer FP64 refinement sweeps (mixed only) */
    int64_t fallback_iterations; /* plain FP64 PCG iterations after a fallback */
    int32_t fallback;          /* 1 when the plain PCG fallback ran */
    int32_t warm_start;        /* learned warm start: 1 kept, -1 rejected (it did not
                                  lower the residual and was undone), 0 none */
    double residual_abs_inf;   /* max|r| of the recurrence residual at exit */
    double residual_l2;        /* ||r||_2 at exit */
    double initial_l2;         /* ||r_0||_2 */
    int32_t converged;         /* 1 when the stop rule fired */
    int32_t status;            /* fl_status of the solve */
    double warm_start_ratio;   /* ||r||_2 after the warm start / ||r_0||_2 (0: none) */
    int32_t fallback_reason;   /* FL_FALLBACK_*: why the preconditioned solve handed over
                                  (also when the iteration cap left the fallback no room:
                                  fallback 0, converged 0) */
    int32_t reserved;
} fl_solve_report;

/* Why a preconditioned (multigrid / learned) solve handed over to the plain
 * FP64 Jacobi-PCG fallback. */
enum {
    FL_FALLBACK_NONE = 0,
    FL_FALLBACK_BUDGET = 1,      /* the inner iteration budget ran out */
    FL_FALLBACK_STAGNATION = 2,  /* ||r|| failed to halve over four iterations */
    FL_FALLBACK_BREAKDOWN = 3    /* a singular search direction or non-finite scalar */
};

/* ---------------------------------------------------------------------- */
/* 3D lid-driven cavity.  Mirrors LidDrivenCavity3DConfig
 * (src/solver/incompressible_3d.rs:8-47) plus the GPU path's knobs. */
typedef struct {
    int64_t nx, ny, nz;
    double length, height, depth;
    double density, lid_velocity, reynolds, dt;
    int64_t pressure_max_iters;
    double pressure_tolerance;     /* AbsInf tolerance, or rtol for RelL2 */
    int32_t stop_rule;             /* FL_STOP_* */
    int32_t initial_guess;         /* FL_GUESS_* */
    int32_t precision;             /* FL_PRECISION_* */
    int32_t preconditioner;        /* FL_PRECOND_* (multigrid needs FL_PRECISION_MIXED) */
    /* FL_PRECISION_MIXED only: relative tolerance of each inner FP32 solve
     * (0 selects 1e-5 for Jacobi-CG, 1e-4 for multigrid FCG).
     * pressure_max_iters caps the inner total. */
    double inner_tolerance;
    int32_t pressure_solver;       /* FL_SOLVER_* (SOR: FP64, Jacobi, AbsInf rule) */
    int32_t reserved;
    double pressure_omega;         /* SOR relaxation factor in (0, 2); reference default 1.7 */
} fl_cavity3d_config;

typedef struct {
    int64_t step;                  /* steps taken so far */
    double time;                   /* step * dt */
    int64_t pressure_iterations;
    int64_t pressure_refinements;
    double pressure_residual;      /* max|r| at exit of the solve */
    double pressure_residual_l2;
    double max_divergence;         /* incompressible_3d.rs:504-517 */
    double divergence_l2;          /* ||div u||_2 over cells */
    double max_speed;              /* incompressible_3d.rs:519-532 */
    int32_t pressure_converged;
    int32_t pressure_fallback;     /* 1 when the plain PCG fallback ran */
    int32_t pressure_warm_start;   /* fl_solve_report.warm_start of the step's solve */
    int32_t pressure_fallback_reason; /* fl_solve_report.fallback_reason */
    double pressure_warm_ratio;    /* fl_solve_report.warm_start_ratio */
} fl_cavity3d_step_report;

typedef struct fl_cavity3d fl_cavity3d;

/* LidDrivenCavity3DSolver::new (incompressible_3d.rs:117-134): validates like
 * LidDrivenCavity3DConfig::validate (:49-84) and allocates device fields (all
 * zero, the solver's initial state). */
fl_status fl_cavity3d_create(fl_ctx *ctx, const fl_cavity3d_config *cfg, fl_cavity3d **out);
void fl_cavity3d_destroy(fl_cavity3d *solver);
/* advance_one_step (incompressible_3d.rs:164-172): predictor, RHS, pressure
 * solve, correction, finite check -- all on device. */
fl_status fl_cavity3d_step(fl_cavity3d *solver, fl_cavity3d_step_report *report);
/* Steps `steps` times; the report is the last step's. */
fl_status fl_cavity3d_run(fl_cavity3d *solver, int64_t steps, fl_cavity3d_step_report *report);
/* Drop-in for solve_pressure_poisson (incompressible_3d.rs:293-332): solves
 * Lap p = rhs with p(0,0,0) = 0 (rhs[0] is ignored and treated as 0).
 * `p_inout` is the initial guess when initial_guess is FL_GUESS_PREVIOUS. */
fl_status fl_cavity3d_pressure_solve(fl_cavity3d *solver, const double *rhs, double *p_inout,
                                     fl_solve_report *report);
/* Host <-> device copies of the state (any pointer may be NULL). */
fl_status fl_cavity3d_get_fields(fl_cavity3d *solver, double *u, double *v, double *w,
                                 double *p);
fl_status fl_cavity3d_set_fields(fl_cavity3d *solver, const double *u, const double *v,
                                 const double *w, const double *p);
/* The predicted velocities u*, v*, w* and the pressure RHS of the last step. */
fl_status fl_cavity3d_get_intermediate(fl_cavity3d *solver, double *u_star, double *v_star,
                                       double *w_star, double *rhs);
/* Per-kernel device timing (no counterpart in the reference; for
 * benchmarking).  With `enable` non-zero, launches of the kernel named
 * `kernel` -- or of every kernel when it is NULL or "" -- are bracketed by
 * HIP events on the solver's stream; launches issued after a device-side
 * stop test has fired (they exit at once) are not counted.  Resets the
 * statistics.  Unknown names are FL_ERR_INVALID_ARGUMENT. */
typedef struct {
    char name[64];    /* kernel name with template arguments, as profilers show it */
    int64_t launches; /* timed launches */
    double total_ms;  /* summed device time of those launches */
    double bytes;     /* summed algorithmic HBM bytes of those launches */
} fl_kernel_stat;
fl_status fl_cavity3d_set_timing(fl_cavity3d *solver, int32_t enable, const char *kernel);
/* Statistics of kernel `index` (0-based, in a fixed order); past the last
 * kernel FL_ERR_INVALID_ARGUMENT. */
fl_status fl_cavity3d_kernel_stats(fl_cavity3d *solver, int32_t index, fl_kernel_stat *out);

/* ---------------------------------------------------------------------- */
/* 2D structured projection solver: IncompressibleSolver with projection
 * coupling (src/solver/incompressible.rs:293-746) on a staggered MAC grid
 * with an optional solid mask.  Boundaries follow SolverBoundaryOverrides
 * (src/preprocess.rs:1362-1407): per side a kind and constant (u, v); only
 * the right side's PressureOutlet makes outlet cells (incompressible.rs:
 * 2066-2082), as in the reference. */
enum { FL_BC_VELOCITY = 0, FL_BC_SYMMETRY = 1, FL_BC_OUTLET = 2 };
enum { FL_SIDE_LEFT = 0, FL_SIDE_RIGHT = 1, FL_SIDE_BOTTOM = 2, FL_SIDE_TOP = 3 };
/* ConvectionScheme (incompressible.rs:14-18). */
enum { FL_UPWIND = 0, FL_CENTRAL = 1 };

/* SimulationConfig (incompressible.rs:58-90), the fields this path uses. */
typedef struct {
    int64_t nx, ny;
    double length, height;
    double density, kinematic_viscosity;
    double dt, t_end;
    int64_t pressure_max_iters;
    double pressure_tolerance;
    int32_t convection;            /* FL_UPWIND / FL_CENTRAL */
    int32_t stop_rule;             /* FL_STOP_* */
    int32_t initial_guess;         /* FL_GUESS_* */
    int32_t adaptive;              /* TimeStepSettings::adaptive */
    int32_t bc_kind[4];            /* per side, FL_SIDE_* order */
    double bc_u[4], bc_v[4];       /* Velocity sides */
    double outlet_pressure[4];     /* PressureOutlet sides */
    /* TimeStepSettings (incompressible.rs:35-56) */
    double min_dt, max_dt, target_cfl, target_diffusion_number, max_growth_factor;
    /* GPU path: FL_PRECISION_FP64 is the reference's Jacobi-PCG;
     * FL_PRECISION_MIXED runs FP32 Jacobi-CG cycles on the correction inside
     * FP64 iterative refinement (x, every true residual, the dots and the
     * stop test stay FP64).  A batch's grids share one precision. */
    int32_t precision;
    int32_t reserved;
    /* FL_PRECISION_MIXED: relative residual drop that ends an inner cycle
     * (0 selects 1e-5).  pressure_max_iters caps the inner total. */
    double inner_tolerance;
    /* PressureSolverKind (incompressible.rs:20-24): FL_SOLVER_PCG, or
     * FL_SOLVER_SOR (FP64, the max-residual rule, relaxation pressure_omega in
     * (0, 2); single solvers and host-stepped batches). */
    int32_t pressure_solver;
    /* Normal velocity of a left Velocity side: FL_INLET_UNIFORM, u = bc_u[LEFT];
     * FL_INLET_PARABOLIC, u(y) = 6 ubar eta (1 - eta) with eta = (y - s) /
     * (H - s) clamped to [0, 1], u = 0 for y <= s, ubar = bc_u[LEFT] and s =
     * inlet_step_height -- Case::boundary_velocity's Channel (s = 0) and
     * BackwardFacingStep inlets (src/cases/mod.rs:141-170), which apply when
     * the side has no velocity override. */
    int32_t inlet_profile;
    double pressure_omega;
    double inlet_step_height;
    /* PressureVelocityCoupling (incompressible.rs:26-33): FL_COUPLING_PROJECTION
     * (transient, the default) or FL_COUPLING_SIMPLE (steady SIMPLE-style
     * pseudo-steps, incompressible.rs:659-712: the predictor adds the relaxed
     * pressure gradient, the solve is for a correction p' from zero, the
     * velocity correction and the right-hand side carry velocity_relaxation,
     * p += pressure_relaxation p'; time = step * dt and no end time; fixed dt
     * only).  The relaxation factors lie in (0, 1]; 0 selects the reference's
     * defaults, 0.7 and 0.3. */
    int32_t coupling;
    int32_t reserved2;
    double velocity_relaxation, pressure_relaxation;
} fl_grid2d_config;
enum { FL_INLET_UNIFORM = 0, FL_INLET_PARABOLIC = 1 };
enum { FL_COUPLING_PROJECTION = 0, FL_COUPLING_SIMPLE = 1 };

/* SolverStep (incompressible.rs:196-214), flow-only fields. */
typedef struct {
    int64_t iteration;
    double time, time_step;
    double continuity_residual;    /* max divergence */
    double pressure_residual;
    int64_t pressure_iterations;
    double momentum_residual;      /* velocity change */
    double max_speed, momentum_cfl, viscous_diffusion_number;
    int32_t pressure_converged;
    int32_t reached_end_time;
} fl_grid2d_step_report;

typedef struct fl_grid2d fl_grid2d;

/* IncompressibleSolver::new (incompressible.rs:320-366): validates, uploads
 * the mask (nx*ny bytes, non-zero = solid; NULL = none) and initialises the
 * velocity (zero field, blocked faces 0, boundary values at t = 0). */
fl_status fl_grid2d_create(fl_ctx *ctx, const fl_grid2d_config *cfg, const uint8_t *solid,
                           fl_grid2d **out);
void fl_grid2d_destroy(fl_grid2d *grid);
/* advance (incompressible.rs:523-555) without file output: one projection
 * step with the pressure solved by Jacobi-PCG on the device. */
fl_status fl_grid2d_step(fl_grid2d *grid, fl_grid2d_step_report *report);
/* solve_pressure_pcg (incompressible.rs:1048-1214) on host arrays: rhs as
 * build_pressure_rhs produces it; p_inout is the warm start. */
fl_status fl_grid2d_pressure_solve(fl_grid2d *grid, const double *rhs, double *p_inout,
                                   fl_solve_report *report);
fl_status fl_grid2d_get_fields(fl_grid2d *grid, double *u, double *v, double *p);
fl_status fl_grid2d_set_fields(fl_grid2d *grid, const double *u, const double *v,
                               const double *p);
/* field_update (incompressible.rs:563-586): cell-centred u, v, speed and
 * vorticity (compute_cell_fields :1381-1439); any pointer may be NULL. */
fl_status fl_grid2d_cell_fields(fl_grid2d *grid, double *u_cell, double *v_cell, double *speed,
                                double *vorticity);

/* Batched 2D cavities of uneven sizes: `count` independent problems advance
 * together, every kernel covering all of them and the pressure solves of all
 * of them running in one persistent cooperative kernel (no host round trip
 * inside a step).  Each problem keeps its own dt (adaptive if requested). */
typedef struct fl_batch2d fl_batch2d;
fl_status fl_batch2d_create(fl_ctx *ctx, int32_t count, const fl_grid2d_config *cfgs,
                            fl_batch2d **out);
void fl_batch2d_destroy(fl_batch2d *batch);
/* One step of every problem that has not reached its t_end; reports[count]
 * may be NULL. */
fl_status fl_batch2d_step(fl_batch2d *batch, fl_grid2d_step_report *reports);
fl_status fl_batch2d_get_fields(fl_batch2d *batch, int32_t index, double *u, double *v,
                                double *p);
/* `steps` steps driven entirely on the device (no reference counterpart; the
 * same numerics as fl_batch2d_step): the dt rule, t_end, counters and error
 * checks run in kernels, each step is the captured pre-solve half, the
 * cooperative persistent PCG and the captured post-solve half (HIP graphs
 * when `use_graphs`), and the host synchronises once.  A grid that fails
 * stops (its report status is FL_ERR_SOLVER) while the others go on; the
 * call then returns FL_ERR_SOLVER with the lowest-index grid's message.
 * reports[count] and step_ms[steps] (device time of each batch step) may be
 * NULL. */
typedef struct {
    int64_t steps;                 /* steps taken in this run */
    double time;                   /* simulated time after the run */
    int64_t pressure_iterations;   /* summed over the run */
    double last_pressure_residual; /* max|r| of the last solve */
    double max_divergence;         /* of the last step */
    double max_speed;
    int32_t reached_end_time;
    int32_t status;                /* FL_OK, or FL_ERR_SOLVER when the grid failed */
} fl_batch2d_run_report;
/* step_ms / pcg_ms (optional, `steps` entries): device time of each step and
 * of its persistent PCG launch (-1 when the step replays as one graph). */
fl_status fl_batch2d_run(fl_batch2d *batch, int64_t steps, int32_t use_graphs,
                         fl_batch2d_run_report *reports, double *step_ms, double *pcg_ms);
/* Launch shape of the batch's persistent PCG (no reference counterpart):
 * `threads` per block (256, 512 or 1024) and `blocks` (clamped to what is
 * co-resident and to the tile count).  The blocks share the iterating
 * grids' cells out equally every iteration; results do not depend on the
 * shape.  A batch holds at most 256 grids. */
typedef struct fl_batch2d_tuning {
    int32_t threads;
    int32_t blocks;
    int32_t reserved[2];
    double ms; /* autotuner: time of the candidate's solve */
} fl_batch2d_tuning;
fl_status fl_batch2d_set_launch(fl_batch2d *batch, const fl_batch2d_tuning *launch);
/* How fl_batch2d_run(use_graphs = 1) replays a step (no reference
 * counterpart): FL_GRAPH_WHOLE_STEP, one captured graph holding the
 * persistent PCG as a cooperative kernel node; FL_GRAPH_SPLIT, captured
 * graphs of the step's two halves around a cooperative launch (where the
 * runtime cannot capture that launch); FL_GRAPH_NONE before the first
 * graph run or after a launch-shape change. */
enum { FL_GRAPH_NONE = 0, FL_GRAPH_WHOLE_STEP = 1, FL_GRAPH_SPLIT = 2 };
fl_status fl_batch2d_graph_mode(const fl_batch2d *batch, int32_t *mode);
fl_status fl_batch2d_get_launch(const fl_batch2d *batch, fl_batch2d_tuning *launch);
/* Autotuner of the launch shape (no reference counterpart): every candidate
 * (threads, co-resident block count) solves the same
 * pseudo-random right-hand sides from zero with each grid's tolerance,
 * capped at `iterations` PCG iterations, in scratch memory; the fastest is
 * kept and returned in `*chosen`.  When `table` (capacity `table_cap`) is
 * given it receives every candidate with its time, `*table_len` their
 * number. */
fl_status fl_batch2d_autotune(fl_batch2d *batch, int32_t iterations, fl_batch2d_tuning *chosen,
                              int32_t table_cap, fl_batch2d_tuning *table, int32_t *table_len);

/* ---------------------------------------------------------------------- */
/* CSR linear algebra: replaces CsrMatrix (src/linear/csr.rs:4-125) and the
 * free functions cg / pcg / bicgstab (src/linear/solver.rs:43-256).  Values
 * are FP64; on the device the row offsets and columns are stored as int32
 * (nnz and ncols must fit in 31 bits). */
typedef struct fl_csr fl_csr;

/* LinearSolverOptions (src/linear/solver.rs:5-20; defaults atol 1e-10,
 * rtol 1e-8, max_iterations 1000). */
typedef struct {
    double absolute_tolerance;
    double relative_tolerance;
    int64_t max_iterations;
} fl_lin_opts;

/* LinearSolveReport (src/linear/solver.rs:29-41): converged = 1 for
 * LinearSolverStatus::Converged, 0 for MaxIterations (a status, not an
 * error).  device_ms is the solve's device time (HIP events, no copies).
 * The mixed path (fl_csr_set_precision) reports its inner iterations over all
 * rounds, the FP64 refinement rounds and whether it handed over to FP64. */
typedef struct {
    int32_t converged;
    int32_t refinements;      /* FP64 residual rounds (mixed path; 0 for FP64) */
    int64_t iterations;
    double initial_residual;
    double final_residual;
    double relative_residual;
    double device_ms;
    int32_t fallback;         /* 1: the mixed solve finished on the FP64 solver */
    int32_t reserved;
} fl_lin_report;

/* The fields of LinearAlgebraError (src/linear/mod.rs:17-26) for the status
 * returned: `context` holds the reference's &'static str (NonFiniteValue
 * context or Breakdown reason) or the InvalidCsr message. */
typedef struct {
    int32_t kind;          /* the fl_status */
    int32_t reserved;
    int64_t index;         /* NonFiniteValue index, InvalidCsr entry */
    int64_t expected;      /* DimensionMismatch */
    int64_t actual;
    int64_t row;           /* MissingDiagonal / ZeroDiagonal / IndexOutOfBounds */
    int64_t column;
    char context[96];
} fl_err_detail;

/* CsrMatrix::new (src/linear/csr.rs:14-29): validates exactly like
 * validate_csr (:195-251) and uploads.  `detail` may be NULL. */
fl_status fl_csr_upload(fl_ctx *ctx, int64_t nrows, int64_t ncols, const int64_t *row_offsets,
                        const int64_t *column_indices, const double *values, int64_t nnz,
                        fl_csr **out, fl_err_detail *detail);
void fl_csr_destroy(fl_csr *matrix);
/* spmv_into (src/linear/csr.rs:89-106): y = A x; a non-finite output row is
 * NonFiniteValue{"SpMV output", row}. */
fl_status fl_spmv(fl_csr *matrix, const double *x, double *y, fl_err_detail *detail);
/* cg (src/linear/solver.rs:43-50).  x_inout is the initial guess (honoured)
 * and receives the solution. */
fl_status fl_cg(fl_csr *matrix, const double *rhs, double *x_inout, const fl_lin_opts *opts,
                fl_lin_report *report, fl_err_detail *detail);
/* pcg with JacobiPreconditioner::new (src/linear/solver.rs:52-60,
 * src/linear/preconditioner.rs:9-28). */
fl_status fl_pcg_jacobi(fl_csr *matrix, const double *rhs, double *x_inout,
                        const fl_lin_opts *opts, fl_lin_report *report, fl_err_detail *detail);
/* No counterpart in the reference.  The arithmetic of the matrix's
 * fl_cg / fl_pcg_jacobi / fl_bicgstab solves: FL_PRECISION_FP64 (the
 * default, the reference's solvers) or FL_PRECISION_MIXED, the same solver in
 * FP32 (values, vectors, SpMV and axpys; dots, alpha/beta and its tests in
 * FP64) on the correction of A e = r inside FP64 iterative refinement: every
 * round forms the true FP64 residual and makes the reference's stop test
 * (||r||_2 <= atol + rtol ||r_0||_2) on it, so an accepted solution meets the
 * FP64 solver's tolerance.  An inner solve ends when its residual is below
 * half that threshold plus inner_tolerance times the round's ||r|| (0
 * selects 1e-5; the round's system is scaled to a unit residual); an FP32
 * breakdown after progress starts a new round, while a diverging or
 * non-finite inner solve, or a round that does not halve the FP64 residual,
 * hands the solve over to the FP64 solver from the current iterate
 * (report.fallback).  fl_pcg_jacobi_noalloc and the device callers
 * (SIMPLE) always solve in FP64. */
fl_status fl_csr_set_precision(fl_csr *matrix, int32_t precision, double inner_tolerance);
/* bicgstab (src/linear/solver.rs:155-256). */
fl_status fl_bicgstab(fl_csr *matrix, const double *rhs, double *x_inout,
                      const fl_lin_opts *opts, fl_lin_report *report, fl_err_detail *detail);
/* No counterpart in the reference.  fl_pcg_jacobi for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops): the same kernels, results bitwise identical, but every device
 * buffer of the solve (work vectors, partial sums, solver state) lives in a
 * library-static array that each GPU holds (42 MiB) until the process ends;
 * the call makes no hipMalloc/hipFree.  Matrices of more than 1048576 rows
 * or columns are rejected (FL_ERR_INVALID_ARGUMENT).  All calls on one GPU
 * share that array: calls that overlap on one GPU (from several threads or
 * contexts) race and must be serialised by the caller. */
fl_status fl_pcg_jacobi_noalloc(fl_csr *matrix, const double *rhs, double *x_inout,
                                const fl_lin_opts *opts, fl_lin_report *report,
                                fl_err_detail *detail);

/* ---------------------------------------------------------------------- */
/* Learned warm start (no counterpart in the reference; the pressure path's
 * learned FCG).  A 3D U-Net (4 levels, 16/32/48/64 channels, 3x3x3 BF16
 * MFMA convolutions, ReLU, no biases) maps the right-hand side of a Poisson
 * problem on an n^3 grid to its zero-mean solution.  Inputs are normalised
 * by their RMS and targets by diag / (RMS G) with the network's output gain
 * G, where `diag` is the Laplacian's diagonal 2 (1/hx^2 + 1/hy^2 + 1/hz^2) at
 * the network's resolution.  Host arrays are [batch][n^3] FP32, x fastest. */
typedef struct {
    int32_t n;         /* resolution per axis: a power of two in [32, 256] */
    int32_t max_batch; /* largest training / inference batch, 1..64 */
    uint64_t seed;     /* weight initialisation stream */
} fl_unet_config;
fl_status fl_unet_create(fl_ctx *ctx, const fl_unet_config *cfg, fl_unet **out);
void fl_unet_destroy(fl_unet *net);
/* Parameters as FP32, layer by layer, each [cout][K padded to 32] with
 * k = tap * cin + c and tap = 9 (dz+1) + 3 (dy+1) + (dx+1), followed by the
 * output gain G (not trained; default n^2 / 2): targets are normalised by
 * diag / (RMS G), predictions are p = (o - mean o) RMS G / diag. */
int64_t fl_unet_num_params(const fl_unet *net);
fl_status fl_unet_get_params(fl_unet *net, float *params);
/* Loads parameters (entries that must be zero are zeroed) and resets Adam. */
fl_status fl_unet_set_params(fl_unet *net, const float *params);
/* One Adam step (beta 0.9 / 0.999) on the mean-squared error of the
 * zero-mean normalised prediction; `loss` receives the batch mean. */
fl_status fl_unet_train_step(fl_unet *net, int32_t batch, const float *rhs, const float *p,
                             double diag, float lr, float *loss);
/* Predicted zero-mean solutions. */
fl_status fl_unet_infer(fl_unet *net, int32_t batch, const float *rhs, double diag, float *p_out);
/* Diagnostics for numerical checks: the raw 16-channel output
 * ([batch][n^3][16]) of the normalised input, and the parameter gradient
 * of the training loss (no update). */
fl_status fl_unet_forward_raw(fl_unet *net, int32_t batch, const float *rhs, float *out16);
fl_status fl_unet_grads(fl_unet *net, int32_t batch, const float *rhs, const float *p, double diag,
                        float *grad, float *loss);
/* MFMA operand-layout check: d (16x16) = a (16x32) b (32x16), row-major,
 * operands rounded to BF16. */
fl_status fl_probe_mfma_bf16(fl_ctx *ctx, const float *a, const float *b, float *d);
/* Attaches `net` (may be NULL) to a solver created with FL_PRECOND_LEARNED:
 * the grid must be a cube of n * 2^L cells per axis with L a multigrid level
 * that keeps the full domain.  With L >= 1 the network is the coarse solve
 * of every V-cycle of the flexible CG, as set by fl_cavity3d_set_learned_mode;
 * with L = 0 it gives a separate warm start ahead of the multigrid FCG.  The
 * network must outlive its use by the solver and must not be used
 * concurrently. */
fl_status fl_cavity3d_set_unet(fl_cavity3d *solver, fl_unet *net);
/* How the network serves as the coarse solve on level L (no reference
 * counterpart).  FL_LEARNED_CORRECTED (the default): the network predicts
 * x = A_L^-1 f_L, the V-cycle continues below L on f_L - A_L x and adds x
 * back on the way up (the multigrid correction of the prediction).
 * FL_LEARNED_REPLACE: levels L and below are replaced by the prediction and
 * two damped-Jacobi sweeps on level L. */
enum { FL_LEARNED_CORRECTED = 0, FL_LEARNED_REPLACE = 1 };
fl_status fl_cavity3d_set_learned_mode(fl_cavity3d *solver, int32_t mode);
/* Training pairs: the last step's pressure right-hand side b = -rhs and
 * pressure p, mean-restricted to n^3 (n = nx / 2^L); `diag` receives the
 * Laplacian diagonal at that resolution. */
fl_status fl_cavity3d_coarse_pair(fl_cavity3d *solver, int32_t n, float *rhs_c, float *p_c,
                                  double *diag);
/* Training utility (no reference counterpart): with n > 0, each later
 * multigrid or learned pressure solve copies the right-hand side that
 * V-cycle number `iteration` (0-based) of its first inner FCG solve receives
 * on the multigrid level of n^3 cells (a level below the finest) aside;
 * n = 0 stops the capture.  fl_cavity3d_captured_rhs returns the last copy
 * (n^3 floats, x fastest) and the rediscretised Laplacian diagonal at that
 * resolution, the inputs of coarse-solve training pairs. */
fl_status fl_cavity3d_capture_coarse(fl_cavity3d *solver, int32_t n, int32_t iteration);
fl_status fl_cavity3d_captured_rhs(fl_cavity3d *solver, float *rhs_c, double *diag);

/* ---------------------------------------------------------------------- */
/* Device assembly of unstructured finite-volume systems (no reference
 * counterpart as a device path: it runs UnstructuredMesh::from_cells,
 * src/mesh.rs:317-423, assemble_poisson + apply_reference,
 * src/poisson.rs:146-263, 303-337, and CsrBuilder::finalize,
 * src/linear/csr.rs:159-192, as GPU sorts and scans).  2D polygon cells of
 * 3..16 vertices; faces are numbered in the reference's creation order and
 * carry its geometry; duplicate CSR entries are summed in insertion order
 * and exact zeros dropped. */
typedef struct fl_mesh2d fl_mesh2d;
enum {                          /* MeshError kinds (src/mesh.rs:270-310) */
    FL_MESH_OK = 0, FL_MESH_EMPTY = 1, FL_MESH_NON_FINITE_POINT = 2, FL_MESH_INVALID_CELL = 3,
    FL_MESH_INVALID_POINT_INDEX = 4, FL_MESH_DEGENERATE_FACE = 5, FL_MESH_NON_MANIFOLD = 6,
    FL_MESH_ZERO_VOLUME = 7, FL_MESH_INCONSISTENT_ORIENTATION = 8
};
enum {                          /* PoissonError kinds (src/poisson.rs:78-88) */
    FL_POISSON_OK = 0, FL_POISSON_MISSING_REFERENCE = 1, FL_POISSON_INCOMPATIBLE_SOURCE = 2,
    FL_POISSON_INVALID_REFERENCE_CELL = 3, FL_POISSON_INVALID_REFERENCE_VALUE = 4,
    FL_POISSON_NON_FINITE_SOURCE = 5, FL_POISSON_DEGENERATE_DISTANCE = 6,
    FL_POISSON_INVALID_GEOMETRY = 7, FL_POISSON_DEGENERATE_FACE_DISTANCE = 8,
    FL_POISSON_MISSING_BOUNDARY = 9, FL_POISSON_NUMERICS = 10, FL_POISSON_LINEAR = 11
};
/* UnstructuredMesh::from_cells for 2D polygons: points xy[2*npts], cell c's
 * vertices cell_vtx[cell_off[c] .. cell_off[c+1]).  A MeshError returns
 * FL_ERR_MESH with its kind in *mesh_error (may be NULL) and the text in
 * fl_last_error. */
fl_status fl_mesh2d_create(fl_ctx *ctx, int64_t npts, const double *xy, int64_t ncells,
                           const int64_t *cell_off, const int64_t *cell_vtx, fl_mesh2d **out,
                           int32_t *mesh_error);
void fl_mesh2d_destroy(fl_mesh2d *mesh);
fl_status fl_mesh2d_info(fl_mesh2d *mesh, int64_t *ncells, int64_t *nfaces);
/* Cell centroids and areas, face owner / neighbour (-1 on the boundary),
 * face centres and area vectors (out of the owner); any pointer may be NULL. */
fl_status fl_mesh2d_get(fl_mesh2d *mesh, double *cx, double *cy, double *volume, int64_t *owner,
                        int64_t *neighbour, double *fcx, double *fcy, double *sx, double *sy);
/* assemble_poisson with Diffusivity::Constant(gamma) (Linear face
 * interpolation, the default), optional cell source (host), per-face
 * boundary kinds (0 ZeroGradient, 1 FixedValue with bc_value; NULL: every
 * face ZeroGradient) and optional reference; the matrix stays on the device
 * as a solver matrix for fl_cg / fl_pcg_jacobi / fl_bicgstab, the rhs comes
 * back to the host.  A PoissonError returns FL_ERR_POISSON with its kind in
 * *poisson_error and, for an incompatible pure-Neumann source, the
 * integrated source in *detail (both may be NULL). */
fl_status fl_mesh2d_assemble_poisson(fl_mesh2d *mesh, double gamma, const double *source,
                                     const int32_t *bc_kind, const double *bc_value,
                                     int32_t has_reference, int64_t ref_cell, double ref_value,
                                     fl_csr **matrix, double *rhs, int32_t *poisson_error,
                                     double *detail);
/* Shape and host copies of a device CSR matrix (any pointer may be NULL). */
fl_status fl_csr_info(fl_csr *matrix, int64_t *nrows, int64_t *ncols, int64_t *nnz);
fl_status fl_csr_download(fl_csr *matrix, int64_t *row_offsets, int64_t *cols, double *vals);

/* ---------------------------------------------------------------------- */
/* Collocated unstructured SIMPLE on the device: solve_simple (src/simple.rs:
 * 148-311) for steady, constant-density flow on a 2D polygon mesh, with
 * every outer iteration's pieces as kernels -- the least-squares pressure
 * gradient (numerics.rs:650-682) and its source, the momentum CSR assembly
 * (momentum.rs:180-341: first-order upwind convection, two-point diffusion,
 * equation relaxation), BiCGSTAB on the x and y systems (momentum.rs:
 * 445-475), rAU, the Rhie-Chow predicted flux (simple.rs:746-805), the
 * pressure-correction CSR (simple.rs:518-638) and its Jacobi-PCG, the
 * pressure, velocity (LSQ gradient of p') and face-flux corrections and the
 * continuity RMS (simple.rs:703-717) -- with no host round trip of a field.
 * The velocity is (x, y): the z system of a 2D mesh has a zero right-hand
 * side and zero guess, and its report is 0 iterations, converged, as the
 * reference's BiCGSTAB returns it.  The pressure-correction matrix keeps the
 * momentum matrix's sparsity: with a reference cell its row and column hold
 * explicit zeros (and 1 on the diagonal) instead of being dropped, which
 * leaves every product and therefore the solve unchanged. */
typedef struct fl_simple2d fl_simple2d;
/* SimpleError kinds (src/simple.rs:17-29) and the NumericsError /
 * MomentumError cases the device path detects */
enum {
    FL_SIMPLE_INVALID_MOMENTUM_DIAGONAL = 1,
    FL_SIMPLE_INVALID_PRESSURE_FACE_COEFFICIENT = 2,
    FL_SIMPLE_INVALID_PRESSURE_REFERENCE = 3,
    FL_SIMPLE_INVALID_PRESSURE_RELAXATION = 4,
    FL_SIMPLE_INVALID_OUTER_ITERATIONS = 5,
    FL_SIMPLE_MOMENTUM_DID_NOT_CONVERGE = 6,
    FL_SIMPLE_PRESSURE_DID_NOT_CONVERGE = 7,
    FL_SIMPLE_SINGULAR_STENCIL = 8,         /* NumericsError::SingularLeastSquaresStencil */
    FL_SIMPLE_DEGENERATE_DISTANCE = 9,      /* owner-neighbour / owner-face distance */
    FL_SIMPLE_MISSING_BOUNDARY = 10,        /* a boundary face without a condition */
    FL_SIMPLE_BACKFLOW = 11,                /* MomentumError: backflow on a ZeroGradient face */
    FL_SIMPLE_NONFINITE_FLUX = 12,          /* MomentumError: non-finite face flux */
    FL_SIMPLE_LINEAR = 13,                  /* a LinearAlgebraError of a solve */
    FL_SIMPLE_INVALID_OPTIONS = 14          /* viscosity or velocity relaxation out of range */
};
/* SimpleOptions (simple.rs:93-131) with Diffusivity::Constant(viscosity). */
typedef struct {
    double viscosity;
    int64_t reference_cell;          /* used when no correction boundary anchors p' */
    int64_t max_outer_iterations;
    double continuity_absolute_tolerance;
    fl_lin_opts momentum_solver, pressure_solver;
    double velocity_relaxation, pressure_relaxation;
} fl_simple2d_options;
/* SimpleReport (simple.rs:133-142) plus the failing item of an error. */
typedef struct {
    int32_t converged;
    int32_t error_kind;              /* FL_SIMPLE_* when the call returned FL_ERR_SIMPLE */
    int64_t outer_iterations;
    double initial_continuity_rms, final_continuity_rms;
    fl_lin_report momentum[3];       /* last outer iteration's x, y, z solves */
    fl_lin_report pressure;          /* last pressure-correction solve */
    int64_t error_index;             /* the cell or face of error_kind */
    double error_value;
    double device_ms;                /* wall time of the outer iterations */
} fl_simple2d_report;
/* Per boundary face (arrays of nfaces, interior entries ignored):
 * velocity_kind 1 FixedValue (velocity_x/y, a wall or inlet velocity) or 0
 * ZeroGradient; pressure_kind / pressure_value, the LSQ stencil's boundary
 * (LeastSquaresGradientStencil::new, numerics.rs:106-200); correction_kind /
 * correction_value (NULL: none, the reference cell pins p'), the
 * pressure-correction boundary whose FixedValue faces anchor p' (an outlet).
 * The mesh must outlive the solver. */
fl_status fl_simple2d_create(fl_mesh2d *mesh, const int32_t *velocity_kind, const double *velocity_x,
                             const double *velocity_y, const int32_t *pressure_kind,
                             const double *pressure_value, const int32_t *correction_kind,
                             const double *correction_value, fl_simple2d **out,
                             fl_simple2d_report *error);
void fl_simple2d_destroy(fl_simple2d *simple);
/* initial_face_flux (simple.rs:719-744): interior faces from the linearly
 * interpolated cell velocity, boundary faces from wall_flux (host arrays). */
fl_status fl_simple2d_initial_face_flux(fl_simple2d *simple, const double *ux, const double *uy,
                                        const double *wall_flux, double *face_flux);
/* SimpleState: cell velocity (x, y), pressure and the authoritative face flux. */
fl_status fl_simple2d_set_state(fl_simple2d *simple, const double *ux, const double *uy,
                                const double *p, const double *face_flux);
fl_status fl_simple2d_get_state(fl_simple2d *simple, double *ux, double *uy, double *p,
                                double *face_flux);
/* solve_simple from the current state; history (may be NULL) receives the
 * continuity RMS before the first and after every outer iteration
 * (max_outer_iterations + 1 entries). */
fl_status fl_simple2d_solve(fl_simple2d *simple, const fl_simple2d_options *options,
                            fl_simple2d_report *report, double *history);

/* ---------------------------------------------------------------------- */
/* On-disk outputs (src/output.rs) in the reference's formats and number
 * formatting (Rust {:.12} / {:.12e}), written from host arrays filled by the
 * device (fl_grid2d_cell_fields, fl_cavity3d_cell_fields).  Host-only calls:
 * they take no context.  `err` (may be NULL) receives the reference's
 * "Cannot create <path>: <reason>" text on FL_ERR_IO.  2D arrays are nx*ny,
 * x fastest; `solid` and `temperature` may be NULL. */
/* write_field_csv (output.rs:31-67): field.csv */
fl_status fl_write_field_csv(const char *path, int64_t nx, int64_t ny, double dx, double dy,
                             const uint8_t *solid, const double *p, const double *u_cell,
                             const double *v_cell, const double *vorticity,
                             const double *temperature, char *err, int32_t errcap);
/* write_legacy_vtk (output.rs:69-125): field.vtk of the 2D solver */
fl_status fl_write_vtk_2d(const char *path, const char *title, int64_t nx, int64_t ny, double dx,
                          double dy, const uint8_t *solid, const double *p, const double *u_cell,
                          const double *v_cell, const double *vorticity, const double *temperature,
                          char *err, int32_t errcap);
/* write_legacy_vtk_3d (output.rs:127-183) with the face averages already
 * taken (cell-centred u, v, w of fl_cavity3d_cell_fields) */
fl_status fl_write_vtk_3d(const char *path, const char *title, int64_t nx, int64_t ny, int64_t nz,
                          double dx, double dy, double dz, const double *p, const double *u_cell,
                          const double *v_cell, const double *w_cell, char *err, int32_t errcap);
/* write_ppm_frame (output.rs:201-255): P6 image, rows top to bottom */
fl_status fl_write_ppm(const char *path, int64_t nx, int64_t ny, const uint8_t *solid,
                       const double *field, int32_t symmetric, char *err, int32_t errcap);
/* One history.csv row (append_histories, incompressible.rs:1744-1767);
 * with `extended` the header and row add cg_iters, hbm_GBps, roofline_frac
 * and fallback_count after the reference's thirteen columns. */
typedef struct {
    int64_t step;
    double time, pressure_residual;
    int64_t pressure_iterations;
    double max_divergence, velocity_change, max_speed, cd, cl, reattachment;
    double momentum_cfl, viscous_diffusion_number, time_step;
    int64_t cg_iters;
    double hbm_GBps, roofline_frac;
    int64_t fallback_count;
} fl_history_row;
/* append_history (output.rs:185-199): the header is written when the file
 * does not exist yet. */
fl_status fl_append_history(const char *path, const fl_history_row *row, int32_t extended,
                            char *err, int32_t errcap);
/* The case-specific profile files of write_case_specific_profiles
 * (incompressible.rs:1821-1913): a header line, then `nrows` rows of
 * `ncols` columns, column c formatted by fmt[c] ('f': {:.12}, 'e': {:.12e}). */
fl_status fl_write_csv_columns(const char *path, const char *header, int32_t ncols, int64_t nrows,
                               const double *const *cols, const char *fmt, char *err,
                               int32_t errcap);
/* Cell-centred fields of the 3D solver for output (incompressible_3d.rs via
 * output.rs:170-180): p and the averages of the two faces of each cell,
 * formed on the device; any pointer may be NULL. */
fl_status fl_cavity3d_cell_fields(fl_cavity3d *solver, double *p, double *u_cell, double *v_cell,
                                  double *w_cell);

#ifdef __cplusplus
}
#endif
#endif
