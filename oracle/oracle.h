/*
 * CPU oracle for the pressure-projection path: a plain-C (C11 + OpenMP), FP64
 * restatement of the FLURSYS reference algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code.  It is used by the tests (as the checker), by the package's
 * smoke() entry (as the checker) and by bench.py's cpu_baseline leg (timed
 * beside the GPU, never as the measured product).
 *
 * Every function cites the reference file:line it restates.  The reference is
 * a Rust crate that cannot be built here (no cargo/rustc, crates.io
 * unreachable, build.rs needs slint-build), so there is no oracle/_ref.
 *
 * Layouts mirror src/field.rs:31-34 (2D, i + nx*j) and src/field.rs:104-107
 * (3D, i + nx*(j + ny*k)) with the staggered extents of the MAC fields:
 *   2D  u (nx+1) x ny, v nx x (ny+1), p nx x ny
 *   3D  u (nx+1) x ny x nz, v nx x (ny+1) x nz, w nx x ny x (nz+1), p nx x ny x nz
 */
#ifndef FLURSYS_ORACLE_H
#define FLURSYS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- shared enums (same numeric values as include/flursys_hip.h) ---- */
enum { OR_STOP_ABS_INF = 0, OR_STOP_REL_L2 = 1 };
enum { OR_GUESS_PREVIOUS = 0, OR_GUESS_ZERO = 1 };
enum { OR_SOLVER_PCG = 0, OR_SOLVER_SOR = 1 };
enum { OR_UPWIND = 0, OR_CENTRAL = 1 };
enum { OR_BC_VELOCITY = 0, OR_BC_SYMMETRY = 1, OR_BC_OUTLET = 2 };
enum { OR_SIDE_LEFT = 0, OR_SIDE_RIGHT = 1, OR_SIDE_BOTTOM = 2, OR_SIDE_TOP = 3 };

/* Status codes (mirror src/linear/mod.rs:17-51 plus the structured-solver
 * string errors, which are reported through `message`). */
enum {
    OR_OK = 0,
    OR_ERR_DIMENSION = 1,
    OR_ERR_INVALID_CSR = 2,
    OR_ERR_INDEX_OOB = 3,
    OR_ERR_NON_FINITE = 4,
    OR_ERR_MISSING_DIAGONAL = 5,
    OR_ERR_ZERO_DIAGONAL = 6,
    OR_ERR_BREAKDOWN = 7,
    OR_ERR_SOLVER = 8,   /* structured-solver String error */
    OR_ERR_CONFIG = 9,
};

typedef struct {
    int64_t iterations;
    double residual_abs_inf;   /* max|r| of the recurrence residual at exit */
    double residual_l2;        /* ||r||_2 at exit */
    double initial_l2;         /* ||r0||_2 */
    int32_t converged;         /* 1 if the stop rule fired */
    int32_t status;            /* OR_OK or an error code */
    char message[160];
} or_solve_report;

/* ======================= 3D lid-driven cavity ======================= */
typedef struct {
    int64_t nx, ny, nz;
    double length, height, depth;
    double density, lid_velocity, reynolds, dt;
    int64_t pressure_max_iters;
    double pressure_tolerance;  /* AbsInf tolerance or rtol for RelL2 */
    double pressure_omega;
    int32_t solver;             /* OR_SOLVER_PCG (pinned PCG) or OR_SOLVER_SOR (reference) */
    int32_t stop_rule;
    int32_t initial_guess;
    int32_t threads;            /* 0 = OpenMP default */
} or3d_config;

typedef struct {
    or3d_config cfg;
    double dx, dy, dz;
    int64_t n;                  /* nx*ny*nz */
    double *p, *rhs, *u, *v, *w, *us, *vs, *ws;
    double *x, *b, *r, *z, *d, *q;  /* PCG workspace */
    int64_t step;
} or3d_state;

int or3d_init(or3d_state *s, const or3d_config *cfg);
void or3d_free(or3d_state *s);
void or3d_predict(or3d_state *s);
void or3d_build_rhs(or3d_state *s);
int or3d_solve_pressure(or3d_state *s, or_solve_report *rep);
void or3d_correct(or3d_state *s);
int or3d_step(or3d_state *s, or_solve_report *rep);
double or3d_max_divergence(const or3d_state *s);
double or3d_max_speed(const or3d_state *s);
double or3d_pressure_residual(const or3d_state *s);
void or3d_apply_operator(const or3d_state *s, const double *x, double *y);
int or3d_ensure_finite(const or3d_state *s, char *message, int cap);

/* ======================= 2D structured solver ======================= */
typedef struct {
    int64_t nx, ny;
    double length, height;
    double density, nu, dt;
    int64_t pressure_max_iters;
    double pressure_tolerance;
    double pressure_omega;
    int32_t convection;          /* OR_UPWIND / OR_CENTRAL */
    int32_t solver;              /* OR_SOLVER_PCG / OR_SOLVER_SOR */
    int32_t stop_rule;
    int32_t initial_guess;
    int32_t threads;
    int32_t bc_kind[4];          /* per side: OR_BC_* */
    double bc_u[4], bc_v[4];     /* Velocity side values */
    double outlet_pressure[4];   /* PressureOutlet value per side */
    /* adaptive time stepping (src/solver/incompressible.rs:35-56) */
    int32_t adaptive;
    double min_dt, max_dt, target_cfl, target_diffusion, max_growth;
    double t_end;
    /* left Velocity side: 0 uniform bc_u, 1 the parabolic Channel /
     * BackwardFacingStep inlet (src/cases/mod.rs:141-170) of mean bc_u[LEFT]
     * above inlet_step_height */
    int32_t inlet_profile;
    double inlet_step_height;
    /* PressureVelocityCoupling (incompressible.rs:26-33): OR_COUPLING_PROJECTION
     * or OR_COUPLING_SIMPLE with its relaxation factors in (0, 1] */
    int32_t coupling;
    double velocity_relaxation, pressure_relaxation;
} or2d_config;
enum { OR_COUPLING_PROJECTION = 0, OR_COUPLING_SIMPLE = 1 };

typedef struct {
    double time_step, pressure_residual, max_divergence, max_speed;
    double momentum_cfl, viscous_diffusion_number, velocity_change;
    int64_t pressure_iterations;
} or2d_diag;

typedef struct {
    or2d_config cfg;
    double dx, dy;
    int64_t n;
    uint8_t *solid;
    double *p, *rhs, *u, *v, *us, *vs, *uold, *vold;
    double *ucell, *vcell, *speed, *vort;
    double *x, *b, *r, *z, *d, *q;
    double *pold;                /* SIMPLE: the pressure before the p' solve */
    int64_t step;
    double time, current_dt;
} or2d_state;

int or2d_init(or2d_state *s, const or2d_config *cfg, const uint8_t *solid);
void or2d_free(or2d_state *s);
void or2d_apply_velocity_boundaries(or2d_state *s, double *u, double *v);
void or2d_predict(or2d_state *s, double dt);
int or2d_reached_end(const or2d_state *s);
void or2d_build_rhs(or2d_state *s, double dt);
int or2d_solve_pressure(or2d_state *s, or_solve_report *rep);
void or2d_correct(or2d_state *s, double dt);
void or2d_cell_fields(or2d_state *s);
int or2d_step(or2d_state *s, or2d_diag *diag, or_solve_report *rep);
double or2d_max_divergence(const or2d_state *s);
void or2d_apply_operator(const or2d_state *s, const double *x, double *y);
double or2d_pressure_residual(const or2d_state *s);
void or2d_stability_numbers(double dt, double nu, double dx, double dy, double umax,
                            double vmax, double *cfl, double *dnu);
int or2d_adaptive_dt(double current_dt, double umax, double vmax, double dx, double dy,
                     double nu, double thermal_diffusivity, double min_dt, double max_dt,
                     double target_cfl, double target_diffusion, double max_growth,
                     double *out_dt, char *message, int cap);

/* ======================= CSR linear algebra ======================= */
typedef struct {
    int32_t status;
    int32_t converged;           /* 1 Converged, 0 MaxIterations */
    int64_t iterations;
    double initial_residual, final_residual, relative_residual;
    int64_t err_index, err_expected, err_actual, err_row, err_col;
    char reason[96];
} or_lin_report;

int or_csr_validate(int64_t nrows, int64_t ncols, const int64_t *offsets, const int64_t *cols,
                    const double *vals, int64_t nnz, or_lin_report *rep);
int or_csr_spmv(int64_t n, const int64_t *offsets, const int64_t *cols, const double *vals,
                const double *x, double *y, or_lin_report *rep);
int or_csr_finalize(int64_t nrows, int64_t ncols, int64_t ntrip, const int64_t *trow,
                    const int64_t *tcol, const double *tval, int64_t *offsets, int64_t *cols,
                    double *vals, int64_t *nnz_out, or_lin_report *rep);
int or_jacobi_inverse(int64_t n, const int64_t *offsets, const int64_t *cols,
                      const double *vals, double *inv_diag, or_lin_report *rep);
int or_cg(int64_t n, const int64_t *offsets, const int64_t *cols, const double *vals,
          const double *rhs, double *x, double atol, double rtol, int64_t max_iter,
          int32_t jacobi, or_lin_report *rep);
int or_bicgstab(int64_t n, const int64_t *offsets, const int64_t *cols, const double *vals,
                const double *rhs, double *x, double atol, double rtol, int64_t max_iter,
                or_lin_report *rep);

/* ============ 2D unstructured mesh + two-point Poisson assembly ============ */
/* UnstructuredMesh (src/mesh.rs:317-423, 533-609), 2D polygons only. */
typedef struct {
    int64_t ncells, nfaces;
    double *cx, *cy, *vol;        /* cell centroid and area */
    int64_t *owner, *neigh;       /* neigh = -1 on boundary faces */
    double *fcx, *fcy, *sx, *sy;  /* face centre and area vector (out of the owner) */
    int64_t *fv0, *fv1;           /* oriented face vertices */
    int64_t *cell_off, *cell_faces; /* Cell::faces: the faces of cell c in its edge order */
} or_mesh2d;
/* MeshError kinds (src/mesh.rs:270-310), the ones 2D polygons can raise */
enum {
    OR_MESH_EMPTY = 20,
    OR_MESH_NON_FINITE_POINT = 21,
    OR_MESH_INVALID_CELL = 22,
    OR_MESH_INVALID_POINT_INDEX = 23,
    OR_MESH_DEGENERATE_FACE = 24,
    OR_MESH_NON_MANIFOLD = 25,
    OR_MESH_ZERO_VOLUME = 26,
    OR_MESH_INCONSISTENT_ORIENTATION = 27,
};
/* ScalarBoundaryCondition (the two the Poisson assembly uses) */
enum { OR_SCALAR_ZERO_GRADIENT = 0, OR_SCALAR_FIXED_VALUE = 1 };
/* PoissonError kinds (src/poisson.rs:78-88) and the NumericsError cases of
 * assemble_poisson */
enum {
    OR_POISSON_MISSING_REFERENCE = 30,
    OR_POISSON_INCOMPATIBLE_SOURCE = 31,
    OR_POISSON_INVALID_REFERENCE_CELL = 32,
    OR_POISSON_INVALID_REFERENCE_VALUE = 33,
    OR_POISSON_NON_FINITE_SOURCE = 34,
    OR_POISSON_DEGENERATE_DISTANCE = 35,
    OR_POISSON_INVALID_GEOMETRY = 36,
    OR_POISSON_DEGENERATE_FACE_DISTANCE = 37,
    OR_POISSON_MISSING_BOUNDARY = 38,
    OR_POISSON_NUMERICS = 39,
    OR_POISSON_LINEAR = 40,
};
/* Cells are polygons: vertex lists cell_vtx[cell_off[c] .. cell_off[c+1]). */
int or_mesh2d_from_cells(int64_t npts, const double *xy, int64_t ncells, const int64_t *cell_off,
                         const int64_t *cell_vtx, or_mesh2d *m, char *message, int cap);
void or_mesh2d_free(or_mesh2d *m);
/* assemble_poisson with Diffusivity::Constant(gamma), optional cell source,
 * per-face boundary conditions (bc_kind/bc_value indexed by face; interior
 * entries ignored) and an optional reference.  Outputs are malloc'ed (free
 * with or_free); `detail` receives the integrated source of an
 * IncompatiblePureNeumannSource error. */
int or_assemble_poisson(const or_mesh2d *m, double gamma, const double *source,
                        const int32_t *bc_kind, const double *bc_value, int32_t has_reference,
                        int64_t ref_cell, double ref_value, int64_t *nnz_out,
                        int64_t **offsets_out, int64_t **cols_out, double **vals_out,
                        double **rhs_out, double *detail, char *message, int cap);

/* ============ caller-shaped systems: SIMPLE pressure correction, momentum ============ */
/* SimpleError / MomentumError kinds (src/simple.rs, src/momentum.rs) */
enum {
    OR_SIMPLE_INVALID_REFERENCE = 50,
    OR_SIMPLE_INVALID_FACE_COEF = 51,
    OR_SIMPLE_DEGENERATE = 52,
    OR_SIMPLE_LINEAR = 53,
    OR_MOM_INVALID_RELAX = 60,
    OR_MOM_INVALID_DT = 61,
    OR_MOM_NONFINITE_SOURCE = 62,
    OR_MOM_NONFINITE_FLUX = 63,
    OR_MOM_DEGENERATE = 64,
    OR_MOM_GEOMETRY = 65,
    OR_MOM_MISSING_BOUNDARY = 66,
    OR_MOM_BACKFLOW = 67,
    OR_MOM_MISSING_RELAX_FIELD = 68,
    OR_MOM_LINEAR = 69,
    OR_MOM_FACE_DIFFUSIVITY = 70,
};
/* interior-face interpolation weights (0 on boundary faces) */
int or_face_weights(const or_mesh2d *m, double *lambda, char *message, int cap);
/* pressure_face_coefficients: d_f of every face (0 on boundary faces) */
int or_pressure_face_coefficients(const or_mesh2d *m, const double *r_au, double *coef,
                                  char *message, int cap);
/* assemble_pressure_correction: CSR + rhs (caller frees with or_free) */
int or_assemble_pressure_correction(const or_mesh2d *m, const double *flux, const double *coef,
                                    int64_t ref_cell, int64_t *nnz_out, int64_t **off_out,
                                    int64_t **col_out, double **val_out, double **rhs_out,
                                    char *message, int cap);
/* assemble_momentum_component with Diffusivity::Constant(nu), per-face
 * boundary conditions (bc_kind/bc_value by face, OR_SCALAR_*), optional
 * source, optional transient (old != NULL with dt) and equation
 * under-relaxation alpha (relax_field needed when alpha < 1).  Returns the
 * relaxed CSR, rhs and the unrelaxed diagonal (caller frees with or_free). */
int or_assemble_momentum(const or_mesh2d *m, double nu, const double *flux, const int32_t *bc_kind,
                         const double *bc_value, const double *source, const double *old, double dt,
                         const double *relax_field, double alpha, int64_t *nnz_out, int64_t **off_out,
                         int64_t **col_out, double **val_out, double **rhs_out, double **diag_out,
                         char *message, int cap);

/* ============ collocated unstructured SIMPLE (simple2d.c) ============ */
enum {
    OR_SIMPLE_INVALID_MOM_DIAG = 54,
    OR_SIMPLE_INVALID_RELAX = 55,
    OR_SIMPLE_INVALID_OUTER = 56,
    OR_SIMPLE_MOM_NOT_CONVERGED = 57,
    OR_SIMPLE_P_NOT_CONVERGED = 58,
    /* NumericsError kinds of the least-squares stencil */
    OR_NUM_SINGULAR_LSQ = 80,
    OR_NUM_DEGENERATE_ON = 81,
    OR_NUM_DEGENERATE_OF = 82,
    OR_NUM_MISSING_BOUNDARY = 83,
};
enum { OR_LSQ_SKIP = 0, OR_LSQ_NEIGHBOUR = 1, OR_LSQ_FIXED = 2 };
/* LeastSquaresGradientStencil for 2D meshes: one equation slot per cell face
 * (the mesh's cell_faces order); ZeroGradient faces are skipped. */
typedef struct {
    int64_t ncells;
    int32_t *kind;       /* OR_LSQ_* */
    int64_t *nb;         /* the neighbouring cell (OR_LSQ_NEIGHBOUR) */
    double *val;         /* the boundary value (OR_LSQ_FIXED) */
    double *rcx, *rcy;   /* rhs coefficient offset / |offset|^2 */
    double *inv;         /* per cell: the inverse 2x2 normal matrix, row-major */
} or_lsq2d;
int or_lsq2d_build(const or_mesh2d *m, const int32_t *bc_kind, const double *bc_value,
                   or_lsq2d *st, int64_t *bad, char *message, int cap);
void or_lsq2d_free(or_lsq2d *st);
void or_lsq2d_gradient(const or_mesh2d *m, const or_lsq2d *st, const double *values, double *gx,
                       double *gy);
int or_initial_face_flux(const or_mesh2d *m, const double *ux, const double *uy,
                         const double *wall_flux, double *flux, int64_t *bad, char *message, int cap);
double or_continuity_rms(const or_mesh2d *m, const double *flux);
int or_rhie_chow_flux(const or_mesh2d *m, const or_lsq2d *pst, const double *usx, const double *usy,
                      const double *p, const double *r_au, const double *wall_flux, double *flux,
                      int64_t *bad, double *badv, char *message, int cap);
int or_correct_face_flux(const or_mesh2d *m, const double *predicted, const double *coef,
                         const double *pc, const int32_t *corr_kind, const double *corr_value,
                         double *out, int64_t *bad, double *badv, char *message, int cap);
int or_correct_cell_velocity(int64_t nc, const double *usx, const double *usy, const double *r_au,
                             const double *gx, const double *gy, double *ux, double *uy, int64_t *bad,
                             double *badv, char *message, int cap);
/* SimpleOptions (simple.rs:93-131) with a constant viscosity; the velocity
 * boundary is per boundary face (OR_SCALAR_* kind, x and y values) and the
 * pressure-correction boundary, when given (corr_kind != NULL), anchors the
 * correction instead of reference_cell. */
typedef struct {
    double viscosity;
    int64_t reference_cell;
    int64_t max_outer_iterations;
    double continuity_tol;
    double mom_atol, mom_rtol;
    int64_t mom_max;
    double p_atol, p_rtol;
    int64_t p_max;
    double alpha_u, alpha_p;
} or_simple_opts;
/* SimpleReport (simple.rs:133-142); the z momentum report is 0 iterations */
typedef struct {
    int32_t converged, pad;
    int64_t outer_iterations;
    double initial_rms, final_rms;
    int64_t mom_iters[3];
    int32_t mom_converged[3], p_converged;
    double mom_final[3];
    int64_t p_iters;
    double p_final;
    int64_t err_index;
    double err_value;
} or_simple_report;
/* solve_simple: advances (ux, uy, p, flux) in place; history[0 .. iterations]
 * receives the continuity RMS (capacity max_outer_iterations + 1). */
int or_simple2d_solve(const or_mesh2d *m, const or_lsq2d *pst, const int32_t *vel_kind,
                      const double *vel_x, const double *vel_y, const int32_t *corr_kind,
                      const double *corr_value, const or_simple_opts *o, double *ux, double *uy,
                      double *p, double *flux, double *history, or_simple_report *rep,
                      char *message, int cap);

/* ============ unstructured triangle-mesh Poisson (config-3 producer) ============ */
/* Builds the jittered structured triangulation of the unit square with m x m
 * squares (2*m*m triangles), assembles the two-point Poisson matrix with
 * ZeroGradient boundaries and Gamma=1, pins `ref_cell` to 0 and returns CSR
 * arrays (caller frees with or_free).  See mesh.c for the cited reference lines. */
int or_mesh_poisson(int64_t m, double jitter, uint64_t seed, int64_t ref_cell,
                    int64_t *ncells_out, int64_t *nnz_out, int64_t **offsets,
                    int64_t **cols, double **vals, double **volumes, int64_t *nfaces_out,
                    int64_t **face_owner, int64_t **face_neigh, double **face_coef,
                    char *message, int cap);
void or_free(void *ptr);

/* Deterministic PRNG shared with the GPU-side input generators. */
uint64_t or_splitmix64(uint64_t *state);
double or_uniform(uint64_t *state);  /* [0,1) with 53 random bits */
double or_normal(uint64_t *state);   /* Box-Muller, consumes two uniforms */

#ifdef __cplusplus
}
#endif
#endif
