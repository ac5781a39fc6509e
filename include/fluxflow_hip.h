/*
 * fluxflow_hip.h -- C-ABI of the MI355X (gfx950) pressure-projection library.
 *
 * This is the drop-in boundary: a host binding (N-API addon, ctypes, cgo ...)
 * maps the reference's JavaScript factories onto these entry points one for
 * one.  Every declaration cites the reference interface it replaces
 * (paths relative to the reference repository).
 *
 * Conventions, kept from the reference:
 *  - fields are first-index-fastest: element (i,j,k) of a field of shape
 *    [n0,n1,n2] lives at i + n0*(j + n1*k)   (packages/tsl_array_n/src/array.js:24-38)
 *  - MAC velocity: U [nx+1,ny,nz], V [nx,ny+1,nz], W [nx,ny,nz+1]; the face
 *    array of an axis at index I is the LOWER face of cell I
 *    (packages/fluxflow/src/grid/grid_data3.js:121-176)
 *  - the library owns every field it creates; objects bind the fields they
 *    are created with ("create once, solve many", linalg.js:42-53)
 *  - every call returns FLUX_OK (0) or an error code; flux_last_error() gives
 *    the message.  Construction errors keep the reference's messages.
 *    Numerical failure never returns an error: it is reported through the
 *    state/diagnostics structs (linalg.js:1056-1103).
 *  - all work is ordered on one HIP stream owned by the library; calls are
 *    blocking unless documented otherwise, and must not interleave.
 */
#ifndef FLUXFLOW_HIP_H
#define FLUXFLOW_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FLUX_OK 0
#define FLUX_ERR_ARG 1     /* invalid argument / shape / type (reference: throw) */
#define FLUX_ERR_HIP 2     /* HIP runtime error */
#define FLUX_ERR_STATE 3   /* not initialised, wrong object state */
#define FLUX_ERR_NOMEM 4

/* Last error message of the calling thread ("" when none). */
const char* flux_last_error(void);

/* ---------------------------------------------------------------- context */
/* replaces tsl_array_n init() / getRenderer()  (packages/tsl_array_n/src/context.js:74-144):
 * selects the device and creates the library's stream. */
int flux_init(int device);
/* releases every library object still alive and the stream */
int flux_shutdown(void);
/* waits for all queued work (the reference's awaited readbacks) */
int flux_sync(void);
/* makes the library's device current on the calling host thread, for
 * bindings that run library calls on worker threads (the N-API addon's
 * async work; the reference has one JS thread, context.js:74-144) */
int flux_thread_bind(void);
/* device name and compute-unit count */
int flux_device_info(char* name, size_t name_len, int* compute_units);

/* ----------------------------------------------------------------- fields */
/* element types: arrayN(type, shape)  (packages/tsl_array_n/src/array.js:94-133).
 * 'float' -> FLUX_F32, 'int' -> FLUX_I32, 'vec3' -> FLUX_VEC3 (3 x f32 per
 * element); FLUX_U8 holds masks, FLUX_F64/F16/BF16 are internal storage. */
enum {
    FLUX_F32 = 0,
    FLUX_I32 = 1,
    FLUX_U8 = 2,
    FLUX_F64 = 3,
    FLUX_F16 = 4,
    FLUX_BF16 = 5,
    FLUX_VEC3 = 6
};
typedef struct flux_field flux_field;
/* arrayN(type, shape): ndim 1..3 */
int flux_field_create(int dtype, int ndim, const int64_t* shape, flux_field** out);
int flux_field_destroy(flux_field* f);
/* fromArray(typed)  (array.js:124-129); bytes must equal the field's size */
int flux_field_upload(flux_field* f, const void* src, size_t bytes);
/* toArray()  (array.js:114-122) */
int flux_field_download(const flux_field* f, void* dst, size_t bytes);
/* fills every element with value (converted to the element type) */
int flux_field_fill(flux_field* f, double value);
/* device-to-device copy of equal-sized fields (createCopyKernel3) */
int flux_field_copy(flux_field* dst, const flux_field* src);
/* .shape / .count / .type  (array.js:94-112) */
int flux_field_info(const flux_field* f, int* dtype, int* ndim, int64_t* shape, int64_t* count);
/* raw device address, for bindings that share device memory */
void* flux_field_data(const flux_field* f);

/* ------------------------------------------------------ Laplacian operator */
/* createLaplacianOperator(shape, gridSpacing, {dirichletMask, faceWeights})
 * (packages/fluxflow/src/linalg/multigrid.js:457-471, stencil laplacianAt
 * 273-339).  ndim 2 or 3; mask: FLUX_U8 cells (1 = Dirichlet) or NULL;
 * face_weights: ndim FLUX_F32 face fields or NULL. */
typedef struct flux_operator flux_operator;
int flux_operator_create(int ndim, const int* shape, const double* spacing, const flux_field* mask,
                         const flux_field* const* face_weights, flux_operator** out);
/* applyLaplacian(input, output) */
int flux_operator_apply(flux_operator* op, const flux_field* in, flux_field* out);
/* the per-cell diagonal laplacianDiagonalAt (multigrid.js:399-437) */
int flux_operator_diagonal(flux_operator* op, flux_field* out);
int flux_operator_destroy(flux_operator* op);

/* --------------------------------------------------------------- multigrid */
/* V-cycle storage precision (no reference counterpart: the reference is FP32) */
enum { FLUX_STORE_F32 = 0, FLUX_STORE_F16 = 1, FLUX_STORE_BF16 = 2 };
/* createMultigridPreconditioner options (multigrid.js:1223-1250) */
typedef struct {
    int levels;        /* numberOfLevels, default 1 */
    int nu_down;       /* numberOfSmoothingIterationsDown, default 2 */
    int nu_up;         /* numberOfSmoothingIterationsUp, default 2 */
    int n_coarsest;    /* numberOfCoarsestIterations, default 20 */
    int n_final;       /* numberOfFinalIterations, default 2 */
    double omega;      /* sorFactor, default 1.0 */
    int storage;       /* FLUX_STORE_*, default FLUX_STORE_F32 */
    int smoother;      /* FLUX_SMOOTH_*, default red-black Gauss-Seidel */
    double jacobi_omega; /* damped-Jacobi factor (multigrid_jacobi.js:1106), default 2/3 */
} flux_mg_options;
/* Smoother: red-black Gauss-Seidel, colour pairs in place (buildRelaxKernel,
 * multigrid.js:626-640), or damped Jacobi, single sweeps ping-ponging through
 * a second buffer (buildJacobiRelaxKernel and relax(),
 * sandbox/jacobi-smoother/multigrid_jacobi.js:632-642, 1271-1285); the
 * Jacobi form requires even sweep counts ("sweep counts must be even").
 * FLUX_SMOOTH_BRICK (no counterpart in the reference): the same red-black
 * relaxation run brick by brick (64 x 8 x 8 cells, values outside a brick
 * held at their values on entry), all of a smoothing's sweeps in one pass
 * over memory; two passes per smoothing, the second on bricks shifted by
 * half a brick, post-smoothing the adjoint (shifted first, colours
 * reversed).  3-D levels above 4096 cells except the coarsest; other levels
 * smooth with red-black Gauss-Seidel. */
enum { FLUX_SMOOTH_RBGS = 0, FLUX_SMOOTH_JACOBI = 1, FLUX_SMOOTH_BRICK = 2 };
/* fills the reference defaults */
void flux_mg_options_default(flux_mg_options* o);
typedef struct flux_multigrid flux_multigrid;
/* createMultigridPreconditioner(shape, gridSpacing, {dirichletMask,
 * coarseDirichletMask, faceWeights, ...}) -- throws the reference's
 * "1D/2D/3D", "gridSpacing length" and "not divisible" errors as FLUX_ERR_ARG. */
int flux_multigrid_create(int ndim, const int* shape, const double* spacing, const flux_mg_options* opts,
                          const flux_field* mask, const flux_field* coarse_mask,
                          const flux_field* const* face_weights, flux_multigrid** out);
/* refreshCoarseLevels() (multigrid.js:1415-1451): re-coarsen masks and face
 * weights after the caller changed the fine ones */
int flux_multigrid_refresh(flux_multigrid* mg);
/* applyMultigridPreconditioner(input, output): one V-cycle from output = 0 */
int flux_multigrid_apply(flux_multigrid* mg, const flux_field* in, flux_field* out);
/* copies level `level`'s coarse mask (u8) and face weights (f32, per axis)
 * to host buffers; any pointer may be NULL.  Shape in shape_out[3]. */
int flux_multigrid_level(flux_multigrid* mg, int level, int* shape_out, uint8_t* mask_out,
                         float* const* weights_out);
int flux_multigrid_destroy(flux_multigrid* mg);

/* --------------------------------------------------------------------- PCG */
enum { FLUX_PREC_MULTIGRID = 0, FLUX_PREC_JACOBI = 1, FLUX_PREC_NONE = 2 };
/* stoppedBy codes (linalg.js:1050-1103, 'residual-growing' 2076) */
enum {
    FLUX_STOP_NONE = 0,
    FLUX_STOP_DEGENERATE_PAP = 1,
    FLUX_STOP_PAP_GROWTH = 2,
    FLUX_STOP_ALPHA_MAGNITUDE = 3,
    FLUX_STOP_DEGENERATE_OLD_RZ = 4,
    FLUX_STOP_CONVERGED = 5,
    FLUX_STOP_RESIDUAL_GROWING = 6
};
/* the solver's `state` (linalg.js:879) plus the FP64 acceptance check */
typedef struct {
    int iterations;
    int stopped_by;          /* FLUX_STOP_* */
    int converged;
    double residual;         /* ||b - A x||_2 of the last true residual (FP64) */
    double threshold;        /* max(tol*||b||, tol) when relative, else tol */
    double b_norm;
    int restarts;            /* beta reset to 0 on a sign flip or |beta| > 10 (linalg.js:2436-2442) */
} flux_pcg_state;
typedef struct flux_pcg flux_pcg;
/* createPreconditionedConjugateGradientSolver(applyOperator, applyPreconditioner, b, x)
 * (linalg.js:782-2469).  The preconditioner is picked by `precond`; mg is
 * required for FLUX_PREC_MULTIGRID.  b and x are FLUX_F32 fields of the
 * operator's shape.  x is the warm start and receives the solution. */
int flux_pcg_create(flux_operator* op, int precond, flux_multigrid* mg, flux_field* b, flux_field* x,
                    flux_pcg** out);
/* solve(tol, maxiter, ..., relativeTolerance)  (linalg.js:2176): the
 * solution is accumulated in FP64, every iteration recomputes the true
 * residual b - A x in FP64 arithmetic, dots are FP64, all scalars stay on
 * the GPU; the host reads one scalar block per chunk of iterations. */
int flux_pcg_solve(flux_pcg* s, double tol, int maxiter, int relative, flux_pcg_state* state);
/* No counterpart in the reference (whose beta is r'z'/rz, linalg.js:1198-1226):
 * the flexible (Polak-Ribiere) beta r'.(z' - z)/rz for an inexact
 * preconditioner.  Off by default. */
int flux_pcg_set_flexible(flux_pcg* s, int flexible);
/* No counterpart in the reference: an additive coarse-space correction of
 * the preconditioner, M^-1 r = V(r) + Phi (Phi^T A Phi)^-1 Phi^T r, where
 * the columns of Phi are the free-cell indicator (free_indicator != 0; the
 * near-null mode of a system pinned at one cell) and up to three FLUX_F32
 * fields of the operator's shape (such as a learned guess), each restricted
 * to the free cells.  Phi^T A Phi is assembled in FP64 at every solve's
 * setup and inverted on the host (degenerate directions dropped); Phi^T r
 * is fused into the residual update and Phi a into the direction update, so
 * an iteration reads each field twice more.  The fields stay bound; pass
 * (0, 0, NULL) to remove the correction. */
int flux_pcg_set_coarse_space(flux_pcg* s, int free_indicator, int nfields, const flux_field* const* fields);
/* No counterpart in the reference (support for an external initial guess,
 * such as the learned one): x <- x + alpha g on the free cells, with
 * alpha = g.(b - A x) / g.(A g) in FP64, the step along g that minimises the
 * energy-norm error, so a poor direction is scaled toward zero instead of
 * adding error.  alpha (0 when g.(A g) is not negative) is returned. */
int flux_pcg_scaled_start(flux_pcg* s, const flux_field* g, double* alpha);
int flux_pcg_destroy(flux_pcg* s);

/* ------------------------------------------------------------ dot reducer */
/* createDotReducer(shape, fieldA, fieldB) (linalg.js:374-512): lane i of a
 * 2-D/3-D shape [n0, ...] sums every element whose first index is i (n0
 * lanes); a 1-D shape of n elements has min(n, 64) lanes striding by the
 * lane count.  The lane sums and their total are FP64 and deterministic (the
 * reference sums lanes in FP32 and the lanes in FP64 on the host).  Both
 * fields are FLUX_F32 with `shape`'s element count; they are bound for the
 * reducer's lifetime. */
typedef struct flux_dot flux_dot;
int flux_dot_create(int ndim, const int* shape, const flux_field* a, const flux_field* b, flux_dot** out);
/* `dispatch` (linalg.js:506): computes the lanes into the `partial` field
 * (FP32 per lane) without reading anything back */
int flux_dot_dispatch(flux_dot* d);
/* read() (linalg.js:490-502): dispatch, then the total */
int flux_dot_read(flux_dot* d, double* total);
/* .lanes and .partial (linalg.js:506) */
int flux_dot_info(const flux_dot* d, int* lanes, flux_field** partial);
int flux_dot_destroy(flux_dot* d);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the reducer's sums for a, b of `shape` without a
 * reducer object and without any allocation.  The per-block partials live in
 * a static device array of 2 MiB (each GPU the library runs on holds its own
 * until the process ends), so a shape may have at most 131071 lanes (the
 * first axis of a 2-D/3-D shape); larger shapes are rejected with
 * FLUX_ERR_ARG.  Writes the lane sums (FP32, `lanes` elements; may be NULL)
 * and the FP64 total (a FLUX_F64 field, element 0) -- the same values, bit
 * for bit, as flux_dot_read and the `partial` field.  Asynchronous on the
 * library stream (flux_sync waits).  Every call on one GPU shares its
 * scratch: calls that overlap on one GPU (from several host threads) race. */
int flux_dot_noalloc(int ndim, const int* shape, const flux_field* a, const flux_field* b, flux_field* lanes_out,
                     flux_field* total_out);

/* ------------------------------------------- plain conjugate gradient */
/* createConjugateGradientSolver(applyOperator, b, x) (linalg.js:563-728):
 * A is the Laplacian `op`, or, with op NULL, the diagonal operator
 * out = diag * in (the operator of examples/04-conjugate-gradient).  b and x
 * are FLUX_F32 fields of one shape; x is the initial guess and the result. */
typedef struct flux_cg flux_cg;
int flux_cg_create(flux_operator* op, const flux_field* diag, flux_field* b, flux_field* x, flux_cg** out);
/* solve(tol, maxiter) (linalg.js:641-724): absolute tolerance on
 * sqrt(r.r), the residual recomputed from x every iteration after the first,
 * the reference's guards; *converged is its final check r.r < tol. */
int flux_cg_solve(flux_cg* s, double tol, int maxiter, int* converged, int* iterations);
/* the solver's p, r, Ap, Ax fields (linalg.js:727) */
int flux_cg_fields(flux_cg* s, flux_field** p, flux_field** r, flux_field** ap, flux_field** ax);
int flux_cg_destroy(flux_cg* s);

/* ---------------------------------------------------- pressure projection */
/* createGridPressureSolver3 options (grid_pressure_solver3.js:66-86).
 * dirichlet(pos) of the reference is a shader function of position; here
 * the caller passes its evaluation per cell: active (u8) and target (f32). */
typedef struct {
    int resolution[3];
    double spacing[3];
    double origin[3];
    const flux_field* dirichlet_active;    /* FLUX_U8 cells or NULL */
    const flux_field* dirichlet_target;    /* FLUX_F32 cells or NULL (target 0) */
    const flux_field* face_weights[3];     /* density beta per face, or NULLs */
    const flux_field* collider_weights[3]; /* open fraction per face, or NULLs */
    flux_mg_options multigrid;
    int preconditioner;                    /* FLUX_PREC_*, default multigrid */
    double tolerance;                      /* default 1e-6 */
    int max_iterations;                    /* <= 0: ceil(factor * max(res)) */
    double max_iterations_factor;          /* default 1.5 */
    int relative_tolerance;                /* default 1 */
    double max_plausible_pressure;         /* default 1e6 */
} flux_pressure3_desc;
void flux_pressure3_desc_default(flux_pressure3_desc* d);
/* diagnostics (grid_pressure_solver3.js:287, 526-551) */
typedef struct {
    int converged;
    int rejected;
    int iterations;
    int stopped_by;
    double residual;
    double b_norm;
    int64_t bad_cells;
} flux_pressure_diag;
typedef struct flux_pressure3 flux_pressure3;
int flux_pressure3_create(const flux_pressure3_desc* d, flux_pressure3** out);
/* project(input, output) + dispatch() (grid_pressure_solver3.js:339-558).
 * The six velocity fields are FLUX_F32 MAC arrays; output may alias input. */
int flux_pressure3_project(flux_pressure3* p, const flux_field* u_in, const flux_field* v_in,
                           const flux_field* w_in, flux_field* u_out, flux_field* v_out,
                           flux_field* w_out, flux_pressure_diag* diag);
/* the two halves of project(): buildSystemBatch + refreshPreconditionerMask
 * (grid_pressure_solver3.js:496-512), then the solve from the current
 * pressure (the warm start), circuit breaker and correction (513-556).  A
 * caller may write an initial guess into the pressure field in between. */
int flux_pressure3_assemble(flux_pressure3* p, const flux_field* u_in, const flux_field* v_in, const flux_field* w_in);
int flux_pressure3_solve(flux_pressure3* p, const flux_field* u_in, const flux_field* v_in, const flux_field* w_in,
                         flux_field* u_out, flux_field* v_out, flux_field* w_out, flux_pressure_diag* diag);
/* flux_pcg_scaled_start on the assembled system (between assemble and
 * solve); no counterpart in the reference. */
int flux_pressure3_scaled_guess(flux_pressure3* p, const flux_field* guess, double* alpha);
/* flux_pcg_set_coarse_space / flux_pcg_set_flexible on the projection's
 * solver; no counterpart in the reference */
int flux_pressure3_set_coarse_space(flux_pressure3* p, int free_indicator, int nfields,
                                    const flux_field* const* fields);
int flux_pressure3_set_flexible(flux_pressure3* p, int flexible);
/* .pressure.data, .b, .dirichletMask, .ventMask (grid_pressure_solver3.js:572) */
flux_field* flux_pressure3_field(flux_pressure3* p, int which);
enum { FLUX_P3_PRESSURE = 0, FLUX_P3_RHS = 1, FLUX_P3_MASK = 2, FLUX_P3_VENT = 3, FLUX_P3_TARGET = 4,
       FLUX_P3_WEIGHT_U = 5, FLUX_P3_WEIGHT_V = 6, FLUX_P3_WEIGHT_W = 7 };
int flux_pressure3_destroy(flux_pressure3* p);

/* ------------------------------------------------- particles and transfers */
typedef struct {
    int n[3];          /* cells */
    float h[3];        /* spacing */
    float origin[3];
} flux_grid3;

/* cellIndexOf + a stable sort by cell (grid_flip_solver3.js:227-235): writes
 * cell[p], perm[rank] = particle id (stable), cell_start[c] (n_cells+1
 * entries).  Particles must lie inside the domain; the count of those that
 * do not is returned in *outside (they are binned to the clamped cell). */
int flux_bin_particles(const flux_grid3* g, const flux_field* positions, int64_t count,
                       flux_field* cell, flux_field* perm, flux_field* cell_start, int64_t* outside);
/* P2G of one component by a deterministic gather over cell-sorted particles
 * (the float counterpart of buildScatter + buildFinalize,
 * grid_flip_solver3.js:508-564): out/valid are the component's face fields. */
int flux_p2g_gather(const flux_grid3* g, const flux_field* positions, const flux_field* velocities,
                    const flux_field* perm, const flux_field* cell_start, int component, float eps,
                    flux_field* out, flux_field* valid);
/* P2G of one component with the reference's int32 fixed-point atomics
 * (scale p2gAtomicScale, default 65536): bit-identical to the reference's
 * arithmetic.  numer/denom are FLUX_I32 face scratch fields (cleared here). */
int flux_p2g_fixed(const flux_grid3* g, const flux_field* positions, const flux_field* velocities,
                   int64_t count, int component, float scale, float eps, flux_field* numer,
                   flux_field* denom, flux_field* out, flux_field* valid);
/* g2pUpdate (grid_flip_solver3.js:903-922) */
int flux_g2p(const flux_grid3* g, const flux_field* positions, flux_field* velocities, int64_t count,
             const flux_field* const* vel_new, const flux_field* const* vel_old, float flip_ratio,
             float damping);
/* advectParticles (grid_flip_solver3.js:933-940), no collider */
int flux_advect_particles(const flux_grid3* g, flux_field* positions, int64_t count,
                          const flux_field* const* vel, float dt, int max_substeps);
/* the same with the FLIP solver's static collider (advectionSolver.trace
 * with `collider`, grid_flip_solver3.js:927-940; advection_solver3.js:20-62):
 * a sub-step that ends inside the collider (its SDF, cell-centred on the
 * grid, sampled <= 0) stops at the SDF-weighted point between its ends.
 * collider_sdf may be NULL (then identical to flux_advect_particles). */
int flux_advect_particles_collider(const flux_grid3* g, flux_field* positions, int64_t count,
                                   const flux_field* const* vel, const flux_field* collider_sdf, float dt,
                                   int max_substeps);
/* advectFaceCentered3, order 1, no collider (advection_solver3.js:130-176) */
int flux_advect_faces(const flux_grid3* g, const flux_field* const* vel_in, flux_field* const* vel_out,
                      float dt, int max_substeps);
/* createSemiLagrangianAdvectionSolver3({velocityGrid, dt, maxSubsteps, order})
 * (advection_solver3.js:67-362), no collider: the velocity fields are bound
 * at creation (their contents are read at each call); order 1 is the
 * monotone tricubic semi-Lagrangian step, order 2 MacCormack with the
 * min/max clamp over the forward trace's trilinear corners (90-128).  Order 2
 * owns 12 bytes of device scratch per face. */
typedef struct flux_advection3 flux_advection3;
int flux_advection3_create(const flux_grid3* g, const flux_field* const* velocity, float dt, int max_substeps,
                           int order, flux_advection3** out);
/* the time step of later calls (dt as a live node) */
int flux_advection3_set_dt(flux_advection3* a, float dt);
/* the options' `collider` (advection_solver3.js:67-82): a static collider's
 * SDF cell-centred on the velocity grid (NULL removes it).  The back-trace
 * then stops at the collider (advection_solver3.js:41-52), sample points
 * inside it keep the output's value, and order 2's clamp ranges over the
 * corners outside it (95-128).  The field is read at every call. */
int flux_advection3_set_collider(flux_advection3* a, const flux_field* collider_sdf);
/* advectFaceCentered3(input, output) (130-282); output must alias neither the
 * input nor the velocity */
int flux_advection3_faces(flux_advection3* a, const flux_field* const* in, flux_field* const* out);
/* advectScalar3(input, output) (285-357): cell-centred fields */
int flux_advection3_scalar(flux_advection3* a, const flux_field* in, flux_field* out);
int flux_advection3_destroy(flux_advection3* a);
/* liquid fraction per cell from sorted particles (count > 0 -> 1), then
 * two-phase face beta = rho_ref / max(rho_face, 1e-4)
 * (grid_two_phase_flip_solver2.js:693-762, in 3-D) */
int flux_two_phase_beta(const flux_grid3* g, const flux_field* cell_start, float rho_gas, float rho_liquid,
                        flux_field* fraction, flux_field* const* beta);
/* MAC divergence at cell centres (grid_math3.js:397-410) */
int flux_divergence(const flux_grid3* g, const flux_field* const* vel, flux_field* out);

/* ------------------------------------------------ FLIP grid stages (3-D) */
/* createExtrapolateToRegion3 (array_utils3.js:88-228): `iterations` Jacobi
 * steps in which an invalid cell takes the mean of its valid 6-neighbours
 * (summed +x,-x,+y,-y,+z,-z) and becomes valid.  value/out: FLUX_F32 of one
 * shape (out may be value); valid: FLUX_I32, FLUX_U8 or FLUX_F32 (nonzero =
 * valid).  The object owns the ping-pong scratch (create once, run many). */
typedef struct flux_extrapolator flux_extrapolator;
int flux_extrapolator3_create(const flux_field* value, const flux_field* valid, flux_field* out,
                              flux_extrapolator** e);
int flux_extrapolator3_run(flux_extrapolator* e, int iterations);
int flux_extrapolator3_destroy(flux_extrapolator* e);
/* clearFluidMask + markFluidCellsKernel (grid_flip_solver3.js:216-245):
 * fluid_mask (FLUX_F32 cells) = 1 in every cell holding a particle, else 0 */
int flux_mark_fluid_cells(const flux_grid3* g, const flux_field* positions, int64_t count, flux_field* fluid_mask);
/* computeU/V/WFluidAdjacent (grid_flip_solver3.js:595-625): out (FLUX_I32,
 * the axis's face shape) = 1 where either neighbouring cell (clamped) is fluid */
int flux_fluid_adjacent(const flux_grid3* g, const flux_field* fluid_mask, int axis, flux_field* out);

/* Static collider primitives (primitive_sdf3.js:21-70): type 0 sphere
 * {center, radius}; type 1 box {center, half extents, optional rotation
 * about `axis` by `angle` radians when rotated != 0}. */
typedef struct {
    int type;
    int rotated;
    double center[3];
    double radius;
    double half[3];
    double axis[3];
    double angle;
} flux_primitive3;
/* FP64 grid geometry for the collider set-up (the reference computes it on
 * the host in double precision) */
typedef struct {
    int n[3];
    double h[3];
    double origin[3];
} flux_grid3d;
/* computeFaceWeights (sdf_collider3.js:131-262 with fractionInsideCpu,
 * level_set_utils.js:89-161): open area fraction of every U/V/W face,
 * floored at 0.01 when open and 0 when all four corners are on or inside
 * the surface; weights[3] are FLUX_F32 face fields. */
int flux_collider_face_weights(const flux_grid3d* g, const flux_primitive3* shapes, int count, int invert,
                               flux_field* const* weights);
/* createOutflowPressureDirichlet3 over `count` (<= 4) createSDFOutflow3
 * objects combined with the caller's own Dirichlet fields by
 * combineDirichlet3(outflows, user) (sdf_inflow_outflow3.js:44-94),
 * evaluated at the cell centres: active (u8) = inside any outflow (its SDF,
 * cell-centred on this grid, sampled < 0) or the caller's mask; target =
 * the caller's target where the caller's mask is set, else 0.  user_active
 * and user_target may be NULL.  The outputs feed flux_pressure3_desc's
 * dirichlet_active / dirichlet_target. */
int flux_outflow3_dirichlet(const flux_grid3* g, const flux_field* const* outflow_sdf, int count,
                            const flux_field* user_active, const flux_field* user_target, flux_field* active,
                            flux_field* target);
/* addShapes (sdf_collider3.js:71-110): the signed distance at cell centres */
int flux_collider_sdf(const flux_grid3d* g, const flux_primitive3* shapes, int count, int invert, flux_field* sdf);

/* constrainVelocity (grid_blocked_boundary_condition_solver3.js:389-477) on
 * the MAC velocity `vel`: with a static collider (its SDF on the velocity
 * grid's cells, optional; its face weights, optional) the marker/project,
 * extrapolation, no-flux projection with friction and blocked-boundary
 * passes; then the closed domain walls (DIRECTION_* bits: left 1, right 2,
 * down 4, up 8, back 16, front 32; default 63) and the +-100 clamp. */
typedef struct flux_boundary3 flux_boundary3;
int flux_boundary3_create(const flux_grid3* g, flux_field* const* vel, const flux_field* sdf,
                          const flux_field* const* face_weights, float friction, flux_boundary3** out);
int flux_boundary3_set_closed(flux_boundary3* b, int flags);
/* setInflows (grid_blocked_boundary_condition_solver3.js:128-190) with
 * createSDFInflow3 objects (sdf_inflow_outflow3.js:15-24): the inflow's SDF
 * on the velocity grid's cells, its velocity and mode (0 'set', 1 'add');
 * applied after the walls, before the clamp, in the order added (at most 4) */
int flux_boundary3_add_inflow(flux_boundary3* b, const flux_field* sdf, float vx, float vy, float vz, int mode);
int flux_boundary3_constrain(flux_boundary3* b, int extrapolation_depth);
int flux_boundary3_destroy(flux_boundary3* b);

/* ------------------------------------------------------ 2-D Eulerian step */
/* a 2-D MAC grid: U [n0+1, n1] at (o.x, o.y + h.y/2), V [n0, n1+1] at
 * (o.x + h.x/2, o.y), cell data at o + h/2 (createFaceCenteredGrid2,
 * grid_data2.js:161-198) */
typedef struct {
    int n[2];
    float h[2];
    float origin[2];
} flux_grid2;
/* createSDFStaticCollider2 / createSDFInflow2 / createSDFOutflow2 .addPolygons
 * (sdf_collider2.js:82-100, polygon_sdf.js:24-97): sdf (FLUX_F32 cells) =
 * the signed distance (negative inside, even-odd) to the union of polygons,
 * evaluated in FP64 at each cell centre; ring r spans the (x, y) points
 * ring_start[r] .. ring_start[r+1]-1 of xy (host arrays) */
int flux_polygon_sdf2(flux_field* sdf, const flux_grid2* g, const double* xy, const int* ring_start, int rings,
                      int invert);
/* advectFaceCentered2 order 1 (advection_solver2.js:81-127, 263-320): the
 * midpoint back-trace stops at the collider (collider_sdf may be NULL); a
 * face inside the collider keeps the output's value.  Output must not alias
 * the input. */
int flux_advect_faces2(const flux_grid2* g, const flux_field* u, const flux_field* v, const flux_field* collider_sdf,
                       float dt, int max_substeps, flux_field* u_out, flux_field* v_out);
/* createSemiLagrangianAdvectionSolver2({velocityGrid, collider, dt,
 * maxSubsteps, order}) (advection_solver2.js:174-526), face grids: the
 * velocity fields and the collider SDF (may be NULL) are bound at creation;
 * order 1, or 2 = MacCormack with the clamp over the forward trace's bilinear
 * corners that lie outside the collider (221-258).  Order 2 owns 12 bytes of
 * device scratch per face. */
typedef struct flux_advection2 flux_advection2;
int flux_advection2_create(const flux_grid2* g, const flux_field* u, const flux_field* v,
                           const flux_field* collider_sdf, float dt, int max_substeps, int order,
                           flux_advection2** out);
int flux_advection2_set_dt(flux_advection2* a, float dt);
/* advectFaceCentered2(input, output): a face inside the collider keeps the
 * output's value; the output must alias neither the input nor the velocity */
int flux_advection2_faces(flux_advection2* a, const flux_field* in_u, const flux_field* in_v, flux_field* out_u,
                          flux_field* out_v);
int flux_advection2_destroy(flux_advection2* a);
/* createMaxAbsReducer(fields).read() (reduction.js:43-77): max |x| */
int flux_max_abs(const flux_field* const* fields, int count, float* out);
/* computeAdaptiveSubSteps (time/cfl.js:38-48): steps = min(max_sub,
 * max(1, ceil(max_velocity * frame_dt / min_spacing / courant))),
 * sub_dt = frame_dt / steps */
int flux_adaptive_substeps(double max_velocity, double min_spacing, double frame_dt, double courant, int max_sub,
                           int* steps, double* sub_dt);
/* createGridBlockedBoundaryConditionSolver2(velocity, ..., collider, inflows)
 * (grid_blocked_boundary_condition_solver2.js:98-597): bound to the two
 * velocity fields; a static collider given by its SDF (may be NULL) with a
 * friction coefficient; inflows (mode 0 'set', 1 'add') are SDF fields with
 * a velocity.  constrain = markAndProject, extrapolation by `depth` Jacobi
 * passes, no-flux projection with friction, blocked faces, closed walls
 * (DIRECTION_* bits: left 1, right 2, down 4, up 8; default 15), inflows,
 * the +-100 clamp. */
typedef struct flux_boundary2 flux_boundary2;
int flux_boundary2_create(const flux_grid2* g, flux_field* const* vel, const flux_field* collider_sdf, float friction,
                          flux_boundary2** out);
int flux_boundary2_set_closed(flux_boundary2* b, int flags);
int flux_boundary2_add_inflow(flux_boundary2* b, const flux_field* sdf, float vx, float vy, int mode);
int flux_boundary2_constrain(flux_boundary2* b, int extrapolation_depth);
int flux_boundary2_destroy(flux_boundary2* b);
/* createGridOutflowSolver2 (grid_outflow_solver2.js:217-396): apply =
 * applyOutflowVelocityBC (faces inside the outflow relax toward the
 * velocity one cell upstream by the outward Courant number, U before V);
 * dirichlet = createOutflowPressureDirichlet2 + combineDirichlet
 * (sdf_inflow_outflow2.js:152-199) at the cell centres: active (u8) =
 * inside any outflow or the caller's mask, target = the caller's target
 * where the caller's mask is set, else 0 */
typedef struct flux_outflow2 flux_outflow2;
int flux_outflow2_create(const flux_grid2* g, const flux_field* outflow_sdf, flux_outflow2** out);
int flux_outflow2_apply(flux_outflow2* o, flux_field* u, flux_field* v, float dt);
int flux_outflow2_destroy(flux_outflow2* o);
int flux_outflow2_dirichlet(const flux_grid2* g, const flux_field* const* outflow_sdf, int count,
                            const flux_field* user_active, const flux_field* user_target, flux_field* active,
                            flux_field* target);
/* createGridPressureSolver2 (grid_pressure_solver2.js:223-981): the 2-D
 * projection -- no collider weights; optional face weights; the same PCG,
 * multigrid, guards and circuit breaker as the 3-D solver; the iteration
 * cap is ceil(factor * max(n) * 4) (398) */
typedef struct {
    int resolution[2];
    double spacing[2];
    double origin[2];
    const flux_field* dirichlet_active;   /* u8 cells, or NULL */
    const flux_field* dirichlet_target;   /* f32 cells, or NULL */
    const flux_field* face_weights[2];    /* f32 U and V faces, or NULL */
    flux_mg_options multigrid;
    int preconditioner;
    double tolerance;
    int max_iterations;
    double max_iterations_factor;
    int relative_tolerance;
    double max_plausible_pressure;
} flux_pressure2_desc;
void flux_pressure2_desc_default(flux_pressure2_desc* d);
typedef struct flux_pressure2 flux_pressure2;
int flux_pressure2_create(const flux_pressure2_desc* d, flux_pressure2** out);
/* project(input, output) + dispatch() (663-954); output may alias input */
int flux_pressure2_project(flux_pressure2* p, const flux_field* u_in, const flux_field* v_in, flux_field* u_out,
                           flux_field* v_out, flux_pressure_diag* diag);
flux_field* flux_pressure2_field(flux_pressure2* p, int which);
enum { FLUX_P2_PRESSURE = 0, FLUX_P2_RHS = 1 };
int flux_pressure2_destroy(flux_pressure2* p);
/* createGridSolver2 (grid_solver2.js:131-255) without external forces:
 * onAdvanceTimeStep(dt) = project in place, constrain; clone, advect the
 * clone into the velocity (the collider forwarded to the advection only
 * when advect_collider is set, as the reference's options.advection),
 * the outflow velocity condition, constrain.  The solver owns its velocity
 * fields (zero at creation; fill them through flux_solver2_field). */
typedef struct {
    flux_grid2 grid;
    const flux_field* collider_sdf;       /* NULL: no collider */
    float friction;                       /* default 0.5 */
    int advect_collider;                  /* default 0 */
    int max_substeps;                     /* advection, default 32 */
    int n_inflows;
    const flux_field* inflow_sdf[4];
    float inflow_velocity[4][2];
    int inflow_mode[4];                   /* 0 'set', 1 'add' */
    int n_outflows;
    const flux_field* outflow_sdf[4];
    int outflow_velocity_bc;              /* default 1 */
    const flux_field* dirichlet_active;   /* the caller's own Dirichlet cells, or NULL */
    const flux_field* dirichlet_target;
    int closed_flags;                     /* default 15 */
    int extrapolation_depth;              /* default 5 */
    flux_mg_options multigrid;
    int preconditioner;
    double tolerance;
    int max_iterations;
    double max_iterations_factor;
    int relative_tolerance;
    double max_plausible_pressure;
    int advection_order;                  /* options.advection.order: 1 (default) or 2 (MacCormack) */
} flux_solver2_desc;
void flux_solver2_desc_default(flux_solver2_desc* d);
typedef struct flux_solver2 flux_solver2;
int flux_solver2_create(const flux_solver2_desc* d, flux_solver2** out);
int flux_solver2_step(flux_solver2* s, float dt, flux_pressure_diag* diag);
flux_field* flux_solver2_field(flux_solver2* s, int which);
enum { FLUX_S2_U = 0, FLUX_S2_V = 1, FLUX_S2_U_PREV = 2, FLUX_S2_V_PREV = 3, FLUX_S2_PRESSURE = 4,
       FLUX_S2_DIRICHLET = 5 };
int flux_solver2_destroy(flux_solver2* s);

/* createGridFlipSolver3 (grid_flip_solver3.js:135-1155): free-surface FLIP
 * with reduced pressure (air cells Dirichlet at -dt g.x), int32 fixed-point
 * P2G, weight and fluid-adjacent extrapolation, constrainVelocity with an
 * optional static collider, deterministic resampling (donor ranks from
 * prefix sums, 640-793) and the pressure projection.  Particles live in the
 * solver's own vec3 fields (all max_particles of them are simulated); the
 * caller writes the initial positions/velocities through flux_flip3_field. */
typedef struct {
    int resolution[3];
    double spacing[3];
    double origin[3];
    int64_t max_particles;
    float dt;
    float gravity[3];              /* default (0, -9.81, 0) */
    int reduced_pressure;          /* default 1 */
    float flip_ratio;              /* default 0.97 */
    float velocity_damping;        /* default 0.02 */
    float p2g_atomic_scale;        /* default 65536 */
    float weight_epsilon;          /* default 1e-4 */
    int closed_domain_flags;       /* default 63 (all six walls) */
    int resample;                  /* default 1 */
    int min_particles_per_cell;    /* default 4 */
    int max_particles_per_cell;    /* default 8 */
    int extrapolation_depth;       /* default 5 */
    int max_substeps;              /* particle advection, default 32 */
    const flux_field* collider_sdf;        /* static collider SDF on the cells, or NULL */
    const flux_field* collider_weights[3]; /* its face weights for constrainVelocity, or NULLs */
    float friction;                        /* default 0.5 */
    /* the projection's options (multigrid, tolerance, ...); resolution,
     * spacing, origin and the Dirichlet fields are set by the solver, and a
     * max_plausible_pressure <= 0 is derived as 8 dt |g| (sum of extents)
     * (grid_flip_solver3.js:425-432) */
    flux_pressure3_desc pressure;
    /* carried concentration (carryConcentration, mixing, fade,
     * concentrationAtomicScale; grid_flip_solver3.js:967-1049): a per-particle
     * value in [0, 1], averaged per cell by fixed-point atomics each step and
     * blended back by mixing / faded by fade after G2P */
    int carry_concentration;           /* default 0 */
    float mixing;                      /* default 0 */
    float fade;                        /* default 0 */
    float concentration_atomic_scale;  /* default 4096 */
    /* ambientDensity / componentDensity (265-357, 848-880): both > 0 couple
     * the cell concentration to the density, the pressure's face weights
     * rho_ref / rho_face and, with reduced pressure, the gravity term */
    float ambient_density;             /* default 0 (off) */
    float component_density;           /* default 0 (off) */
    int mass_weighted_transfer;        /* massWeightedTransfer (493-531), default 0 */
    /* pressurePin (grid_two_phase_flip_solver2.js:133-169): when the free
     * surface leaves no Dirichlet cell (a closed domain full of fluid), this
     * cell is pinned to 0: 0 none (default), 1 top centre (nx/2, ny-1, nz/2),
     * 2 the cell pin_cell */
    int pressure_pin;
    int pin_cell[3];
    /* phase-aware resampling (grid_two_phase_flip_solver2.js:1131-1503, in
     * 3-D): donors split by phase, a recipient takes its majority phase;
     * needs carry_concentration.  Default 0 */
    int phase_resample;
    /* colliderPushThresh / colliderPushShift (pushOutOfCollider,
     * grid_flip_solver3.js:945-964): with a collider, after the particle
     * advection a particle where the collider's SDF is below the threshold
     * moves along the SDF's normalised gradient by (thresh - sdf + shift).
     * Defaults 0 and 0 */
    float collider_push_thresh;
    float collider_push_shift;
} flux_flip3_desc;
void flux_flip3_desc_default(flux_flip3_desc* d);
typedef struct flux_flip3 flux_flip3;
int flux_flip3_create(const flux_flip3_desc* d, flux_flip3** out);
/* onAdvanceTimeStep(): advect, P2G, extrapolate, constrain, mark fluid,
 * resample, constrain, project, constrain, extrapolate, G2P */
int flux_flip3_step(flux_flip3* s, flux_pressure_diag* diag);
/* fields 10-14 are the last projection's Dirichlet fields and face weights
 * (the density beta; NULL without density coupling); 15-17 its input
 * velocity, kept only while flux_flip3_capture_projection_input is on */
enum { FLUX_FLIP3_POSITIONS = 0, FLUX_FLIP3_VELOCITIES = 1, FLUX_FLIP3_FLUID_MASK = 2, FLUX_FLIP3_U = 3,
       FLUX_FLIP3_V = 4, FLUX_FLIP3_W = 5, FLUX_FLIP3_PRESSURE = 6, FLUX_FLIP3_CONCENTRATION = 7,
       FLUX_FLIP3_CELL_CONCENTRATION = 8, FLUX_FLIP3_CELL_DENSITY = 9, FLUX_FLIP3_DIRICHLET_ACTIVE = 10,
       FLUX_FLIP3_DIRICHLET_TARGET = 11, FLUX_FLIP3_BETA_U = 12, FLUX_FLIP3_BETA_V = 13, FLUX_FLIP3_BETA_W = 14,
       FLUX_FLIP3_PROJECTION_INPUT_U = 15, FLUX_FLIP3_PROJECTION_INPUT_V = 16, FLUX_FLIP3_PROJECTION_INPUT_W = 17 };
/* Keep a copy of each step's projection input (the grid velocity after P2G,
 * gravity and constrainVelocity), so the pressure system a frame solved can
 * be captured with its Dirichlet fields and weights -- the frame capture of
 * the reference's sandbox (sandbox/stalled-system/export_system.mjs) for
 * this solver.  Costs three face-field copies per step while on. */
int flux_flip3_capture_projection_input(flux_flip3* s, int on);
flux_field* flux_flip3_field(flux_flip3* s, int which);
/* The projection solver the step uses (owned by the FLIP solver), so a caller
 * can set its coarse-space correction (flux_pressure3_set_coarse_space) or
 * flexible beta.  No counterpart in the reference, whose FLIP solver builds
 * its pressure solver internally (grid_flip_solver3.js:425-432) and exposes
 * none of these options. */
flux_pressure3* flux_flip3_pressure(flux_flip3* s);
/* The step's dt from the next step on: the reference reads dt as a live node
 * each step (createGridAdaptiveTimeStep3 writes the sub-step dt into it,
 * grid_adaptive_timestep3.js:26-40); the projection's plausible-pressure
 * bound keeps the value derived from the creation dt. */
int flux_flip3_set_dt(flux_flip3* s, float dt);
int flux_flip3_destroy(flux_flip3* s);

/* createGridSmokeSolver3 (grid_smoke_solver3.js:25-154 over createGridSolver3,
 * grid_solver3.js:45-149): buoyancy up * (a rho + b (T - T_amb)) applied as an
 * external force, constrainVelocity (closed walls by flag), projection,
 * semi-Lagrangian velocity advection (order 1 or 2), then density and temperature
 * advected through the new velocity and decayed into the other halves of
 * their ping-pong pairs.  No collider, inflow or outflow objects. */
typedef struct {
    int resolution[3];
    double spacing[3];
    double origin[3];
    float dt;
    float buoyancy_smoke_density_factor;  /* default -0.000625 */
    float buoyancy_temperature_factor;    /* default 5.0 */
    float ambient_temperature;            /* default 0 */
    float smoke_decay;                    /* default 0.001 */
    float temperature_decay;              /* default 0.001 */
    float up[3];                          /* default (0, 1, 0) */
    int closed_domain_flags;              /* default 63 */
    int max_substeps;                     /* default 32 */
    int extrapolation_depth;              /* constrainVelocity's, default 5 */
    flux_pressure3_desc pressure;         /* resolution/spacing/origin set by the solver */
    int advection_order;                  /* options.advection.order: 1 (default) or 2, velocity and scalars */
} flux_smoke3_desc;
void flux_smoke3_desc_default(flux_smoke3_desc* d);
typedef struct flux_smoke3 flux_smoke3;
int flux_smoke3_create(const flux_smoke3_desc* d, flux_smoke3** out);
int flux_smoke3_step(flux_smoke3* s, flux_pressure_diag* diag);
/* a ball source re-injected with max into the current density/temperature
 * (examples/34-smoke-plume-3d/main.js:146-158) */
int flux_smoke3_source_sphere(flux_smoke3* s, const float* center, float radius, float density, float temperature);
/* a per-cell source re-injected with max: state = max(state, source) for the
 * current density and/or temperature (NULL leaves that state alone); the
 * example's createSourceKernel (examples/34-smoke-plume-3d/main.js:146-158)
 * with the source value read from a field, e.g. a seeded noisy ball */
int flux_smoke3_source_field(flux_smoke3* s, const flux_field* density, const flux_field* temperature);
enum { FLUX_SMOKE3_U = 0, FLUX_SMOKE3_V = 1, FLUX_SMOKE3_W = 2, FLUX_SMOKE3_DENSITY = 3, FLUX_SMOKE3_TEMPERATURE = 4,
       FLUX_SMOKE3_PRESSURE = 5 };
/* DENSITY / TEMPERATURE return the current state (density.current) */
flux_field* flux_smoke3_field(flux_smoke3* s, int which);
int flux_smoke3_destroy(flux_smoke3* s);

/* -------------------------------------------------- learned initial guess */
/* No reference counterpart (the reference has no learned component): a 3-D
 * conv U-Net (4 levels, 16/32/64/64 channels, 3x3x3, ~0.69M parameters)
 * predicting the pressure from (b / rms(b), Dirichlet mask, log face weight),
 * trained by flux_unet_train_step on pairs the MGPCG solver produces.
 * Parameters are FP32, flattened layer by layer as w[co][ci][27] then b[co]. */
typedef struct flux_unet flux_unet;
int64_t flux_unet_param_count(void);
/* grid divisible by 8; training != 0 also allocates gradients and Adam state */
int flux_unet_create(int nx, int ny, int nz, int training, flux_unet** out);
int flux_unet_destroy(flux_unet* u);
int flux_unet_set_params(flux_unet* u, const float* host, int64_t count);
/* which = 0: parameters, 1: gradients of the last train step */
int flux_unet_get_params(flux_unet* u, float* host, int64_t count, int which);
/* Arithmetic of the inference entry points (forward, guess, guess_into):
 * FLUX_UNET_F32 direct FP32 convolutions (the default); FLUX_UNET_BF16
 * implicit-GEMM convolutions on the matrix cores with BF16 weights and
 * channels-last BF16 activations, FP32 accumulation and an FP32 output.
 * Training always runs in FP32. */
enum { FLUX_UNET_F32 = 0, FLUX_UNET_BF16 = 1 };
int flux_unet_set_precision(flux_unet* u, int precision);
/* input FLUX_F32 [3][nz][ny][nx] -> out FLUX_F32 [nz][ny][nx] */
int flux_unet_forward(flux_unet* u, const flux_field* input, flux_field* out);
/* forward, MSE against target, backward, and an Adam step when lr > 0 */
int flux_unet_train_step(flux_unet* u, const flux_field* input, const flux_field* target, float lr, float beta1,
                         float beta2, float eps, double* loss);
/* No counterpart in the reference: a training step on the energy-norm loss
 * L = 1 - (g.r)^2 / ((g.B g) e_star) of the guess g = network output on the
 * free cells (mask: the system's Dirichlet mask), B = -A the system's
 * operator `op`, r the residual of the start (b on the free cells when the
 * start is 0 there) and e_star = x*.(-r) > 0 for the exact correction x*:
 * the relative energy error after the solver's optimal step along g.  The
 * gradient uses one application of `op`; Adam when lr > 0.  *loss = L. */
int flux_unet_train_step_energy(flux_unet* u, const flux_field* input, flux_operator* op, const flux_field* r,
                                const flux_field* mask, double e_star, float lr, float beta1, float beta2, float eps,
                                double* loss);
/* input channels from an assembled system (b, Dirichlet mask, face weights) */
int flux_unet_features(flux_unet* u, const flux_field* b, const flux_field* mask, const flux_field* const* weights,
                       float density_ratio, flux_field* input);
/* pressure guess = network output x rms(b) */
int flux_unet_guess(flux_unet* u, const flux_field* input, flux_field* guess);
/* the same guess written into `pressure` at the free cells only (Dirichlet
 * cells keep their targets), ready for flux_pressure3_solve */
int flux_unet_guess_into(flux_unet* u, const flux_field* input, const flux_field* mask, flux_field* pressure);

/* ----------------------------------------------------------------- timing */
/* HIP-event timing of the library stream */
int flux_timer_start(void);
int flux_timer_stop(double* ms);
/* per-kernel event timing: when enabled, every launch of the kernel family
 * `name` ("apply", "update", "relax", "residual", "restrict", "correct") on
 * the finest level is bracketed by events; read returns the summed
 * duration and the launch count, then resets. */
int flux_kprof_enable(const char* name);
int flux_kprof_read(double* total_ms, int64_t* launches);

#ifdef __cplusplus
}
#endif
#endif
