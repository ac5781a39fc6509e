/*
 * CPU oracle for the fluxflow pressure-projection path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, in plain C, the
 * algorithms of the reference's WebGPU/TSL code so that the HIP product path
 * can be checked against it.  Only the test-suite, the smoke check and the
 * CPU-baseline leg of the benchmark may load it, and there only as the
 * checker (or as the CPU baseline being timed), never as a fallback.
 *
 * Layout conventions (packages/tsl_array_n/src/array.js:24-38): every field is
 * first-index-fastest, cell (i,j,k) lives at i + nx*(j + ny*k).  MAC faces:
 *   U [nx+1, ny, nz]   V [nx, ny+1, nz]   W [nx, ny, nz+1]
 * and the face array of axis a at index I is the LOWER face of cell I
 * (packages/fluxflow/src/linalg/multigrid.js:316-324).
 *
 * Every real-valued routine exists twice: suffix _f64 (double storage and
 * arithmetic: the FP64 truth) and _f32 (float storage and arithmetic; dot
 * products accumulate in double, like the GPU path).
 */
#ifndef FLUX_ORACLE_H
#define FLUX_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_LEVELS 10

/* PCG stop codes: packages/fluxflow/src/linalg/linalg.js:1056-1103 */
enum {
    ORC_STOP_NONE = 0,
    ORC_STOP_DEGENERATE_PAP = 1,
    ORC_STOP_PAP_GROWTH = 2,
    ORC_STOP_ALPHA_MAGNITUDE = 3,
    ORC_STOP_DEGENERATE_OLD_RZ = 4,
    ORC_STOP_CONVERGED = 5,
    ORC_STOP_RESIDUAL_GROWING = 6
};

enum { ORC_PREC_MULTIGRID = 0, ORC_PREC_JACOBI = 1, ORC_PREC_NONE = 2 };

typedef struct {
    int ndim;        /* 2 or 3; a 2-D grid has n[2] == 1 */
    int n[3];
    double h[3];
} orc_grid;

typedef struct {
    int levels;          /* numberOfLevels (default 1) */
    int nu_down;         /* numberOfSmoothingIterationsDown (2) */
    int nu_up;           /* numberOfSmoothingIterationsUp (2) */
    int n_coarsest;      /* numberOfCoarsestIterations (20) */
    int n_final;         /* numberOfFinalIterations (2) */
    double omega;        /* sorFactor (1.0); the Jacobi smoother's damping (2/3) */
    int smoother;        /* ORC_SMOOTH_RBGS (multigrid.js), ORC_SMOOTH_JACOBI or ORC_SMOOTH_BRICK
                            (sandbox/jacobi-smoother/multigrid_jacobi.js) */
} orc_mg_options;

/* ORC_SMOOTH_BRICK: red-black sweeps brick by brick, the values outside a
 * brick held at their values on entry (orc_brick_relax; the library's
 * one-pass smoother, no counterpart in the reference) */
enum { ORC_SMOOTH_RBGS = 0, ORC_SMOOTH_JACOBI = 1, ORC_SMOOTH_BRICK = 2 };
#define ORC_BRICK_X 64
#define ORC_BRICK_Y 8
#define ORC_BRICK_Z 8
/* levels of more than this many cells smooth brick-wise */
#define ORC_BRICK_MIN_CELLS 4096

typedef struct {
    int iterations;
    int stopped_by;      /* ORC_STOP_* ; CONVERGED when the threshold was met */
    int converged;
    double residual_sq;  /* r.r of the last true residual */
    double threshold;    /* absolute stop threshold used */
    double b_norm;
    int restarts;        /* beta resets: sign flip or |beta| > 10 (linalg.js:2436-2442) */
} orc_pcg_state;

/* ---------------- deterministic inputs ---------------- */
/* mulberry32: packages/fluxflow/src/grid/jitter_random.js:29-42 */
typedef struct { uint32_t state; } orc_rng;
void orc_rng_seed(orc_rng* r, uint32_t seed);
uint32_t orc_rng_next_u32(orc_rng* r);
double orc_rng_next(orc_rng* r);           /* [0,1) exactly as the JS */
void orc_rng_fill(uint32_t seed, double* out, int64_t n);

/* computeFlipBoxSeed3: packages/fluxflow/src/grid/grid_flip_solver3.js:70-130.
 * Writes count*3 floats (positions); returns the count.  Pass out=NULL to
 * only count. */
int64_t orc_flip_box_seed3(const double box_min[3], const double box_max[3],
                           const double h[3], int ppa, double jitter,
                           uint32_t seed, float* out);

/* cellIndexOf: grid_flip_solver3.js:227-235 (f32, IEEE division). */
void orc_cell_index3(const float* pos, int64_t count, const float origin[3],
                     const float h[3], const int n[3], int32_t* cell_out);

/* Stable counting sort of particles by cell (the deterministic binning the
 * HIP path builds with a radix sort): perm_out[rank] = particle id. */
void orc_sort_by_cell(const int32_t* cell, int64_t count, int64_t n_cells,
                      int32_t* perm_out, int32_t* cell_start_out);

/* P2G, fixed-point int32 mode exactly as buildScatter/buildFinalize
 * (grid_flip_solver3.js:508-564); component = 0,1,2 (U,V,W).  numer/denom are
 * int32 face accumulators (not cleared here). */
void orc_p2g_fixed3(const float* pos, const float* vel, int64_t count, int component,
                    const int face_n[3], const float data_origin[3], const float h[3],
                    float scale, int32_t* numer, int32_t* denom);
void orc_p2g_finalize_fixed(const int32_t* numer, const int32_t* denom, int64_t n_faces,
                            float scale, float eps, float* out, int32_t* valid);
/* P2G, deterministic float mode: gather over cell-sorted particles (perm and
 * cell_start from orc_sort_by_cell), f32 sums in a fixed order, finalized
 * with the same eps rule. */
void orc_p2g_gather3(const float* pos, const float* vel, const int32_t* perm,
                     const int32_t* cell_start, int component, const int n[3],
                     const float data_origin[3], const float h[3], float eps,
                     float* out, int32_t* valid);
/* P2G float mode: the same trilinear weights, summed in double in particle
 * order, then divided (the value the sorted segmented sum must reproduce). */
void orc_p2g_float3(const float* pos, const float* vel, int64_t count, int component,
                    const int face_n[3], const float data_origin[3], const float h[3],
                    double* numer, double* denom);
void orc_p2g_finalize_float(const double* numer, const double* denom, int64_t n_faces,
                            float eps, float* out, int32_t* valid);

/* trilinear MAC sample: grid_math3.js:61-126 */
float orc_sample_trilinear3(const float* data, const int shape[3], const float data_origin[3],
                            const float h[3], const float pos[3]);
/* G2P: grid_flip_solver3.js:903-922 */
void orc_g2p3(const float* pos, float* vel, int64_t count,
              const float* u, const float* v, const float* w,
              const float* u_old, const float* v_old, const float* w_old,
              const int n[3], const float h[3], const float origin[3],
              float flip_ratio, float damping);
/* particle advection, RK2 back-trace with no collider:
 * grid_flip_solver3.js:933-940 and advection_solver3.js:20-62 */
void orc_advect_particles3(float* pos, int64_t count,
                           const float* u, const float* v, const float* w,
                           const int n[3], const float h[3], const float origin[3],
                           float dt, int max_substeps);
/* semi-Lagrangian MAC advection, order 1, no collider:
 * advection_solver3.js:130-176 with grid_math3.js:137-219 */
void orc_advect_face3(const float* u, const float* v, const float* w,
                      float* u_out, float* v_out, float* w_out,
                      const int n[3], const float h[3], const float origin[3],
                      float dt, int max_substeps);
/* two-phase face beta: grid_two_phase_flip_solver2.js:693-762 (3-D form) */
void orc_two_phase_beta3(const float* liquid_fraction, const int n[3],
                         float rho_gas, float rho_liquid,
                         float* beta_u, float* beta_v, float* beta_w);

/* ---------------- real-templated solver (suffix _f64 / _f32) ---------------- */
#define ORC_DECL(real, S)                                                                        \
    void orc_apply_##S(const orc_grid* g, const real* x, real* out, const uint8_t* mask,          \
                       const real* const* w);                                                     \
    void orc_diag_##S(const orc_grid* g, real* out, const uint8_t* mask, const real* const* w);   \
    void orc_relax_##S(const orc_grid* g, real* x, const real* b, const uint8_t* mask,            \
                       const real* const* w, double omega, int color);                            \
    void orc_jacobi_##S(const orc_grid* g, const real* xr, real* xw, const real* b,                \
                        const uint8_t* mask, const real* const* w, double omega);                 \
    void orc_brick_relax_##S(const orc_grid* g, const real* xin, real* xout, const real* b,        \
                             const uint8_t* mask, const real* const* w, double omega, int sweeps, \
                             int reversed);                                                       \
    void orc_residual_##S(const orc_grid* g, const real* x, const real* b, real* r,               \
                          const uint8_t* mask, const real* const* w);                             \
    void orc_restrict_##S(const orc_grid* fine, const real* r_fine, const uint8_t* fine_mask,     \
                          real* b_coarse, const uint8_t* coarse_mask);                            \
    void orc_correct_##S(const orc_grid* fine, const real* x_coarse, real* x_fine,                \
                         const uint8_t* fine_mask);                                               \
    void orc_coarsen_weights_##S(const orc_grid* fine, const real* const* w_fine,                 \
                                 real* const* w_coarse);                                          \
    void orc_coarsen_mask_##S(const orc_grid* fine, const uint8_t* fine_src,                      \
                              const real* const* w_coarse, uint8_t* coarse_mask);                 \
    void* orc_mg_create_##S(const orc_grid* g, const orc_mg_options* o, const uint8_t* mask,      \
                            const uint8_t* coarse_src_mask, const real* const* w);                \
    void orc_mg_destroy_##S(void* mg);                                                            \
    void orc_mg_apply_##S(void* mg, const real* r, real* z);                                      \
    int orc_mg_level_info_##S(void* mg, int level, int* n_out, uint8_t** mask, real** w);         \
    int orc_pcg_solve_##S(const orc_grid* g, const uint8_t* mask, const real* const* w,           \
                          int prec, void* mg, const real* b, real* x, double tol, int maxiter,    \
                          int relative, orc_pcg_state* st);                                       \
    double orc_true_residual_##S(const orc_grid* g, const uint8_t* mask, const real* const* w,    \
                                 const real* b, const real* x);                                   \
    void orc_pressure3_assemble_##S(const orc_grid* g, const uint8_t* active, const real* target, \
                                    const real* const* beta, const real* const* cw,               \
                                    const real* const* vel, uint8_t* mask_out, uint8_t* vent_out, \
                                    real* target_out, real* const* wsys_out, real* b_out,         \
                                    real* pressure);                                              \
    void orc_pressure3_correct_##S(const orc_grid* g, const real* pressure,                       \
                                   const real* const* beta, const real* const* cw,                \
                                   const real* const* vel_in, real* const* vel_out);              \
    int64_t orc_count_bad_##S(const real* p, int64_t n, double max_plausible);                    \
    void orc_divergence3_##S(const orc_grid* g, const real* const* vel, real* out);

ORC_DECL(double, f64)
ORC_DECL(float, f32)

/* the HIP plain path's arithmetic: f32 vectors, double solution accumulator,
 * double true residual, double dots (see oracle_f32.c) */
/* Post-smoothing with the colours in the pre-smoothing order (not reversed):
 * the non-symmetric V-cycle that the reference's captured frame-112 stalled
 * on, a host-side switch of the stop-code tests. */
void orc_set_post_forward(int on);
int orc_post_forward_flag(void);
/* brick smoother shape (cells per brick along each axis) and passes per
 * smoothing (2: the second pass on bricks shifted by half a brick) */
void orc_set_brick(int bx, int by, int bz, int passes);
const int* orc_brick_dims(void);
int orc_brick_passes(void);
void orc_set_flexible(int on);
int orc_pcg_solve_mixed(const orc_grid* g, const uint8_t* mask, const float* const* w, int prec,
                        void* mg, const float* b, float* x, double* x64_out, double tol, int maxiter,
                        int relative, orc_pcg_state* st);

/* threads for the OpenMP arms (1 = the reference's serial model) */
void orc_set_threads(int n);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif

/* ---------------- FLIP grid stages (oracle_flip.c) ---------------- */
/* createExtrapolateToRegion3: array_utils3.js:88-228 (in may equal out) */
void orc_extrapolate3(const int n[3], const float* in, const int32_t* valid, float* out, int iterations);
/* clearFluidMask + markFluidCellsKernel: grid_flip_solver3.js:216-245 */
void orc_mark_fluid3(const float* pos, int64_t count, const float origin[3], const float h[3], const int n[3],
                     float* mask);
/* computeU/V/WFluidAdjacent: grid_flip_solver3.js:595-625 */
void orc_fluid_adjacent3(const float* mask, const int n[3], int axis, int32_t* out);
/* sphere (type 0) / box (type 1, optional rotation about `axis` by `angle`):
 * primitive_sdf3.js:21-70, 149-208 */
typedef struct {
    int type;
    int rotated;
    double center[3];
    double radius;
    double half[3];
    double axis[3];
    double angle;
} orc_primitive3;
double orc_primitives_distance3(const orc_primitive3* shapes, int count, double x, double y, double z);
/* fractionInsideCpu: level_set_utils.js:89-161 */
double orc_fraction_inside(double bl, double br, double tl, double tr);
/* computeFaceWeights: sdf_collider3.js:131-262 */
void orc_face_weights3(const orc_primitive3* shapes, int count, int invert, const int n[3], const double h[3],
                       const double origin[3], float* wu, float* wv, float* ww);
/* addShapes: the collider SDF at cell centres, sdf_collider3.js:71-110 */
void orc_collider_sdf3(const orc_primitive3* shapes, int count, int invert, const int n[3], const double h[3],
                       const double origin[3], float* sdf);
/* constrainVelocity: grid_blocked_boundary_condition_solver3.js:389-477 for a
 * static collider (sdf on the velocity grid's cells, may be NULL; face
 * weights may be NULL) and the closed-wall flags */
void orc_constrain_velocity3(float* u, float* v, float* w, const int n[3], const float h[3], const float origin[3],
                             int closed_flags, const float* sdf, const float* wu, const float* wv, const float* ww,
                             float friction, int depth);
void orc_constrain_velocity3_inflows(float* u, float* v, float* w, const int n[3], const float h[3],
                                     const float origin[3], int closed_flags, const float* sdf, const float* wu,
                                     const float* wv, const float* ww, float friction, int depth, int n_inflows,
                                     const float* const* inflow_sdf, const float* inflow_vel,
                                     const int* inflow_mode);
void orc_outflow_dirichlet3(const float* osdf, const int n[3], const float h[3], const float origin[3],
                            uint8_t* active);
void orc_push_out3(float* pos, int64_t count, const float* sdf, const int n[3], const float h[3],
                   const float origin[3], float thresh, float shift);

/* the resample pass: grid_flip_solver3.js:640-793 */
void orc_resample3(float* pos, float* vel, int64_t count, const float* u, const float* v, const float* w,
                   const int n[3], const float h[3], const float origin[3], int min_pc, int max_pc);

/* carried concentration and density coupling of createGridFlipSolver3
 * (grid_flip_solver3.js:265-357, 493-531, 848-880, 967-1049) */
void orc_conc_accumulate3(const float* pos, const float* conc, int64_t count, const int n[3], const float h[3],
                          const float origin[3], float scale, int32_t* accum, int32_t* cnt);
void orc_conc_resolve3(const int32_t* accum, const int32_t* cnt, int64_t cells, float scale, int coupling, float d0,
                       float d1, float* cell_conc, float* cell_density);
void orc_density_beta3(const float* density, const float* fluid_mask, const int n[3], float d0, float d1,
                       float* bu, float* bv, float* bw);
void orc_reduced_gravity3(float* u, float* v, float* w, const float* const beta[3], const float* density,
                          const int n[3], const float h[3], const float origin[3], const float g[3], float dt, float d0,
                          float d1);
void orc_mix_fade3(const float* pos, float* conc, int64_t count, const float* cell_conc, const int n[3],
                   const float h[3], const float origin[3], float mixing, float fade);
void orc_p2g_fixed_mass3(const float* pos, const float* vel, const float* conc, int64_t count, int comp,
                         const int face_n[3], const float data_origin[3], const float h[3], float scale, float d0,
                         float d1, int32_t* numer, int32_t* denom);
/* phase-aware resampling (grid_two_phase_flip_solver2.js:1131-1503, in 3-D) */
void orc_resample_phase3(float* pos, float* vel, const float* conc, int64_t count, const float* u, const float* v,
                         const float* w, const int n[3], const float h[3], const float origin[3], int min_pc,
                         int max_pc, const int32_t* cnt, const int32_t* accum, float scale);
/* advectScalar3, order 1, no collider: advection_solver3.js:285-300 */
void orc_advect_scalar3(const float* src, float* dst, const float* u, const float* v, const float* w, const int n[3],
                        const float h[3], const float origin[3], float dt, int max_sub);
/* order 2 (MacCormack), no collider: advection_solver3.js:90-128 (clamp),
 * 180-282 (faces), 302-357 (scalar) */
void orc_advect_faces3_vd(const float* const vel[3], const float* const data[3], float* const out[3], const int n[3],
                          const float h[3], const float origin[3], float dt, int max_sub);
void orc_advect_faces3_mc_vd(const float* const vel[3], const float* const data[3], float* const out[3],
                             const int n[3], const float h[3], const float origin[3], float dt, int max_sub);
/* trace (advection_solver3.js:84-88) of count points, in place */
void orc_trace3(float* pos, int64_t count, const float* u, const float* v, const float* w, const int n[3],
                const float h[3], const float origin[3], float dt, int max_sub, float direction);
/* advection through a static collider (csdf: its SDF cell-centred on the
 * velocity grid, or NULL): advection_solver3.js:20-62 (the trace stops at
 * the collider), 130-357 (points inside keep the output's value; order 2
 * clamps over the corners outside the collider) */
void orc_advect_faces3_c(const float* const vel[3], const float* const data[3], float* const out[3], const int n[3],
                         const float h[3], const float origin[3], const float* csdf, float dt, int max_sub,
                         int order);
void orc_advect_scalar3_c(const float* src, float* dst, const float* u, const float* v, const float* w,
                          const int n[3], const float h[3], const float origin[3], const float* csdf, float dt,
                          int max_sub, int order);
void orc_advect_particles3_c(float* pos, int64_t count, const float* u, const float* v, const float* w,
                             const int n[3], const float h[3], const float origin[3], const float* csdf, float dt,
                             int max_sub);
void orc_trace3_c(float* pos, int64_t count, const float* u, const float* v, const float* w, const int n[3],
                  const float h[3], const float origin[3], const float* csdf, float dt, int max_sub,
                  float direction);
void orc_advect_face3_mc(const float* u, const float* v, const float* w, float* u_out, float* v_out, float* w_out,
                         const int n[3], const float h[3], const float origin[3], float dt, int max_sub);
void orc_advect_scalar3_mc(const float* src, float* dst, const float* u, const float* v, const float* w,
                           const int n[3], const float h[3], const float origin[3], float dt, int max_sub);
/* external force with the smoke buoyancy: external_force_solver3.js:19-47,
 * grid_smoke_solver3.js:66-81 */
void orc_buoyancy3(float* u, float* v, float* w, const float* dens, const float* temp, const int n[3],
                   const float h[3], const float origin[3], float a, float b, float tamb, const float up[3], float dt);
/* ball source with max: examples/34-smoke-plume-3d/main.js:146-158 */
void orc_source_sphere3(float* d, const int n[3], const float h[3], const float origin[3], const float c[3], float r,
                        float value);

/* ---------------- 2-D Eulerian step (oracle_grid2.c) ---------------- */
/* polygonsSignedDistance: polygon_sdf.js:24-97 (rings of (x, y) pairs;
 * ring r spans points ring_start[r] .. ring_start[r+1]-1) */
double orc_polygons_distance2(const double* xy, const int* ring_start, int rings, double x, double y);
/* addPolygons: sdf_collider2.js:82-100 */
void orc_polygon_sdf2(const double* xy, const int* ring_start, int rings, int invert, const int n[2],
                      const double h[2], const double origin[2], float* sdf);
/* collocatedValueAtPosition2 / collocatedCubicValueAtPosition2: grid_math.js:35-76, 102-167 */
float orc_sample_bilinear2(const float* d, const int s[2], const float o[2], const float h[2], float px, float py);
float orc_sample_cubic2(const float* d, const int s[2], const float o[2], const float h[2], float px, float py);
/* advectFaceCentered2, order 1, with an optional collider SDF:
 * advection_solver2.js:81-127, 263-320 (uo/vo hold the pre-advection copy) */
void orc_advect_face2(const float* u, const float* v, const int n[2], const float h[2], const float origin[2],
                      const float* sdf, float dt, int max_sub, float* uo, float* vo);
/* advectFaceCentered2, order 2 (MacCormack) with the collider:
 * advection_solver2.js:221-258, 310-430; the _vd form advects data (du, dv)
 * through the velocity (u, v), order 1 or 2 */
void orc_advect_faces2_vd(const float* u, const float* v, const float* du, const float* dv, const int n[2],
                          const float h[2], const float origin[2], const float* sdf, float dt, int max_sub, int order,
                          float* uo, float* vo);
/* trace with the collider (advection_solver2.js:197-201), in place */
void orc_trace2(float* pos, int64_t count, const float* u, const float* v, const int n[2], const float h[2],
                const float origin[2], const float* sdf, float dt, int max_sub, float direction);
void orc_advect_face2_mc(const float* u, const float* v, const int n[2], const float h[2], const float origin[2],
                         const float* sdf, float dt, int max_sub, float* uo, float* vo);
/* constrainVelocity: grid_blocked_boundary_condition_solver2.js:75-575 */
void orc_constrain_velocity2(float* u, float* v, const int n[2], const float h[2], const float origin[2],
                             int closed_flags, const float* sdf, float friction, int depth, int n_inflows,
                             const float* const* inflow_sdf, const float* inflow_vel, const int* inflow_mode);
/* applyOutflowVelocityBC: grid_outflow_solver2.js:245-365 */
void orc_outflow_bc2(float* u, float* v, const int n[2], const float h[2], const float origin[2], float dt,
                     const float* osdf);
/* createGridPressureSolver2 build and correction: grid_pressure_solver2.js:677-834 */
void orc_pressure2_assemble(const int n[2], const float h[2], const float* u, const float* v, const uint8_t* active,
                            const float* target, float* b, float* p);
void orc_pressure2_correct(const int n[2], const float h[2], const float* p, float* u, float* v);
/* createOutflowPressureDirichlet2 at the cell centres: sdf_inflow_outflow2.js:152-166 */
void orc_outflow_dirichlet2(const float* osdf, const int n[2], const float h[2], const float origin[2],
                            uint8_t* active);
/* createMaxAbsReducer: reduction.js:43-77 */
float orc_max_abs(const float* const* arrays, const int64_t* counts, int narrays);
/* computeAdaptiveSubSteps: time/cfl.js:38-48 */
int orc_adaptive_substeps(double max_velocity, double min_spacing, double frame_dt, double courant, int max_sub,
                          double* sub_dt);

#endif
