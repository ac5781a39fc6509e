/*
 * Real-templated body of the oracle's solver (TEST INFRASTRUCTURE ONLY).
 * Included by oracle_f64.c and oracle_f32.c with REAL and SUF defined.
 *
 * Each routine follows the reference function cited beside it.  Arithmetic is
 * done in REAL, in the reference's operation order, so that the f32 arm is
 * the reference's own single-precision arithmetic and the f64 arm the FP64
 * truth.  Dot products accumulate in double in both arms.
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "oracle_util.h"

#define OCAT2(a, b) a##_##b
#define OCAT(a, b) OCAT2(a, b)
#define FN(name) OCAT(name, SUF)

/* ------------------------------------------------------------------------
 * One row of A·x: laplacianAt, packages/fluxflow/src/linalg/multigrid.js:273-339.
 *   missing neighbour      -> its difference is 0 (Neumann)
 *   Dirichlet neighbour    -> dLower = c, dUpper = -c (elimination)
 *   face weight            -> multiplies each difference after that
 *   masked (Dirichlet) row -> -c
 * ------------------------------------------------------------------------ */
static inline REAL FN(row_apply)(const orc_grid* g, const REAL* x, const uint8_t* m,
                                 const REAL* const* w, int i, int j, int k) {
    const int* n = g->n;
    int64_t c = orc_cidx(n, i, j, k);
    REAL center = x[c];
    if (m && m[c]) return -center;
    int I[3] = {i, j, k};
    REAL sum = 0;
    for (int a = 0; a < g->ndim; ++a) {
        int idx = I[a];
        REAL dLo = 0, dUp = 0;
        int lo[3] = {i, j, k}, up[3] = {i, j, k};
        lo[a] -= 1;
        up[a] += 1;
        if (idx > 0) {
            int64_t l = orc_cidx(n, lo[0], lo[1], lo[2]);
            dLo = (m && m[l]) ? center : (REAL)(center - x[l]);
        }
        if (idx < n[a] - 1) {
            int64_t u = orc_cidx(n, up[0], up[1], up[2]);
            dUp = (m && m[u]) ? -center : (REAL)(x[u] - center);
        }
        if (w) {
            dLo = dLo * w[a][orc_fidx(n, a, i, j, k)];
            dUp = dUp * w[a][orc_fidx(n, a, up[0], up[1], up[2])];
        }
        REAL h2 = (REAL)(g->h[a] * g->h[a]);
        REAL term = (REAL)(dUp - dLo) / h2;
        sum = (a == 0) ? term : (REAL)(sum + term);
    }
    return sum;
}

/* laplacianDiagonalAt, multigrid.js:399-437: Dirichlet neighbours are NOT
 * dropped from the diagonal; a masked row's diagonal is -1. */
static inline REAL FN(row_diag)(const orc_grid* g, const uint8_t* m, const REAL* const* w,
                                int i, int j, int k) {
    const int* n = g->n;
    int64_t c = orc_cidx(n, i, j, k);
    if (m && m[c]) return (REAL)-1;
    int I[3] = {i, j, k};
    REAL sum = 0;
    for (int a = 0; a < g->ndim; ++a) {
        int idx = I[a];
        int up[3] = {i, j, k};
        up[a] += 1;
        REAL hasLo = idx > 0 ? (REAL)1 : (REAL)0;
        REAL hasUp = idx < n[a] - 1 ? (REAL)1 : (REAL)0;
        if (w) {
            hasLo = hasLo * w[a][orc_fidx(n, a, i, j, k)];
            hasUp = hasUp * w[a][orc_fidx(n, a, up[0], up[1], up[2])];
        }
        REAL scale = (REAL)(-1.0 / (g->h[a] * g->h[a]));
        REAL term = (REAL)(hasLo + hasUp) * scale;
        sum = (a == 0) ? term : (REAL)(sum + term);
    }
    return sum;
}

/* createLaplacianOperator, multigrid.js:457-471 */
void FN(orc_apply)(const orc_grid* g, const REAL* x, REAL* out, const uint8_t* mask,
                   const REAL* const* w) {
    const int* n = g->n;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i)
                out[orc_cidx(n, i, j, k)] = FN(row_apply)(g, x, mask, w, i, j, k);
}

void FN(orc_diag)(const orc_grid* g, REAL* out, const uint8_t* mask, const REAL* const* w) {
    const int* n = g->n;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i)
                out[orc_cidx(n, i, j, k)] = FN(row_diag)(g, mask, w, i, j, k);
}

/* buildRelaxKernel, multigrid.js:626-640, colour = (i+j+k) mod 2 (colorOf,
 * 442-448): x_c += ((b - A x)_c / diag_c) * omega on cells of `color`.
 * Neighbours of a cell always have the other colour, so the sweep reads
 * only values this sweep does not write. */
void FN(orc_relax)(const orc_grid* g, REAL* x, const REAL* b, const uint8_t* mask,
                   const REAL* const* w, double omega, int color) {
    const int* n = g->n;
    REAL om = (REAL)omega;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = ((j + k + color) & 1); i < n[0]; i += 2) {
                int64_t c = orc_cidx(n, i, j, k);
                REAL ax = FN(row_apply)(g, x, mask, w, i, j, k);
                REAL d = FN(row_diag)(g, mask, w, i, j, k);
                REAL upd = (REAL)((REAL)(b[c] - ax) / d) * om;
                x[c] = (REAL)(x[c] + upd);
            }
}

/* buildResidualKernel, multigrid.js:741-749 */
void FN(orc_residual)(const orc_grid* g, const REAL* x, const REAL* b, REAL* r,
                      const uint8_t* mask, const REAL* const* w) {
    const int* n = g->n;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                int64_t c = orc_cidx(n, i, j, k);
                r[c] = (REAL)(b[c] - FN(row_apply)(g, x, mask, w, i, j, k));
            }
}

static void FN(coarse_shape)(const orc_grid* fine, int* nc) {
    for (int a = 0; a < 3; ++a) nc[a] = (a < fine->ndim) ? fine->n[a] / 2 : fine->n[a];
}

/* restrictionTapsForAxis (766-780) + buildRestrictKernel (866-914): per axis
 * taps 1/8,3/8,3/8,1/8 at 2c-1 (clamped to 2c), 2c, 2c+1, 2c+2 (clamped to
 * 2c+1); a masked fine cell contributes 0; a masked coarse cell gets 0. */
void FN(orc_restrict)(const orc_grid* fine, const REAL* r_fine, const uint8_t* fine_mask,
                      REAL* b_coarse, const uint8_t* coarse_mask) {
    int nc[3];
    FN(coarse_shape)(fine, nc);
    const int* nf = fine->n;
    static const double tw[4] = {0.125, 0.375, 0.375, 0.125};
#pragma omp parallel for schedule(static)
    for (int K = 0; K < nc[2]; ++K)
        for (int J = 0; J < nc[1]; ++J)
            for (int I = 0; I < nc[0]; ++I) {
                int C[3] = {I, J, K};
                int tap[3][4];
                int ntap[3];
                for (int a = 0; a < 3; ++a) {
                    if (a < fine->ndim) {
                        int c = C[a];
                        tap[a][0] = c > 0 ? 2 * c - 1 : 2 * c;
                        tap[a][1] = 2 * c;
                        tap[a][2] = 2 * c + 1;
                        tap[a][3] = (c + 1 < nc[a]) ? 2 * c + 2 : 2 * c + 1;
                        ntap[a] = 4;
                    } else {
                        tap[a][0] = 0;
                        ntap[a] = 1;
                    }
                }
                REAL sum = 0;
                int first = 1;
                for (int t0 = 0; t0 < ntap[0]; ++t0)
                    for (int t1 = 0; t1 < ntap[1]; ++t1)
                        for (int t2 = 0; t2 < ntap[2]; ++t2) {
                            double wgt = 1.0 * tw[t0];
                            wgt = wgt * tw[t1];
                            if (ntap[2] == 4) wgt = wgt * tw[t2];
                            int64_t f = orc_cidx(nf, tap[0][t0], tap[1][t1], tap[2][t2]);
                            REAL v = (fine_mask && fine_mask[f]) ? (REAL)0 : r_fine[f];
                            REAL term = v * (REAL)wgt;
                            sum = first ? term : (REAL)(sum + term);
                            first = 0;
                        }
                int64_t cc = orc_cidx(nc, I, J, K);
                b_coarse[cc] = (coarse_mask && coarse_mask[cc]) ? (REAL)0 : sum;
            }
}

/* correctionTapsForAxis (921-934) + buildCorrectKernel (990-1040): bilinear /
 * trilinear prolongation ADDED to the fine x except at masked fine cells. */
void FN(orc_correct)(const orc_grid* fine, const REAL* x_coarse, REAL* x_fine,
                     const uint8_t* fine_mask) {
    int nc[3];
    FN(coarse_shape)(fine, nc);
    const int* nf = fine->n;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nf[2]; ++k)
        for (int j = 0; j < nf[1]; ++j)
            for (int i = 0; i < nf[0]; ++i) {
                int F[3] = {i, j, k};
                int ti[3][2];
                REAL tw[3][2];
                int nt[3];
                for (int a = 0; a < 3; ++a) {
                    if (a < fine->ndim) {
                        int ci = F[a] / 2;
                        int even = (F[a] % 2) == 0;
                        int lower = ci > 0 ? ci - 1 : ci;
                        int upper = ci < nc[a] - 1 ? ci + 1 : ci;
                        ti[a][0] = even ? lower : ci;
                        tw[a][0] = even ? (REAL)0.25 : (REAL)0.75;
                        ti[a][1] = even ? ci : upper;
                        tw[a][1] = even ? (REAL)0.75 : (REAL)0.25;
                        nt[a] = 2;
                    } else {
                        ti[a][0] = 0;
                        tw[a][0] = 1;
                        nt[a] = 1;
                    }
                }
                REAL sum = 0;
                int first = 1;
                for (int t0 = 0; t0 < nt[0]; ++t0)
                    for (int t1 = 0; t1 < nt[1]; ++t1)
                        for (int t2 = 0; t2 < nt[2]; ++t2) {
                            REAL wgt = (REAL)1 * tw[0][t0];
                            wgt = wgt * tw[1][t1];
                            if (nt[2] == 2) wgt = wgt * tw[2][t2];
                            REAL term = x_coarse[orc_cidx(nc, ti[0][t0], ti[1][t1], ti[2][t2])] * wgt;
                            sum = first ? term : (REAL)(sum + term);
                            first = 0;
                        }
                int64_t f = orc_cidx(nf, i, j, k);
                if (!(fine_mask && fine_mask[f])) x_fine[f] = (REAL)(x_fine[f] + sum);
            }
}

/* buildFaceWeightCoarsenKernel, multigrid.js:1138-1184: coarse face = mean of
 * the 2^(d-1) fine faces at index 2I along the face axis. */
void FN(orc_coarsen_weights)(const orc_grid* fine, const REAL* const* w_fine,
                             REAL* const* w_coarse) {
    int nc[3];
    FN(coarse_shape)(fine, nc);
    const int* nf = fine->n;
    for (int a = 0; a < fine->ndim; ++a) {
        int fs[3] = {nc[0] + (a == 0), nc[1] + (a == 1), nc[2] + (a == 2)};
        int count = 1 << (fine->ndim - 1);
#pragma omp parallel for schedule(static)
        for (int K = 0; K < fs[2]; ++K)
            for (int J = 0; J < fs[1]; ++J)
                for (int I = 0; I < fs[0]; ++I) {
                    int C[3] = {I, J, K};
                    int lo[3], cnt[3];
                    for (int b = 0; b < 3; ++b) {
                        lo[b] = (b < fine->ndim) ? 2 * C[b] : 0;
                        cnt[b] = (b < fine->ndim && b != a) ? 2 : 1;
                    }
                    REAL sum = 0;
                    int first = 1;
                    for (int d0 = 0; d0 < cnt[0]; ++d0)
                        for (int d1 = 0; d1 < cnt[1]; ++d1)
                            for (int d2 = 0; d2 < cnt[2]; ++d2) {
                                REAL v = w_fine[a][orc_fidx(nf, a, lo[0] + d0, lo[1] + d1, lo[2] + d2)];
                                sum = first ? v : (REAL)(sum + v);
                                first = 0;
                            }
                    w_coarse[a][orc_fidx(nc, a, I, J, K)] = sum / (REAL)count;
                }
    }
}

/* buildMaskCoarsenKernel, multigrid.js:1066-1122: coarse cell masked when ALL
 * of its fine children in `fine_src` are masked, OR (with weights) when the
 * six coarse face weights sum below 1e-6. */
void FN(orc_coarsen_mask)(const orc_grid* fine, const uint8_t* fine_src,
                          const REAL* const* w_coarse, uint8_t* coarse_mask) {
    int nc[3];
    FN(coarse_shape)(fine, nc);
    const int* nf = fine->n;
#pragma omp parallel for schedule(static)
    for (int K = 0; K < nc[2]; ++K)
        for (int J = 0; J < nc[1]; ++J)
            for (int I = 0; I < nc[0]; ++I) {
                int C[3] = {I, J, K};
                int masked = 0, have = 0;
                if (fine_src) {
                    int cnt[3];
                    for (int b = 0; b < 3; ++b) cnt[b] = (b < fine->ndim) ? 2 : 1;
                    masked = 1;
                    for (int d0 = 0; d0 < cnt[0]; ++d0)
                        for (int d1 = 0; d1 < cnt[1]; ++d1)
                            for (int d2 = 0; d2 < cnt[2]; ++d2) {
                                int f0 = (cnt[0] == 2) ? 2 * I + d0 : I;
                                int f1 = (cnt[1] == 2) ? 2 * J + d1 : J;
                                int f2 = (cnt[2] == 2) ? 2 * K + d2 : K;
                                masked = masked && fine_src[orc_cidx(nf, f0, f1, f2)];
                            }
                    have = 1;
                }
                if (w_coarse) {
                    REAL total = 0;
                    for (int a = 0; a < fine->ndim; ++a) {
                        int up[3] = {C[0], C[1], C[2]};
                        up[a] += 1;
                        REAL pair = (REAL)(w_coarse[a][orc_fidx(nc, a, I, J, K)] +
                                           w_coarse[a][orc_fidx(nc, a, up[0], up[1], up[2])]);
                        total = (a == 0) ? pair : (REAL)(total + pair);
                    }
                    int enclosed = total < (REAL)1e-6;
                    masked = have ? (masked || enclosed) : enclosed;
                    have = 1;
                }
                coarse_mask[orc_cidx(nc, I, J, K)] = (uint8_t)(have && masked);
            }
}

/* ------------------------------------------------------------------------
 * Multigrid preconditioner: createMultigridPreconditioner, multigrid.js:1213-1735
 * ------------------------------------------------------------------------ */
typedef struct {
    orc_mg_options o;
    int levels;
    orc_grid g[ORC_MAX_LEVELS];
    REAL *x[ORC_MAX_LEVELS], *b[ORC_MAX_LEVELS], *buf[ORC_MAX_LEVELS];
    uint8_t* mask[ORC_MAX_LEVELS]; /* level 0: caller's (not owned) */
    REAL* w[ORC_MAX_LEVELS][3];    /* level 0: caller's (not owned) */
    int has_w, has_mask_levels;
} FN(orc_mg);

void* FN(orc_mg_create)(const orc_grid* g, const orc_mg_options* o, const uint8_t* mask,
                        const uint8_t* coarse_src_mask, const REAL* const* w) {
    if (o->levels < 1 || o->levels > ORC_MAX_LEVELS) return NULL;
    if (o->smoother == ORC_SMOOTH_JACOBI &&
        ((o->nu_down & 1) || (o->nu_up & 1) || (o->n_coarsest & 1) || (o->n_final & 1)))
        return NULL; /* multigrid_jacobi.js:1273-1276: sweep counts must be even */
    int div = 1 << (o->levels - 1);
    for (int a = 0; a < g->ndim; ++a)
        if (g->n[a] % div) return NULL; /* computeLevelShapes 176-202 */
    FN(orc_mg)* mg = (FN(orc_mg)*)calloc(1, sizeof(FN(orc_mg)));
    mg->o = *o;
    mg->levels = o->levels;
    mg->has_w = (w != NULL);
    mg->has_mask_levels = (mask != NULL) || (w != NULL); /* coarseningMasks 1261-1262 */
    for (int l = 0; l < mg->levels; ++l) {
        orc_grid* gl = &mg->g[l];
        gl->ndim = g->ndim;
        for (int a = 0; a < 3; ++a) {
            gl->n[a] = (a < g->ndim) ? (g->n[a] >> l) : g->n[a];
            gl->h[a] = g->h[a] * (double)(1 << l);
        }
        int64_t cells = orc_cells(gl->n);
        if (l > 0) {
            mg->x[l] = (REAL*)calloc(cells, sizeof(REAL));
            mg->b[l] = (REAL*)calloc(cells, sizeof(REAL));
        }
        mg->buf[l] = (REAL*)calloc(cells, sizeof(REAL));
        if (l == 0) {
            mg->mask[0] = (uint8_t*)mask;
            for (int a = 0; a < 3; ++a) mg->w[0][a] = w ? (REAL*)w[a] : NULL;
        } else {
            if (mg->has_mask_levels) mg->mask[l] = (uint8_t*)calloc(cells, 1);
            if (mg->has_w)
                for (int a = 0; a < g->ndim; ++a) {
                    int64_t nf = orc_faces(gl->n, a);
                    mg->w[l][a] = (REAL*)malloc(nf * sizeof(REAL));
                    for (int64_t q = 0; q < nf; ++q) mg->w[l][a][q] = 1;
                }
        }
    }
    /* refreshCoarseLevels (1415-1437): weights of level l first, then the
     * level-l mask from level l-1's source (level 1 reads the vent-only mask). */
    for (int l = 1; l < mg->levels; ++l) {
        if (mg->has_w) FN(orc_coarsen_weights)(&mg->g[l - 1], (const REAL* const*)mg->w[l - 1], mg->w[l]);
        if (!mg->has_mask_levels) continue;
        const uint8_t* src = NULL;
        if (mask) src = (l == 1) ? (coarse_src_mask ? coarse_src_mask : mask) : mg->mask[l - 1];
        FN(orc_coarsen_mask)(&mg->g[l - 1], src, mg->has_w ? (const REAL* const*)mg->w[l] : NULL, mg->mask[l]);
    }
    return mg;
}

void FN(orc_mg_destroy)(void* p) {
    FN(orc_mg)* mg = (FN(orc_mg)*)p;
    if (!mg) return;
    for (int l = 0; l < mg->levels; ++l) {
        free(mg->buf[l]);
        if (l > 0) {
            free(mg->x[l]);
            free(mg->b[l]);
            free(mg->mask[l]);
            for (int a = 0; a < 3; ++a) free(mg->w[l][a]);
        }
    }
    free(mg);
}

int FN(orc_mg_level_info)(void* p, int level, int* n_out, uint8_t** mask, REAL** w) {
    FN(orc_mg)* mg = (FN(orc_mg)*)p;
    if (level < 0 || level >= mg->levels) return -1;
    for (int a = 0; a < 3; ++a) n_out[a] = mg->g[level].n[a];
    if (mask) *mask = mg->mask[level];
    if (w)
        for (int a = 0; a < 3; ++a) w[a] = mg->w[level][a];
    return 0;
}

/* One damped-Jacobi sweep (buildJacobiRelaxKernel,
 * sandbox/jacobi-smoother/multigrid_jacobi.js:632-642): every cell reads the
 * old iterate `xr` and writes xr_c + ((b - A xr)_c / diag_c) * omega to `xw`. */
void FN(orc_jacobi)(const orc_grid* g, const REAL* xr, REAL* xw, const REAL* b, const uint8_t* mask,
                    const REAL* const* w, double omega) {
    const int* n = g->n;
    REAL om = (REAL)omega;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                int64_t c = orc_cidx(n, i, j, k);
                REAL ax = FN(row_apply)(g, xr, mask, w, i, j, k);
                REAL d = FN(row_diag)(g, mask, w, i, j, k);
                REAL upd = (REAL)((REAL)(b[c] - ax) / d) * om;
                xw[c] = (REAL)(xr[c] + upd);
            }
}

/* Brick-wise red-black Gauss-Seidel (no counterpart in the reference: the
 * library's one-pass smoother).  The grid is cut into bricks of
 * ORC_BRICK_X x ORC_BRICK_Y x ORC_BRICK_Z cells (origin at cell 0, partial
 * bricks at the upper ends); every brick runs `sweeps` colour pairs of the
 * reference's relax (multigrid.js:626-640, same row, same colours) on its own
 * cells, reading a neighbour outside the brick from `xin` (the field on
 * entry, unchanged throughout) and one inside from `xout`.  `reversed` runs
 * colour 1 before colour 0, which makes the post-smoother the adjoint of the
 * pre-smoother, so the V-cycle stays symmetric.  xin == NULL means the field
 * is zero on entry. */
void FN(orc_brick_relax_shape)(const orc_grid* g, const REAL* xin, REAL* xout, const REAL* b, const uint8_t* mask,
                         const REAL* const* w, double omega, int sweeps, int reversed, const int* B, int shifted);
static inline REAL FN(brick_nbr)(const REAL* xin, const REAL* xout, int64_t q, int in_brick) {
    if (in_brick) return xout[q];
    return xin ? xin[q] : (REAL)0;
}

static inline REAL FN(row_apply_brick)(const orc_grid* g, const REAL* xin, const REAL* xout, const int* lo,
                                       const int* hi, const uint8_t* m, const REAL* const* w, int i, int j,
                                       int k) {
    const int* n = g->n;
    int64_t c = orc_cidx(n, i, j, k);
    REAL center = xout[c];
    if (m && m[c]) return -center;
    int I[3] = {i, j, k};
    REAL sum = 0;
    for (int a = 0; a < g->ndim; ++a) {
        int idx = I[a];
        REAL dLo = 0, dUp = 0;
        int lo3[3] = {i, j, k}, up3[3] = {i, j, k};
        lo3[a] -= 1;
        up3[a] += 1;
        if (idx > 0) {
            int64_t l = orc_cidx(n, lo3[0], lo3[1], lo3[2]);
            REAL xl = FN(brick_nbr)(xin, xout, l, lo3[a] >= lo[a]);
            dLo = (m && m[l]) ? center : (REAL)(center - xl);
        }
        if (idx < n[a] - 1) {
            int64_t u = orc_cidx(n, up3[0], up3[1], up3[2]);
            REAL xu = FN(brick_nbr)(xin, xout, u, up3[a] < hi[a]);
            dUp = (m && m[u]) ? -center : (REAL)(xu - center);
        }
        if (w) {
            dLo = dLo * w[a][orc_fidx(n, a, i, j, k)];
            dUp = dUp * w[a][orc_fidx(n, a, up3[0], up3[1], up3[2])];
        }
        REAL h2 = (REAL)(g->h[a] * g->h[a]);
        REAL term = (REAL)(dUp - dLo) / h2;
        sum = (a == 0) ? term : (REAL)(sum + term);
    }
    return sum;
}

void FN(orc_brick_relax)(const orc_grid* g, const REAL* xin, REAL* xout, const REAL* b, const uint8_t* mask,
                         const REAL* const* w, double omega, int sweeps, int reversed) {
    FN(orc_brick_relax_shape)(g, xin, xout, b, mask, w, omega, sweeps, reversed, orc_brick_dims(), 0);
}

void FN(orc_brick_relax_shape)(const orc_grid* g, const REAL* xin, REAL* xout, const REAL* b, const uint8_t* mask,
                         const REAL* const* w, double omega, int sweeps, int reversed, const int* B, int shifted) {
    const int* n = g->n;
    int nb[3], off[3];
    for (int a = 0; a < 3; ++a) {
        off[a] = shifted ? B[a] / 2 : 0;
        nb[a] = (n[a] + off[a] + B[a] - 1) / B[a];
    }
    REAL om = (REAL)omega;
    const int64_t bricks = (int64_t)nb[0] * nb[1] * nb[2];
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < bricks; ++q) {
        int bi = (int)(q % nb[0]), bj = (int)((q / nb[0]) % nb[1]), bk = (int)(q / ((int64_t)nb[0] * nb[1]));
        int lo[3] = {bi * B[0] - off[0], bj * B[1] - off[1], bk * B[2] - off[2]};
        int hi[3];
        for (int a = 0; a < 3; ++a) {
            hi[a] = lo[a] + B[a] < n[a] ? lo[a] + B[a] : n[a];
            if (lo[a] < 0) lo[a] = 0;
        }
        for (int k = lo[2]; k < hi[2]; ++k)
            for (int j = lo[1]; j < hi[1]; ++j)
                for (int i = lo[0]; i < hi[0]; ++i) {
                    int64_t c = orc_cidx(n, i, j, k);
                    xout[c] = xin ? xin[c] : (REAL)0;
                }
        for (int s = 0; s < sweeps; ++s)
            for (int h = 0; h < 2; ++h) {
                int color = reversed ? 1 - h : h;
                for (int k = lo[2]; k < hi[2]; ++k)
                    for (int j = lo[1]; j < hi[1]; ++j)
                        for (int i = lo[0] + ((lo[0] + j + k + color) & 1); i < hi[0]; i += 2) {
                            int64_t c = orc_cidx(n, i, j, k);
                            REAL ax = FN(row_apply_brick)(g, xin, xout, lo, hi, mask, w, i, j, k);
                            REAL d = FN(row_diag)(g, mask, w, i, j, k);
                            REAL upd = (REAL)((REAL)(b[c] - ax) / d) * om;
                            xout[c] = (REAL)(xout[c] + upd);
                        }
            }
    }
}

/* The levels the brick smoother covers: every 3-D level but the coarsest
 * with more than ORC_BRICK_MIN_CELLS cells (smaller levels smooth globally,
 * as the library runs them in one workgroup). */
static int FN(brick_level)(const FN(orc_mg)* mg, int l) {
    return mg->o.smoother == ORC_SMOOTH_BRICK && mg->g[l].ndim == 3 && l < mg->levels - 1 &&
           orc_cells(mg->g[l].n) > (int64_t)ORC_BRICK_MIN_CELLS;
}

/* relax (multigrid.js:1589-1607; the Jacobi form multigrid_jacobi.js:1271-1285):
 * red-black takes `iters` colour pairs, post-smoothing colours reversed; the
 * Jacobi form takes `iters` single sweeps ping-ponging through buf (an even
 * count lands back in x; the reference rejects odd counts) and ignores the
 * order, being symmetric by construction. */
static void FN(mg_relax_n)(FN(orc_mg)* mg, int l, int iters, int reversed) {
    if (orc_post_forward_flag()) reversed = 0; /* the non-symmetric V-cycle of the stop-code tests */
    if (FN(brick_level)(mg, l)) {
        /* all `iters` sweeps in one brick pass; buf holds the field on entry */
        if (iters <= 0) return;
        const REAL* const* w = mg->w[l][0] ? (const REAL* const*)mg->w[l] : NULL;
        const int passes = orc_brick_passes();
        for (int ps = 0; ps < passes; ++ps) {
            const int shifted = reversed ? (passes - 1 - ps) : ps;
            memcpy(mg->buf[l], mg->x[l], sizeof(REAL) * orc_cells(mg->g[l].n));
            FN(orc_brick_relax_shape)(&mg->g[l], mg->buf[l], mg->x[l], mg->b[l], mg->mask[l], w, mg->o.omega, iters,
                                      reversed, orc_brick_dims(), shifted);
        }
        return;
    }
    if (mg->o.smoother == ORC_SMOOTH_JACOBI) {
        const REAL* const* w = mg->w[l][0] ? (const REAL* const*)mg->w[l] : NULL;
        for (int it = 0; it < iters; ++it) {
            const REAL* src = (it & 1) ? mg->buf[l] : mg->x[l];
            REAL* dst = (it & 1) ? mg->x[l] : mg->buf[l];
            FN(orc_jacobi)(&mg->g[l], src, dst, mg->b[l], mg->mask[l], w, mg->o.omega);
        }
        return;
    }
    for (int it = 0; it < iters; ++it) {
        const REAL* const* w = mg->w[l][0] ? (const REAL* const*)mg->w[l] : NULL;
        FN(orc_relax)(&mg->g[l], mg->x[l], mg->b[l], mg->mask[l], w, mg->o.omega, reversed ? 1 : 0);
        FN(orc_relax)(&mg->g[l], mg->x[l], mg->b[l], mg->mask[l], w, mg->o.omega, reversed ? 0 : 1);
    }
}

/* vCycle, multigrid.js:1609-1642 */
static void FN(mg_vcycle)(FN(orc_mg)* mg, int l) {
    if (l == mg->levels - 1) {
        /* coarsest: colour 0 then 1, numberOfCoarsestIterations times
         * (buildCoarseSweepKernel 679-737 and relax() agree on this order) */
        FN(mg_relax_n)(mg, l, mg->o.n_coarsest, 0);
        return;
    }
    const REAL* const* w = mg->w[l][0] ? (const REAL* const*)mg->w[l] : NULL;
    FN(mg_relax_n)(mg, l, mg->o.nu_down, 0);
    FN(orc_residual)(&mg->g[l], mg->x[l], mg->b[l], mg->buf[l], mg->mask[l], w);
    FN(orc_restrict)(&mg->g[l], mg->buf[l], mg->mask[l], mg->b[l + 1], mg->mask[l + 1]);
    memset(mg->x[l + 1], 0, sizeof(REAL) * orc_cells(mg->g[l + 1].n));
    FN(mg_vcycle)(mg, l + 1);
    FN(orc_correct)(&mg->g[l], mg->x[l + 1], mg->x[l], mg->mask[l]);
    FN(mg_relax_n)(mg, l, l > 0 ? mg->o.nu_up : mg->o.n_final, 1);
}

/* applyMultigridPreconditioner: z = one V-cycle from z = 0 (buildForm 1683-1711) */
void FN(orc_mg_apply)(void* p, const REAL* r, REAL* z) {
    FN(orc_mg)* mg = (FN(orc_mg)*)p;
    mg->x[0] = z;
    mg->b[0] = (REAL*)r;
    memset(z, 0, sizeof(REAL) * orc_cells(mg->g[0].n));
    FN(mg_vcycle)(mg, 0);
    mg->x[0] = NULL;
    mg->b[0] = NULL;
}

/* ------------------------------------------------------------------------
 * PCG.  Normative semantics: host solve(), linalg.js:2176-2465, with the true
 * residual recomputed every iteration (RESIDUAL_RECOMPUTE_INTERVAL = 1, line
 * 184; the GPU-resident self-stopping list 1366-1370 recomputes from the
 * first iteration, which is what is done here).  Guards: linalg.js:226, 244,
 * 284, 547-551, 1095.
 * ------------------------------------------------------------------------ */
/* fixed 64K-element chunks summed in order: the same value for any thread count */
static double FN(dot)(const REAL* a, const REAL* b, int64_t n) {
    const int64_t CH = 65536;
    int64_t nch = (n + CH - 1) / CH;
    double* part = (double*)malloc(sizeof(double) * (nch > 0 ? nch : 1));
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < nch; ++c) {
        double s = 0;
        int64_t e = (c + 1) * CH < n ? (c + 1) * CH : n;
        for (int64_t q = c * CH; q < e; ++q) s += (double)a[q] * (double)b[q];
        part[c] = s;
    }
    double s = 0;
    for (int64_t c = 0; c < nch; ++c) s += part[c];
    free(part);
    return s;
}

static int FN(degenerate)(double v) { return !isfinite(v) || v == 0.0; }

double FN(orc_true_residual)(const orc_grid* g, const uint8_t* mask, const REAL* const* w,
                             const REAL* b, const REAL* x) {
    /* ||b - A x||_2 evaluated in double from the stored values */
    const int* n = g->n;
    double s = 0;
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                int64_t c = orc_cidx(n, i, j, k);
                double ax;
                if (mask && mask[c]) {
                    ax = -(double)x[c];
                } else {
                    int I[3] = {i, j, k};
                    ax = 0;
                    for (int a = 0; a < g->ndim; ++a) {
                        int lo[3] = {i, j, k}, up[3] = {i, j, k};
                        lo[a] -= 1;
                        up[a] += 1;
                        double dLo = 0, dUp = 0, cc = (double)x[c];
                        if (I[a] > 0) {
                            int64_t l = orc_cidx(n, lo[0], lo[1], lo[2]);
                            dLo = (mask && mask[l]) ? cc : cc - (double)x[l];
                        }
                        if (I[a] < n[a] - 1) {
                            int64_t u = orc_cidx(n, up[0], up[1], up[2]);
                            dUp = (mask && mask[u]) ? -cc : (double)x[u] - cc;
                        }
                        if (w) {
                            dLo *= (double)w[a][orc_fidx(n, a, i, j, k)];
                            dUp *= (double)w[a][orc_fidx(n, a, up[0], up[1], up[2])];
                        }
                        ax += (dUp - dLo) / (g->h[a] * g->h[a]);
                    }
                }
                double r = (double)b[c] - ax;
                s += r * r;
            }
    return sqrt(s);
}

int FN(orc_pcg_solve)(const orc_grid* g, const uint8_t* mask, const REAL* const* w, int prec,
                      void* mg, const REAL* b, REAL* x, double tol, int maxiter, int relative,
                      orc_pcg_state* st) {
    int64_t n = orc_cells(g->n);
    REAL* r = (REAL*)calloc(n, sizeof(REAL));
    REAL* z = (REAL*)calloc(n, sizeof(REAL));
    REAL* p = (REAL*)calloc(n, sizeof(REAL));
    REAL* Ap = (REAL*)calloc(n, sizeof(REAL));
    REAL* Ax = (REAL*)calloc(n, sizeof(REAL));
    REAL* dg = NULL;
    if (prec == ORC_PREC_JACOBI) {
        dg = (REAL*)calloc(n, sizeof(REAL));
        FN(orc_diag)(g, dg, mask, w);
    }
#define APPLY_PREC()                                                                   \
    do {                                                                               \
        if (prec == ORC_PREC_MULTIGRID) FN(orc_mg_apply)(mg, r, z);                    \
        else if (prec == ORC_PREC_JACOBI) {                                            \
            for (int64_t q = 0; q < n; ++q) z[q] = (dg[q] != 0) ? r[q] / dg[q] : r[q]; \
        } else memcpy(z, r, sizeof(REAL) * n);                                         \
    } while (0)

    memset(st, 0, sizeof(*st));
    FN(orc_apply)(g, x, Ax, mask, w);
    for (int64_t q = 0; q < n; ++q) r[q] = (REAL)(b[q] - Ax[q]);
    double initRTr = FN(dot)(r, r, n);
    double bb = FN(dot)(b, b, n);
    double thr = relative ? fmax(tol * sqrt(fabs(bb)), tol) : tol;
    st->threshold = thr;
    st->b_norm = sqrt(fabs(bb));
    double newRTr = initRTr;
    APPLY_PREC();
    double oldRZ = FN(dot)(r, z, n);
    memcpy(p, z, sizeof(REAL) * n);
    double base = fmax(fabs(initRTr), 1e-12);
    double best = -1;
    if (sqrt(fabs(initRTr)) >= thr) {
        for (int it = 0; it < maxiter; ++it) {
            st->iterations = it + 1;
            FN(orc_apply)(g, p, Ap, mask, w);
            double pAp = FN(dot)(p, Ap, n);
            if (FN(degenerate)(pAp)) { st->stopped_by = ORC_STOP_DEGENERATE_PAP; break; }
            if (fabs(pAp) > base * 1e8) { st->stopped_by = ORC_STOP_PAP_GROWTH; break; }
            double alpha = oldRZ / pAp;
            if (!(fabs(alpha) <= 1e6)) { st->stopped_by = ORC_STOP_ALPHA_MAGNITUDE; break; }
            REAL al = (REAL)alpha;
            for (int64_t q = 0; q < n; ++q) x[q] = (REAL)(x[q] + (REAL)(p[q] * al));
            FN(orc_apply)(g, x, Ax, mask, w);
            for (int64_t q = 0; q < n; ++q) r[q] = (REAL)(b[q] - Ax[q]);
            newRTr = FN(dot)(r, r, n);
            if (sqrt(fabs(newRTr)) < thr) break;
            if (best < 0 || fabs(newRTr) < best) best = fabs(newRTr);
            else if (best > 0 && fabs(newRTr) > best * 100) { st->stopped_by = ORC_STOP_RESIDUAL_GROWING; break; }
            APPLY_PREC();
            double newRZ = FN(dot)(r, z, n);
            if (FN(degenerate)(oldRZ)) { st->stopped_by = ORC_STOP_DEGENERATE_OLD_RZ; break; }
            int flip = (oldRZ > 0) != (newRZ > 0);
            double betaRaw = newRZ / oldRZ;
            double beta = (flip || fabs(betaRaw) > 10) ? 0.0 : betaRaw;
            st->restarts += (flip || fabs(betaRaw) > 10) ? 1 : 0;
            REAL be = (REAL)beta;
            for (int64_t q = 0; q < n; ++q) p[q] = (REAL)(z[q] + (REAL)(p[q] * be));
            oldRZ = newRZ;
        }
    }
#undef APPLY_PREC
    st->residual_sq = newRTr;
    st->converged = sqrt(fabs(newRTr)) < thr;
    /* a guard that fires after convergence is a non-event (guardIsANonEvent) */
    if (st->converged) st->stopped_by = ORC_STOP_CONVERGED;
    free(r); free(z); free(p); free(Ap); free(Ax); free(dg);
    return st->converged;
}

/* ------------------------------------------------------------------------
 * Pressure system assembly and velocity correction:
 * packages/fluxflow/src/grid/grid_pressure_solver3.js
 * ------------------------------------------------------------------------ */
void FN(orc_pressure3_assemble)(const orc_grid* g, const uint8_t* active, const REAL* target,
                                const REAL* const* beta, const REAL* const* cw,
                                const REAL* const* vel, uint8_t* mask_out, uint8_t* vent_out,
                                REAL* target_out, REAL* const* wsys_out, REAL* b_out,
                                REAL* pressure) {
    const int* n = g->n;
    int has_mask = (active != NULL) || (cw != NULL);
    /* updateDirichletFields 171-206 */
    if (has_mask) {
        for (int k = 0; k < n[2]; ++k)
            for (int j = 0; j < n[1]; ++j)
                for (int i = 0; i < n[0]; ++i) {
                    int64_t c = orc_cidx(n, i, j, k);
                    int enclosed = 0;
                    if (cw) {
                        REAL t = cw[0][orc_fidx(n, 0, i, j, k)];
                        t = t + cw[0][orc_fidx(n, 0, i + 1, j, k)];
                        t = t + cw[1][orc_fidx(n, 1, i, j, k)];
                        t = t + cw[1][orc_fidx(n, 1, i, j + 1, k)];
                        t = t + cw[2][orc_fidx(n, 2, i, j, k)];
                        t = t + cw[2][orc_fidx(n, 2, i, j, k + 1)];
                        enclosed = t < (REAL)1e-6;
                    }
                    if (active) {
                        int act = active[c] != 0;
                        mask_out[c] = (uint8_t)(act || enclosed);
                        target_out[c] = enclosed ? (act ? target[c] : (REAL)0) : target[c];
                        vent_out[c] = (uint8_t)act;
                    } else {
                        mask_out[c] = (uint8_t)enclosed;
                        target_out[c] = 0;
                        vent_out[c] = 0;
                    }
                }
    }
    /* combineAxisWeight 217-229 */
    int has_w = (beta != NULL) || (cw != NULL);
    if (has_w)
        for (int a = 0; a < 3; ++a) {
            int64_t nf = orc_faces(n, a);
            for (int64_t q = 0; q < nf; ++q)
                wsys_out[a][q] = (beta && cw) ? (REAL)(beta[a][q] * cw[a][q]) : (beta ? beta[a][q] : cw[a][q]);
        }
    REAL ih2[3];
    for (int a = 0; a < 3; ++a) ih2[a] = (REAL)(1.0 / (g->h[a] * g->h[a]));
    /* dispatchBuildSystem 429-453 with weightedDivergence 412-427 and
     * neighbourDirichletContribution 350-394 */
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                int64_t c = orc_cidx(n, i, j, k);
                REAL lu = vel[0][orc_fidx(n, 0, i, j, k)], ru = vel[0][orc_fidx(n, 0, i + 1, j, k)];
                REAL bv = vel[1][orc_fidx(n, 1, i, j, k)], tv = vel[1][orc_fidx(n, 1, i, j + 1, k)];
                REAL bw = vel[2][orc_fidx(n, 2, i, j, k)], fw = vel[2][orc_fidx(n, 2, i, j, k + 1)];
                if (cw) {
                    lu = lu * cw[0][orc_fidx(n, 0, i, j, k)];
                    ru = ru * cw[0][orc_fidx(n, 0, i + 1, j, k)];
                    bv = bv * cw[1][orc_fidx(n, 1, i, j, k)];
                    tv = tv * cw[1][orc_fidx(n, 1, i, j + 1, k)];
                    bw = bw * cw[2][orc_fidx(n, 2, i, j, k)];
                    fw = fw * cw[2][orc_fidx(n, 2, i, j, k + 1)];
                }
                REAL div = (REAL)(ru - lu) / (REAL)g->h[0];
                div = div + (REAL)((REAL)(tv - bv) / (REAL)g->h[1]);
                div = div + (REAL)((REAL)(fw - bw) / (REAL)g->h[2]);
                if (!has_mask) {
                    b_out[c] = div;
                    continue;
                }
                if (mask_out[c]) {
                    b_out[c] = -target_out[c];
                    pressure[c] = target_out[c];
                    continue;
                }
                REAL nb = 0;
                int I[3] = {i, j, k};
                for (int a = 0; a < 3; ++a) {
                    for (int side = 0; side < 2; ++side) {
                        int q[3] = {i, j, k};
                        q[a] += side ? 1 : -1;
                        int inb = side ? (I[a] < n[a] - 1) : (I[a] > 0);
                        REAL term = 0;
                        if (inb) {
                            int64_t nc = orc_cidx(n, q[0], q[1], q[2]);
                            if (mask_out[nc]) {
                                int fi[3] = {i, j, k};
                                if (side) fi[a] += 1;
                                REAL wf = has_w ? wsys_out[a][orc_fidx(n, a, fi[0], fi[1], fi[2])] : (REAL)1;
                                term = (REAL)(wf * target_out[nc]) * ih2[a];
                            }
                        }
                        nb = (a == 0 && side == 0) ? term : (REAL)(nb + term);
                    }
                }
                b_out[c] = (REAL)(div - nb);
            }
}

/* dispatchCorrectU/V/W 457-494: interior faces only, u -= beta * grad p,
 * gated off where the collider weight is 0. */
void FN(orc_pressure3_correct)(const orc_grid* g, const REAL* pressure,
                               const REAL* const* beta, const REAL* const* cw,
                               const REAL* const* vel_in, REAL* const* vel_out) {
    const int* n = g->n;
    for (int a = 0; a < 3; ++a) {
        int fs[3] = {n[0] + (a == 0), n[1] + (a == 1), n[2] + (a == 2)};
        REAL hh = (REAL)g->h[a];
        for (int k = 0; k < fs[2]; ++k)
            for (int j = 0; j < fs[1]; ++j)
                for (int i = 0; i < fs[0]; ++i) {
                    int F[3] = {i, j, k};
                    int m = F[a];
                    if (!(m > 0 && m < n[a])) continue;
                    int lo[3] = {i, j, k};
                    lo[a] -= 1;
                    int64_t f = orc_fidx(n, a, i, j, k);
                    REAL grad = (REAL)(pressure[orc_cidx(n, i, j, k)] -
                                       pressure[orc_cidx(n, lo[0], lo[1], lo[2])]) / hh;
                    REAL scaled = beta ? (REAL)(grad * beta[a][f]) : grad;
                    REAL gated = cw ? (REAL)(scaled * (cw[a][f] > 0 ? (REAL)1 : (REAL)0)) : scaled;
                    vel_out[a][f] = (REAL)(vel_in[a][f] - gated);
                }
    }
}

/* countBadPressureCells, grid_pressure_solver3.js:297-304 with
 * isNonFiniteOrAbove (float_guards.js:107-111) */
int64_t FN(orc_count_bad)(const REAL* p, int64_t n, double max_plausible) {
    int64_t bad = 0;
    for (int64_t q = 0; q < n; ++q)
        if (!(fabs((double)p[q]) <= max_plausible)) bad++;
    return bad;
}

/* faceCenteredDivergenceAtCenter3, grid_math3.js:397-410 */
void FN(orc_divergence3)(const orc_grid* g, const REAL* const* vel, REAL* out) {
    const int* n = g->n;
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                REAL d = (REAL)(vel[0][orc_fidx(n, 0, i + 1, j, k)] - vel[0][orc_fidx(n, 0, i, j, k)]) / (REAL)g->h[0];
                d = d + (REAL)((REAL)(vel[1][orc_fidx(n, 1, i, j + 1, k)] - vel[1][orc_fidx(n, 1, i, j, k)]) / (REAL)g->h[1]);
                d = d + (REAL)((REAL)(vel[2][orc_fidx(n, 2, i, j, k + 1)] - vel[2][orc_fidx(n, 2, i, j, k)]) / (REAL)g->h[2]);
                out[orc_cidx(n, i, j, k)] = d;
            }
}
