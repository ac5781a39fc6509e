/* FP32 instantiation of the oracle solver (test infrastructure only). */
#define REAL float
#define SUF f32
#include "solver_body.h"

/* ------------------------------------------------------------------------
 * Mixed arm: the arithmetic of the HIP plain path.  Search directions,
 * residual and preconditioner in f32; the solution accumulated in double;
 * the true residual b - A x evaluated in double arithmetic every iteration;
 * dots in double.  Same guards and stop rules as orc_pcg_solve_f32 above
 * (linalg.js:2176-2465 with the GPU-resident recompute-every-iteration list
 * 1366-1370).  x is the f32 warm start on entry and the rounded solution on
 * exit; x64_out (optional) receives the double accumulator.
 * ------------------------------------------------------------------------ */
static void apply_mixed(const orc_grid* g, const double* x, float* r, const float* b, const uint8_t* m,
                        const float* const* w, double* rr_out) {
    const int* n = g->n;
    double* part = (double*)calloc(n[2], sizeof(double));
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                int64_t c = orc_cidx(n, i, j, k);
                double ax;
                if (m && m[c]) ax = -x[c];
                else {
                    int I[3] = {i, j, k};
                    ax = 0;
                    for (int a = 0; a < g->ndim; ++a) {
                        int lo[3] = {i, j, k}, up[3] = {i, j, k};
                        lo[a] -= 1;
                        up[a] += 1;
                        double dLo = 0, dUp = 0, cc = x[c];
                        if (I[a] > 0) {
                            int64_t l = orc_cidx(n, lo[0], lo[1], lo[2]);
                            dLo = (m && m[l]) ? cc : cc - x[l];
                        }
                        if (I[a] < n[a] - 1) {
                            int64_t u = orc_cidx(n, up[0], up[1], up[2]);
                            dUp = (m && m[u]) ? -cc : x[u] - cc;
                        }
                        if (w) {
                            dLo *= (double)w[a][orc_fidx(n, a, i, j, k)];
                            dUp *= (double)w[a][orc_fidx(n, a, up[0], up[1], up[2])];
                        }
                        ax += (dUp - dLo) * (1.0 / (g->h[a] * g->h[a]));
                    }
                }
                double rv = (double)b[c] - ax;
                r[c] = (float)rv;
                part[k] += rv * rv;
            }
    double s = 0;
    for (int k = 0; k < n[2]; ++k) s += part[k];
    free(part);
    *rr_out = s;
}

/* Flexible (Polak-Ribiere) beta = r_{k+1}.(z_{k+1} - z_k) / r_k.z_k instead of
 * the reference's r_{k+1}.z_{k+1} / r_k.z_k (linalg.js:1198-1226): the
 * library's choice for mixed-precision V-cycles, restated here so the GPU
 * path can be compared with it (orc_set_flexible). */
static int g_flexible = 0;
void orc_set_flexible(int on) { g_flexible = on ? 1 : 0; }

static int g_brick[3] = {ORC_BRICK_X, ORC_BRICK_Y, ORC_BRICK_Z};
static int g_brick_passes = 2;
void orc_set_brick(int bx, int by, int bz, int passes) {
    g_brick[0] = bx; g_brick[1] = by; g_brick[2] = bz; g_brick_passes = passes;
}
const int* orc_brick_dims(void) { return g_brick; }
int orc_brick_passes(void) { return g_brick_passes; }

static int g_post_forward = 0;
void orc_set_post_forward(int on) { g_post_forward = on ? 1 : 0; }
int orc_post_forward_flag(void) { return g_post_forward; }

int orc_pcg_solve_mixed(const orc_grid* g, const uint8_t* mask, const float* const* w, int prec, void* mg,
                        const float* b, float* x, double* x64_out, double tol, int maxiter, int relative,
                        orc_pcg_state* st) {
    int64_t n = orc_cells(g->n);
    float* r = (float*)calloc(n, sizeof(float));
    float* z = (float*)calloc(n, sizeof(float));
    float* p = (float*)calloc(n, sizeof(float));
    float* Ap = (float*)calloc(n, sizeof(float));
    double* xd = (double*)calloc(n, sizeof(double));
    float* dg = NULL;
    if (prec == ORC_PREC_JACOBI) {
        dg = (float*)calloc(n, sizeof(float));
        orc_diag_f32(g, dg, mask, w);
    }
#define MIX_PREC()                                                                           \
    do {                                                                                     \
        if (prec == ORC_PREC_MULTIGRID) orc_mg_apply_f32(mg, r, z);                          \
        else if (prec == ORC_PREC_JACOBI) {                                                  \
            for (int64_t q = 0; q < n; ++q) z[q] = (fabsf(dg[q]) > 0) ? r[q] / dg[q] : r[q]; \
        } else memcpy(z, r, sizeof(float) * n);                                              \
    } while (0)
    memset(st, 0, sizeof(*st));
    for (int64_t q = 0; q < n; ++q) xd[q] = x[q];
    double rr;
    apply_mixed(g, xd, r, b, mask, w, &rr);
    double initRTr = rr;
    double bb = dot_f32(b, b, n);
    double thr = relative ? fmax(tol * sqrt(fabs(bb)), tol) : tol;
    st->threshold = thr;
    st->b_norm = sqrt(fabs(bb));
    double newRTr = initRTr;
    if (sqrt(fabs(initRTr)) < thr) {
        st->stopped_by = ORC_STOP_CONVERGED;
        st->converged = 1;
    } else {
        MIX_PREC();
        double oldRZ = dot_f32(r, z, n);
        double base = fmax(fabs(initRTr), 1e-12);
        double best = -1, beta = 0;
        for (int it = 0; it < maxiter; ++it) {
            st->iterations = it + 1;
            float be = (float)beta;
#pragma omp parallel for schedule(static)
            for (int64_t q = 0; q < n; ++q) p[q] = z[q] + p[q] * be;
            orc_apply_f32(g, p, Ap, mask, w);
            double pAp = dot_f32(p, Ap, n);
            double alpha = oldRZ / pAp;
            int code = 0;
            if (degenerate_f32(pAp)) code = ORC_STOP_DEGENERATE_PAP;
            else if (fabs(pAp) > base * 1e8) code = ORC_STOP_PAP_GROWTH;
            else if (!(fabs(alpha) <= 1e6)) code = ORC_STOP_ALPHA_MAGNITUDE;
            if (code) { st->stopped_by = code; break; }
#pragma omp parallel for schedule(static)
            for (int64_t q = 0; q < n; ++q) xd[q] = xd[q] + alpha * (double)p[q];
            apply_mixed(g, xd, r, b, mask, w, &rr);
            newRTr = rr;
            const double rzo = g_flexible ? dot_f32(r, z, n) : 0.0; /* r_{k+1}.z_k */
            double res = sqrt(fabs(rr));
            if (res <= thr) { st->stopped_by = ORC_STOP_CONVERGED; break; }
            if (best < 0 || fabs(rr) < best) best = fabs(rr);
            else if (best > 0 && fabs(rr) > best * 100) { st->stopped_by = ORC_STOP_RESIDUAL_GROWING; break; }
            MIX_PREC();
            double newRZ = dot_f32(r, z, n);
            if (degenerate_f32(oldRZ)) { st->stopped_by = ORC_STOP_DEGENERATE_OLD_RZ; break; }
            int flip = (oldRZ > 0) != (newRZ > 0);
            double braw = (g_flexible ? newRZ - rzo : newRZ) / oldRZ;
            beta = (flip || !(fabs(braw) <= 10)) ? 0.0 : braw;
            st->restarts += (flip || !(fabs(braw) <= 10)) ? 1 : 0;
            oldRZ = newRZ;
        }
        st->converged = st->stopped_by == ORC_STOP_CONVERGED;
    }
#undef MIX_PREC
    st->residual_sq = newRTr;
    for (int64_t q = 0; q < n; ++q) x[q] = (float)xd[q];
    if (x64_out) memcpy(x64_out, xd, sizeof(double) * n);
    free(r); free(z); free(p); free(Ap); free(xd); free(dg);
    return st->converged;
}
