// 3-D pressure projection on the MAC grid:
// createGridPressureSolver3 (packages/fluxflow/src/grid/grid_pressure_solver3.js:66-574).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "flux_internal.h"
#include "flux_solver.h"

namespace flux {

static constexpr int PTPB = 256;

struct Grid3 {
    int n0, n1, n2;
};

__device__ __forceinline__ void cell3(const Grid3& g, int c, int& i, int& j, int& k) {
    int row = c / g.n0;
    i = c - row * g.n0;
    k = row / g.n1;
    j = row - k * g.n1;
}
__device__ __forceinline__ int fu(const Grid3& g, int i, int j, int k) { return i + (g.n0 + 1) * (j + g.n1 * k); }
__device__ __forceinline__ int fv(const Grid3& g, int i, int j, int k) { return i + g.n0 * (j + (g.n1 + 1) * k); }
__device__ __forceinline__ int fw(const Grid3& g, int i, int j, int k) { return i + g.n0 * (j + g.n1 * k); }

// updateDirichletFields (grid_pressure_solver3.js:171-206): enclosed = the six
// collider weights sum below 1e-6; mask = active OR enclosed; target is the
// user's, or 0 for a cell only enclosure pins; vent = active only.
__global__ void k_p3_dirichlet(Grid3 g, const uint8_t* __restrict__ active, const float* __restrict__ target,
                               const float* __restrict__ cu, const float* __restrict__ cv, const float* __restrict__ cw,
                               uint8_t* __restrict__ mask, uint8_t* __restrict__ vent, float* __restrict__ tout) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.n0 * g.n1 * g.n2) return;
    int i, j, k;
    cell3(g, c, i, j, k);
    bool enclosed = false;
    if (cu) {
        float t = cu[fu(g, i, j, k)];
        t = t + cu[fu(g, i + 1, j, k)];
        t = t + cv[fv(g, i, j, k)];
        t = t + cv[fv(g, i, j + 1, k)];
        t = t + cw[fw(g, i, j, k)];
        t = t + cw[fw(g, i, j, k + 1)];
        enclosed = t < 1e-6f;
    }
    if (active) {
        bool act = active[c] != 0;
        float tg = target ? target[c] : 0.0f;
        mask[c] = (uint8_t)(act || enclosed);
        tout[c] = enclosed ? (act ? tg : 0.0f) : tg;
        vent[c] = (uint8_t)act;
    } else {
        mask[c] = (uint8_t)enclosed;
        tout[c] = 0.0f;
        vent[c] = 0;
    }
}

// combineAxisWeight (217-229): density beta x collider open fraction
__global__ void k_p3_combine(int n, const float* __restrict__ a, const float* __restrict__ b, float* __restrict__ out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = a[t] * b[t];
}

// dispatchBuildSystem (429-453): fluid row b = collider-weighted divergence
// (weightedDivergence 412-427) minus the Dirichlet neighbours' w*target/h^2
// (neighbourDirichletContribution 350-394); Dirichlet row b = -target and
// the pressure is seeded with the target.
__global__ void k_p3_build(Grid3 g, float hx, float hy, float hz, float ih2x, float ih2y, float ih2z,
                           const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ w,
                           const float* __restrict__ cu, const float* __restrict__ cv, const float* __restrict__ cwz,
                           const uint8_t* __restrict__ mask, const float* __restrict__ target,
                           const float* __restrict__ wu, const float* __restrict__ wv, const float* __restrict__ ww,
                           float* __restrict__ b, float* __restrict__ p) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.n0 * g.n1 * g.n2) return;
    int i, j, k;
    cell3(g, c, i, j, k);
    int ul = fu(g, i, j, k), ur = ul + 1;
    int vb = fv(g, i, j, k), vt = vb + g.n0;
    int wb = c, wf = c + g.n0 * g.n1;
    float lu = u[ul], ru = u[ur], bv = v[vb], tv = v[vt], bw = w[wb], fwv = w[wf];
    if (cu) {
        lu *= cu[ul];
        ru *= cu[ur];
        bv *= cv[vb];
        tv *= cv[vt];
        bw *= cwz[wb];
        fwv *= cwz[wf];
    }
    float div = (ru - lu) / hx;
    div = div + (tv - bv) / hy;
    div = div + (fwv - bw) / hz;
    if (!mask) {
        b[c] = div;
        return;
    }
    if (mask[c]) {
        float tg = target[c];
        b[c] = -tg;
        p[c] = tg;
        return;
    }
    float nb = 0.0f;
    const int n[3] = {g.n0, g.n1, g.n2};
    const int I[3] = {i, j, k};
    const int st[3] = {1, g.n0, g.n0 * g.n1};
    const float* wsa[3] = {wu, wv, ww};
    const int flo[3] = {ul, vb, wb};
    const int fst[3] = {1, g.n0, g.n0 * g.n1};
    const float ih2[3] = {ih2x, ih2y, ih2z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (I[a] > 0 && mask[c - st[a]]) nb += (wsa[a] ? wsa[a][flo[a]] : 1.0f) * target[c - st[a]] * ih2[a];
        if (I[a] < n[a] - 1 && mask[c + st[a]]) nb += (wsa[a] ? wsa[a][flo[a] + fst[a]] : 1.0f) * target[c + st[a]] * ih2[a];
    }
    b[c] = div - nb;
}

// countBadPressureCells (297-304): non-finite or |p| > maxPlausible
__global__ void k_p3_count_bad(const float* __restrict__ p, int n, float maxp, unsigned long long* count) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = false;
    if (c < n) bad = !(fabsf(p[c]) <= maxp);
    unsigned long long m = __ballot(bad);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(count, (unsigned long long)__popcll(m));
}

// dispatchCorrectU/V/W (457-494): interior faces 0 < m < n:
// out = in - beta * (p_m - p_{m-1}) / h, gated off where the collider weight is 0.
__global__ void k_p3_correct(Grid3 g, int axis, float h, const float* __restrict__ p, const float* __restrict__ beta,
                             const float* __restrict__ cwt, const float* __restrict__ in, float* __restrict__ out) {
    int s0 = g.n0 + (axis == 0), s1 = g.n1 + (axis == 1), s2 = g.n2 + (axis == 2);
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= s0 * s1 * s2) return;
    int row = t / s0;
    int i = t - row * s0;
    int k = row / s1;
    int j = row - k * s1;
    int F[3] = {i, j, k};
    int nn[3] = {g.n0, g.n1, g.n2};
    int m = F[axis];
    if (!(m > 0 && m < nn[axis])) return;  // boundary faces are left as they are
    int hi = i + g.n0 * (j + g.n1 * k);
    int lo = hi - (axis == 0 ? 1 : (axis == 1 ? g.n0 : g.n0 * g.n1));
    float grad = (p[hi] - p[lo]) / h;
    float scaled = beta ? grad * beta[t] : grad;
    float gated = cwt ? scaled * (cwt[t] > 0.0f ? 1.0f : 0.0f) : scaled;
    out[t] = in[t] - gated;
}

__global__ void k_p3_copy(const float* __restrict__ a, float* __restrict__ b, int n) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) b[t] = a[t];
}

}  // namespace flux

using namespace flux;

struct flux_pressure3 {
    flux_pressure3_desc d;
    int n[3];
    double h[3];
    flux_field* pressure = nullptr;
    flux_field* b = nullptr;
    flux_field* mask = nullptr;
    flux_field* vent = nullptr;
    flux_field* target = nullptr;
    flux_field* wsys[3] = {nullptr, nullptr, nullptr};  // owned when both beta and collider weights exist
    flux_field* snapshot = nullptr;
    const flux_field* wop[3] = {nullptr, nullptr, nullptr};  // weights seen by the operator
    bool has_mask = false, has_w = false, own_w = false;
    flux_operator* op = nullptr;
    flux_multigrid* mg = nullptr;
    flux_pcg* pcg = nullptr;
    unsigned long long* bad = nullptr;
    int maxiter = 0;
};

extern "C" {

void flux_pressure3_desc_default(flux_pressure3_desc* d) {
    memset(d, 0, sizeof *d);
    for (int a = 0; a < 3; ++a) d->spacing[a] = 1.0;
    flux_mg_options_default(&d->multigrid);
    d->preconditioner = FLUX_PREC_MULTIGRID;
    d->tolerance = 1e-6;
    d->max_iterations = 0;
    d->max_iterations_factor = 1.5;
    d->relative_tolerance = 1;
    d->max_plausible_pressure = 1e6;
}

int flux_pressure3_destroy(flux_pressure3* p) {
    if (!p) return FLUX_OK;
    flux_pcg_destroy(p->pcg);
    flux_multigrid_destroy(p->mg);
    flux_operator_destroy(p->op);
    flux_field_destroy(p->pressure);
    flux_field_destroy(p->b);
    flux_field_destroy(p->mask);
    flux_field_destroy(p->vent);
    flux_field_destroy(p->target);
    flux_field_destroy(p->snapshot);
    for (int a = 0; a < 3; ++a) flux_field_destroy(p->wsys[a]);
    dfree(p->bad);
    delete p;
    return FLUX_OK;
}

int flux_pressure3_create(const flux_pressure3_desc* d, flux_pressure3** out) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(d && out, "createGridPressureSolver3: NULL argument");
    FLUX_TRY(check_shape(3, d->resolution, d->spacing, 3, "createGridPressureSolver3"));
    FLUX_REQUIRE(d->preconditioner >= 0 && d->preconditioner <= 2,
                 "grid_pressure_solver3: unknown preconditioner '%d' (expected multigrid, jacobi, none).", d->preconditioner);
    flux_pressure3* p = new flux_pressure3();
    p->d = *d;
    for (int a = 0; a < 3; ++a) {
        p->n[a] = d->resolution[a];
        p->h[a] = d->spacing[a];
    }
    int64_t cells = cells_of(p->n);
    int rc = FLUX_OK;
#define P3_CHECK(expr)            \
    do {                          \
        rc = (expr);              \
        if (rc) {                 \
            flux_pressure3_destroy(p); \
            return rc;            \
        }                         \
    } while (0)
    if (d->dirichlet_active) P3_CHECK(check_field(d->dirichlet_active, FLUX_U8, cells, "dirichlet active"));
    if (d->dirichlet_target) P3_CHECK(check_field(d->dirichlet_target, FLUX_F32, cells, "dirichlet target"));
    bool has_beta = d->face_weights[0] != nullptr;
    bool has_cw = d->collider_weights[0] != nullptr;
    for (int a = 0; a < 3; ++a) {
        if (has_beta) P3_CHECK(check_field(d->face_weights[a], FLUX_F32, faces_of(p->n, a), "faceWeights"));
        if (has_cw) P3_CHECK(check_field(d->collider_weights[a], FLUX_F32, faces_of(p->n, a), "colliderWeights"));
    }
    int64_t shape[3] = {p->n[0], p->n[1], p->n[2]};
    P3_CHECK(flux_field_create(FLUX_F32, 3, shape, &p->pressure));
    P3_CHECK(flux_field_create(FLUX_F32, 3, shape, &p->b));
    P3_CHECK(flux_field_create(FLUX_F32, 3, shape, &p->snapshot));
    P3_CHECK(flux_field_create(FLUX_F32, 3, shape, &p->target));
    p->has_mask = d->dirichlet_active || has_cw;
    if (p->has_mask) {
        P3_CHECK(flux_field_create(FLUX_U8, 3, shape, &p->mask));
        P3_CHECK(flux_field_create(FLUX_U8, 3, shape, &p->vent));
    }
    p->has_w = has_beta || has_cw;
    if (has_beta && has_cw) {
        p->own_w = true;
        for (int a = 0; a < 3; ++a) {
            int64_t fs[3] = {p->n[0] + (a == 0), p->n[1] + (a == 1), p->n[2] + (a == 2)};
            P3_CHECK(flux_field_create(FLUX_F32, 3, fs, &p->wsys[a]));
            p->wop[a] = p->wsys[a];
        }
    } else if (has_beta) {
        for (int a = 0; a < 3; ++a) p->wop[a] = d->face_weights[a];
    } else if (has_cw) {
        for (int a = 0; a < 3; ++a) p->wop[a] = d->collider_weights[a];
    }
    const flux_field* const* wop = p->has_w ? p->wop : nullptr;
    P3_CHECK(flux_operator_create(3, p->n, p->h, p->mask, wop, &p->op));
    if (d->preconditioner == FLUX_PREC_MULTIGRID)
        P3_CHECK(flux_multigrid_create(3, p->n, p->h, &d->multigrid, p->mask, p->vent, wop, &p->mg));
    P3_CHECK(flux_pcg_create(p->op, d->preconditioner, p->mg, p->b, p->pressure, &p->pcg));
    p->bad = (unsigned long long*)dmalloc(sizeof(unsigned long long));
    if (!p->bad) P3_CHECK(set_error(FLUX_ERR_NOMEM, "createGridPressureSolver3: allocation failed"));
    int mr = std::max(p->n[0], std::max(p->n[1], p->n[2]));
    p->maxiter = d->max_iterations > 0 ? d->max_iterations : (int)std::ceil(d->max_iterations_factor * mr * 1);
#undef P3_CHECK
    *out = p;
    return FLUX_OK;
}

int flux_pressure3_project(flux_pressure3* p, const flux_field* u_in, const flux_field* v_in, const flux_field* w_in,
                           flux_field* u_out, flux_field* v_out, flux_field* w_out, flux_pressure_diag* diag) {
    FLUX_TRY(flux_pressure3_assemble(p, u_in, v_in, w_in));
    return flux_pressure3_solve(p, u_in, v_in, w_in, u_out, v_out, w_out, diag);
}

flux_field* flux_pressure3_field(flux_pressure3* p, int which) {
    if (!p) return nullptr;
    switch (which) {
        case FLUX_P3_PRESSURE: return p->pressure;
        case FLUX_P3_RHS: return p->b;
        case FLUX_P3_MASK: return p->mask;
        case FLUX_P3_VENT: return p->vent;
        case FLUX_P3_TARGET: return p->target;
        case FLUX_P3_WEIGHT_U: return (flux_field*)p->wop[0];
        case FLUX_P3_WEIGHT_V: return (flux_field*)p->wop[1];
        case FLUX_P3_WEIGHT_W: return (flux_field*)p->wop[2];
        default: return nullptr;
    }
}

static int check_velocity(flux_pressure3* p, const flux_field* const* vin, flux_field* const* vout) {
    for (int a = 0; a < 3; ++a) {
        int rc = check_field(vin[a], FLUX_F32, faces_of(p->n, a), "project input");
        if (rc)
            return set_error(FLUX_ERR_ARG,
                             "createGridPressureSolver3: project() input resolution does not match constructed resolution [%d, %d, %d].",
                             p->n[0], p->n[1], p->n[2]);
        if (vout) FLUX_TRY(check_field(vout[a], FLUX_F32, faces_of(p->n, a), "project output"));
    }
    return FLUX_OK;
}

int flux_pressure3_assemble(flux_pressure3* p, const flux_field* u_in, const flux_field* v_in, const flux_field* w_in) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(p, "project: NULL argument");
    const flux_field* vin[3] = {u_in, v_in, w_in};
    FLUX_TRY(check_velocity(p, vin, nullptr));
    const flux_pressure3_desc& d = p->d;
    Grid3 g{p->n[0], p->n[1], p->n[2]};
    int64_t cells = cells_of(p->n);
    unsigned gc = blocks_for(cells, PTPB);
    const float* cw[3] = {nullptr, nullptr, nullptr};
    const float* beta[3] = {nullptr, nullptr, nullptr};
    for (int a = 0; a < 3; ++a) {
        if (d.collider_weights[0]) cw[a] = (const float*)d.collider_weights[a]->d;
        if (d.face_weights[0]) beta[a] = (const float*)d.face_weights[a]->d;
    }
    // buildSystemBatch: updateDirichletFields + dispatchBuildSystem (496-504)
    if (p->has_mask)
        k_p3_dirichlet<<<gc, PTPB, 0, g_stream>>>(g, d.dirichlet_active ? (const uint8_t*)d.dirichlet_active->d : nullptr,
                                                   d.dirichlet_target ? (const float*)d.dirichlet_target->d : nullptr,
                                                   cw[0], cw[1], cw[2], (uint8_t*)p->mask->d, (uint8_t*)p->vent->d,
                                                   (float*)p->target->d);
    if (p->own_w)
        for (int a = 0; a < 3; ++a) {
            int64_t nf = faces_of(p->n, a);
            k_p3_combine<<<blocks_for(nf, PTPB), PTPB, 0, g_stream>>>((int)nf, beta[a], cw[a], (float*)p->wsys[a]->d);
        }
    const float* ws[3] = {nullptr, nullptr, nullptr};
    if (p->has_w)
        for (int a = 0; a < 3; ++a) ws[a] = (const float*)p->wop[a]->d;
    k_p3_build<<<gc, PTPB, 0, g_stream>>>(
        g, (float)p->h[0], (float)p->h[1], (float)p->h[2], (float)(1.0 / (p->h[0] * p->h[0])),
        (float)(1.0 / (p->h[1] * p->h[1])), (float)(1.0 / (p->h[2] * p->h[2])), (const float*)u_in->d,
        (const float*)v_in->d, (const float*)w_in->d, cw[0], cw[1], cw[2],
        p->has_mask ? (const uint8_t*)p->mask->d : nullptr, (const float*)p->target->d, ws[0], ws[1], ws[2],
        (float*)p->b->d, (float*)p->pressure->d);
    FLUX_LAUNCH_CHECK();
    // refreshPreconditionerMask (multigrid.js:1449-1451)
    if (p->mg) FLUX_TRY(mg_refresh(p->mg));
    FLUX_HIP(hipStreamSynchronize(g_stream));
    return FLUX_OK;
}

int flux_pressure3_scaled_guess(flux_pressure3* p, const flux_field* guess, double* alpha) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(p && guess, "scaledGuess: NULL argument");
    return flux_pcg_scaled_start(p->pcg, guess, alpha);
}

int flux_pressure3_set_coarse_space(flux_pressure3* p, int free_indicator, int nfields,
                                    const flux_field* const* fields) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(p, "coarseSpace: NULL solver");
    return flux_pcg_set_coarse_space(p->pcg, free_indicator, nfields, fields);
}

int flux_pressure3_set_flexible(flux_pressure3* p, int flexible) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(p, "flexible: NULL solver");
    return flux_pcg_set_flexible(p->pcg, flexible);
}

int flux_pressure3_solve(flux_pressure3* p, const flux_field* u_in, const flux_field* v_in, const flux_field* w_in,
                         flux_field* u_out, flux_field* v_out, flux_field* w_out, flux_pressure_diag* diag) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(p && diag, "project: NULL argument");
    const flux_field* vin[3] = {u_in, v_in, w_in};
    flux_field* vout[3] = {u_out, v_out, w_out};
    FLUX_TRY(check_velocity(p, vin, vout));
    const flux_pressure3_desc& d = p->d;
    Grid3 g{p->n[0], p->n[1], p->n[2]};
    int64_t cells = cells_of(p->n);
    unsigned gc = blocks_for(cells, PTPB);
    const float* cw[3] = {nullptr, nullptr, nullptr};
    const float* beta[3] = {nullptr, nullptr, nullptr};
    for (int a = 0; a < 3; ++a) {
        if (d.collider_weights[0]) cw[a] = (const float*)d.collider_weights[a]->d;
        if (d.face_weights[0]) beta[a] = (const float*)d.face_weights[a]->d;
    }
    flux_pcg_state st;
    FLUX_TRY(pcg_solve_raw(p->pcg, d.tolerance, p->maxiter, d.relative_tolerance, &st));
    // circuit breaker (289-336, 545-554)
    FLUX_HIP(hipMemsetAsync(p->bad, 0, sizeof(unsigned long long), g_stream));
    k_p3_count_bad<<<gc, PTPB, 0, g_stream>>>((const float*)p->pressure->d, (int)cells, (float)d.max_plausible_pressure,
                                               p->bad);
    unsigned long long bad = 0;
    FLUX_HIP(hipMemcpyAsync(&bad, p->bad, sizeof bad, hipMemcpyDeviceToHost, g_stream));
    FLUX_HIP(hipStreamSynchronize(g_stream));
    bool rejected = bad > 0;
    if (rejected)
        k_p3_copy<<<gc, PTPB, 0, g_stream>>>((const float*)p->snapshot->d, (float*)p->pressure->d, (int)cells);
    else
        k_p3_copy<<<gc, PTPB, 0, g_stream>>>((const float*)p->pressure->d, (float*)p->snapshot->d, (int)cells);
    for (int a = 0; a < 3; ++a) {
        int64_t nf = faces_of(p->n, a);
        k_p3_correct<<<blocks_for(nf, PTPB), PTPB, 0, g_stream>>>(g, a, (float)p->h[a], (const float*)p->pressure->d,
                                                                   beta[a], cw[a], (const float*)vin[a]->d,
                                                                   (float*)vout[a]->d);
    }
    FLUX_LAUNCH_CHECK();
    FLUX_HIP(hipStreamSynchronize(g_stream));
    diag->converged = st.converged;
    diag->rejected = rejected ? 1 : 0;
    diag->iterations = st.iterations;
    diag->stopped_by = st.stopped_by;
    diag->residual = st.residual;
    diag->b_norm = st.b_norm;
    diag->bad_cells = (int64_t)bad;
    return FLUX_OK;
}

}  // extern "C"
