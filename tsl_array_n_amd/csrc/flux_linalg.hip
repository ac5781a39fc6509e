// Dot reducer and plain conjugate gradient for gfx950:
// createDotReducer (packages/fluxflow/src/linalg/linalg.js:374-512) and
// createConjugateGradientSolver (linalg.js:563-728).
//
// The reducer keeps the reference's lane structure -- lane i of a 2-D/3-D
// field of shape [n0, ...] sums every element whose first index is i; a 1-D
// field of n elements has min(n, 64) lanes that stride by the lane count --
// and the reference's `partial` field (one FP32 value per lane).  The sums
// are FP64 and deterministic: each block of the first kernel adds a fixed
// set of rows for 256 consecutive lanes (coalesced loads), and a one-block
// kernel adds those per-block partials per lane in block order, then the
// lanes in lane order.
#include <algorithm>
#include <cmath>
#include <vector>

#include "flux_internal.h"
#include "flux_stencil.h"

namespace flux {

static constexpr int DOT_TPB = 256;
static constexpr int DOT_FIN_TPB = 1024;
static constexpr unsigned DOT_MAX_ROW_BLOCKS = 1024;

// part[x * lanes + i] = sum over rows r = x, x + gridDim.x, ... of
// a[i + lanes r] * b[i + lanes r] (elements past `count` are absent)
__global__ void __launch_bounds__(DOT_TPB) k_dot_lanes(const float* __restrict__ a, const float* __restrict__ b,
                                                       int64_t count, int lanes, int rows, double* __restrict__ part) {
    const int i = blockIdx.y * DOT_TPB + threadIdx.x;
    if (i >= lanes) return;
    double acc = 0.0;
    for (int r = blockIdx.x; r < rows; r += gridDim.x) {
        const int64_t e = (int64_t)i + (int64_t)lanes * r;
        if (e < count) acc += (double)a[e] * (double)b[e];
    }
    part[(size_t)blockIdx.x * lanes + i] = acc;
}

// lane64[i] = sum_x part[x * lanes + i] (x in order); lane32[i] its FP32
// value (the reference's `partial` field); lane64[lanes] = the lane total,
// also stored at *total when given
__global__ void __launch_bounds__(DOT_FIN_TPB) k_dot_finish(const double* __restrict__ part, unsigned gx, int lanes,
                                                            float* __restrict__ lane32, double* __restrict__ lane64,
                                                            double* __restrict__ total) {
    double mine = 0.0;
    for (int i = threadIdx.x; i < lanes; i += DOT_FIN_TPB) {
        double s = 0.0;
        for (unsigned x = 0; x < gx; ++x) s += part[(size_t)x * lanes + i];
        lane64[i] = s;
        if (lane32) lane32[i] = (float)s;
        mine += s;
    }
    const double tot = block_sum(mine);
    if (threadIdx.x == 0) {
        lane64[lanes] = tot;
        if (total) *total = tot;
    }
}

struct DotGeom {
    int lanes = 0, rows = 0;
    int64_t count = 0;
};

static int dot_geometry(int ndim, const int* shape, DotGeom* g, const char* who) {
    if (ndim < 1 || ndim > 3)
        return set_error(FLUX_ERR_ARG, "%s: only 1D/2D/3D shapes are supported, got %dD.", who, ndim);
    int64_t count = 1;
    for (int a = 0; a < ndim; ++a) {
        if (shape[a] <= 0) return set_error(FLUX_ERR_ARG, "%s: shape[%d] must be positive", who, a);
        count *= shape[a];
    }
    if (count >= (1ll << 31)) return set_error(FLUX_ERR_ARG, "%s: %lld elements exceed the 2^31 index range", who,
                                               (long long)count);
    g->count = count;
    if (ndim == 1) {
        g->lanes = std::min(shape[0], 64);  // MAX_1D_DOT_LANES
        g->rows = (int)((count + g->lanes - 1) / g->lanes);
    } else {
        g->lanes = shape[0];
        g->rows = (int)(count / shape[0]);
    }
    return FLUX_OK;
}

// Scratch of the allocation-free entry point: the per-block partials and the
// lane sums fit in DOT_SCRATCH_DOUBLES (2 MiB), so a shape has at most
// (DOT_SCRATCH_DOUBLES - 1) / 2 lanes there.  Both entry points use the same
// row-block count, so their sums run in the same order.
static constexpr size_t DOT_SCRATCH_DOUBLES = (size_t)1 << 18;

static bool dot_fits_scratch(const DotGeom& g) { return 2 * (size_t)g.lanes + 1 <= DOT_SCRATCH_DOUBLES; }

// Row blocks: at most DOT_MAX_ROW_BLOCKS, and within the scratch when the
// shape fits it.
static unsigned dot_row_blocks(const DotGeom& g) {
    size_t by_cap = dot_fits_scratch(g) ? (DOT_SCRATCH_DOUBLES - (size_t)g.lanes - 1) / (size_t)g.lanes : 1;
    return (unsigned)std::max<size_t>(1, std::min<size_t>({(size_t)g.rows, (size_t)DOT_MAX_ROW_BLOCKS, by_cap}));
}

static void dot_launch(const DotGeom& g, const float* a, const float* b, double* part, unsigned gx, float* lane32,
                       double* lane64, double* total) {
    dim3 grid(gx, (unsigned)((g.lanes + DOT_TPB - 1) / DOT_TPB));
    k_dot_lanes<<<grid, DOT_TPB, 0, g_stream>>>(a, b, g.count, g.lanes, g.rows, part);
    k_dot_finish<<<1, DOT_FIN_TPB, 0, g_stream>>>(part, gx, g.lanes, lane32, lane64, total);
}

// ---------------------------------------------------------------- CG kernels
__global__ void k_cg_diag_apply(const float* __restrict__ d, const float* __restrict__ in, float* __restrict__ out,
                                int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = d[t] * in[t];
}

// init (linalg.js:603-610): r = b - Ax, p = 0, Ap = 0
__global__ void k_cg_init(const float* __restrict__ b, const float* __restrict__ ax, float* __restrict__ r,
                          float* __restrict__ p, float* __restrict__ ap, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    r[t] = b[t] - ax[t];
    p[t] = 0.0f;
    ap[t] = 0.0f;
}

// y += s * x (updateX: x += p alpha; updateR with s = -alpha: r -= Ap alpha)
__global__ void k_cg_axpy(float* __restrict__ y, const float* __restrict__ x, float s, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) y[t] = y[t] + x[t] * s;
}

// recomputeR (linalg.js:616-620): r = b - Ax
__global__ void k_cg_sub(const float* __restrict__ b, const float* __restrict__ ax, float* __restrict__ r, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) r[t] = b[t] - ax[t];
}

// updateP (linalg.js:632-636): p = r + p beta
__global__ void k_cg_update_p(float* __restrict__ p, const float* __restrict__ r, float beta, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = r[t] + p[t] * beta;
}

}  // namespace flux

using namespace flux;

// ------------------------------------------------ allocation-free reducer
// The partials and lane sums of flux_dot_noalloc live in this array; each
// device holds its own copy until the process ends.  In host code `scratch`
// names a host-side stand-in: kernels receive the device address obtained
// once per device with hipGetSymbolAddress.
static __device__ __align__(256) double scratch[DOT_SCRATCH_DOUBLES];

static constexpr int SCRATCH_MAX_DEVICES = 64;
static double* s_scratch_ptr[SCRATCH_MAX_DEVICES];

static int scratch_for_current_device(double** out) {
    int dev = 0;
    FLUX_HIP(hipGetDevice(&dev));
    FLUX_REQUIRE(dev >= 0 && dev < SCRATCH_MAX_DEVICES, "dotNoAlloc: device %d out of range", dev);
    if (!s_scratch_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return set_error(FLUX_ERR_HIP, "dotNoAlloc: hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        s_scratch_ptr[dev] = (double*)ptr;
    }
    *out = s_scratch_ptr[dev];
    return FLUX_OK;
}

struct flux_dot {
    const flux_field* a = nullptr;
    const flux_field* b = nullptr;
    DotGeom g;
    unsigned gx = 1;
    flux_field* partial = nullptr;  // lanes FP32 (the reference's `partial`)
    double* part = nullptr;         // per-block per-lane partials
    double* lane64 = nullptr;       // lanes + 1 (the total)
};

struct flux_cg {
    flux_operator* op = nullptr;
    const flux_field* diag = nullptr;
    flux_field* b = nullptr;
    flux_field* x = nullptr;
    int64_t n = 0;
    flux_field *p = nullptr, *r = nullptr, *ap = nullptr, *ax = nullptr;
    flux_dot *dot_rr = nullptr, *dot_pap = nullptr;
    float beta = 0.0f;  // the reference's `beta` array0, read by updateP
};

extern "C" {

int flux_dot_create(int ndim, const int* shape, const flux_field* a, const flux_field* b, flux_dot** out) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(shape && a && b && out, "createDotReducer: NULL argument");
    DotGeom g;
    FLUX_TRY(dot_geometry(ndim, shape, &g, "createDotReducer"));
    FLUX_TRY(check_field(a, FLUX_F32, g.count, "createDotReducer: fieldA"));
    FLUX_TRY(check_field(b, FLUX_F32, g.count, "createDotReducer: fieldB"));
    flux_dot* d = new flux_dot();
    d->a = a;
    d->b = b;
    d->g = g;
    d->gx = dot_row_blocks(g);
    const int64_t lanes64 = g.lanes;
    int rc = flux_field_create(FLUX_F32, 1, &lanes64, &d->partial);
    d->part = (double*)dmalloc(sizeof(double) * (size_t)d->gx * (size_t)g.lanes);
    d->lane64 = (double*)dmalloc(sizeof(double) * ((size_t)g.lanes + 1));
    if (rc || !d->part || !d->lane64) {
        flux_dot_destroy(d);
        return rc ? rc : set_error(FLUX_ERR_NOMEM, "createDotReducer: allocation failed");
    }
    *out = d;
    return FLUX_OK;
}

int flux_dot_dispatch(flux_dot* d) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(d, "createDotReducer: NULL reducer");
    dot_launch(d->g, (const float*)d->a->d, (const float*)d->b->d, d->part, d->gx, (float*)d->partial->d, d->lane64,
               nullptr);
    FLUX_LAUNCH_CHECK();
    return FLUX_OK;
}

int flux_dot_read(flux_dot* d, double* total) {
    FLUX_TRY(flux_dot_dispatch(d));
    double t = 0.0;
    FLUX_HIP(hipMemcpyAsync(&t, d->lane64 + d->g.lanes, sizeof(double), hipMemcpyDeviceToHost, g_stream));
    FLUX_HIP(hipStreamSynchronize(g_stream));
    if (total) *total = t;
    return FLUX_OK;
}

int flux_dot_info(const flux_dot* d, int* lanes, flux_field** partial) {
    FLUX_REQUIRE(d, "createDotReducer: NULL reducer");
    if (lanes) *lanes = d->g.lanes;
    if (partial) *partial = d->partial;
    return FLUX_OK;
}

int flux_dot_noalloc(int ndim, const int* shape, const flux_field* a, const flux_field* b, flux_field* lanes_out,
                     flux_field* total_out) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(shape && a && b && total_out, "dotNoAlloc: NULL argument");
    DotGeom g;
    FLUX_TRY(dot_geometry(ndim, shape, &g, "dotNoAlloc"));
    FLUX_TRY(check_field(a, FLUX_F32, g.count, "dotNoAlloc: fieldA"));
    FLUX_TRY(check_field(b, FLUX_F32, g.count, "dotNoAlloc: fieldB"));
    if (lanes_out) FLUX_TRY(check_field(lanes_out, FLUX_F32, g.lanes, "dotNoAlloc: lanes"));
    FLUX_REQUIRE(total_out->dtype == FLUX_F64 && total_out->count >= 1, "dotNoAlloc: total must be an f64 field");
    FLUX_REQUIRE(dot_fits_scratch(g),
                 "dotNoAlloc: %d lanes exceed the %zu-double scratch (at most %zu lanes); use createDotReducer", g.lanes,
                 DOT_SCRATCH_DOUBLES, (DOT_SCRATCH_DOUBLES - 1) / 2);
    double* base = nullptr;
    FLUX_TRY(scratch_for_current_device(&base));
    const unsigned gx = dot_row_blocks(g);
    double* part = base;
    double* lane64 = base + (size_t)gx * (size_t)g.lanes;
    dot_launch(g, (const float*)a->d, (const float*)b->d, part, gx, lanes_out ? (float*)lanes_out->d : nullptr, lane64,
               (double*)total_out->d);
    FLUX_LAUNCH_CHECK();
    return FLUX_OK;
}

int flux_dot_destroy(flux_dot* d) {
    if (!d) return FLUX_OK;
    if (g_init) (void)hipStreamSynchronize(g_stream);
    if (d->partial) flux_field_destroy(d->partial);
    dfree(d->part);
    dfree(d->lane64);
    delete d;
    return FLUX_OK;
}

// ------------------------------------------------------------------- CG
static int cg_apply(flux_cg* s, const flux_field* in, flux_field* out) {
    if (s->op) return flux_operator_apply(s->op, in, out);
    k_cg_diag_apply<<<blocks_for(s->n, 256), 256, 0, g_stream>>>((const float*)s->diag->d, (const float*)in->d,
                                                                  (float*)out->d, (int)s->n);
    FLUX_LAUNCH_CHECK();
    return FLUX_OK;
}

int flux_cg_destroy(flux_cg* s);

int flux_cg_create(flux_operator* op, const flux_field* diag, flux_field* b, flux_field* x, flux_cg** out) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(b && x && out && (op || diag), "conjugateGradient: NULL argument");
    FLUX_REQUIRE(b->dtype == x->dtype, "conjugateGradient: element type mismatch, b.type(%d) != x.type(%d).", b->dtype,
                 x->dtype);
    FLUX_REQUIRE(b->count == x->count && b->ndim == x->ndim,
                 "conjugateGradient: shape mismatch, b has %lld elements, x has %lld.", (long long)b->count,
                 (long long)x->count);
    FLUX_REQUIRE(b->dtype == FLUX_F32, "conjugateGradient: GPU atomic dot product only supports type \"float\", got %d.",
                 b->dtype);
    if (diag) FLUX_TRY(check_field(diag, FLUX_F32, b->count, "conjugateGradient: diagonal operator"));
    flux_cg* s = new flux_cg();
    s->op = op;
    s->diag = op ? nullptr : diag;
    s->b = b;
    s->x = x;
    s->n = b->count;
    int shape[3] = {(int)b->shape[0], (int)b->shape[1], (int)b->shape[2]};
    int rc = FLUX_OK;
    for (flux_field** f : {&s->p, &s->r, &s->ap, &s->ax})
        if (!rc) rc = flux_field_create(FLUX_F32, b->ndim, b->shape, f);
    if (!rc) rc = flux_dot_create(b->ndim, shape, s->r, s->r, &s->dot_rr);
    if (!rc) rc = flux_dot_create(b->ndim, shape, s->p, s->ap, &s->dot_pap);
    if (rc) {
        flux_cg_destroy(s);
        return rc;
    }
    *out = s;
    return FLUX_OK;
}

static bool degenerate_denominator(double v) { return !std::isfinite(v) || v == 0.0; }

// solve(tol, maxiter) (linalg.js:641-724): absolute tolerance on sqrt(r.r),
// the residual recomputed from x every iteration after the first
// (RESIDUAL_RECOMPUTE_INTERVAL = 1, line 184), alpha and beta handed to the
// kernels as FP32 scalars (setScalar writes a Float32Array); returns the
// reference's final check r.r < tol in *converged.
int flux_cg_solve(flux_cg* s, double tol, int maxiter, int* converged, int* iterations) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(s && converged, "conjugateGradient: NULL argument");
    const unsigned g = blocks_for(s->n, 256);
    const int n = (int)s->n;
    float* p = (float*)s->p->d;
    float* r = (float*)s->r->d;
    float* ap = (float*)s->ap->d;
    float* ax = (float*)s->ax->d;
    float* x = (float*)s->x->d;
    const float* b = (const float*)s->b->d;
    FLUX_TRY(cg_apply(s, s->x, s->ax));
    k_cg_init<<<g, 256, 0, g_stream>>>(b, ax, r, p, ap, n);
    double init_rtr = 0;
    FLUX_TRY(flux_dot_read(s->dot_rr, &init_rtr));
    double old_rtr = init_rtr, new_rtr = init_rtr;
    k_cg_update_p<<<g, 256, 0, g_stream>>>(p, r, s->beta, n);
    bool force = false;
    int it = 0;
    if (std::sqrt(std::fabs(init_rtr)) >= tol) {
        for (int iter = 0; iter < maxiter; ++iter) {
            it = iter + 1;
            FLUX_TRY(cg_apply(s, s->p, s->ap));
            double pap = 0;
            FLUX_TRY(flux_dot_read(s->dot_pap, &pap));
            if (degenerate_denominator(pap)) break;
            const float alpha = (float)(old_rtr / pap);
            k_cg_axpy<<<g, 256, 0, g_stream>>>(x, p, alpha, n);
            if (force || iter > 0) {
                FLUX_TRY(cg_apply(s, s->x, s->ax));
                k_cg_sub<<<g, 256, 0, g_stream>>>(b, ax, r, n);
                force = false;
            } else {
                k_cg_axpy<<<g, 256, 0, g_stream>>>(r, ap, -alpha, n);
            }
            FLUX_TRY(flux_dot_read(s->dot_rr, &new_rtr));
            if (std::sqrt(std::fabs(new_rtr)) < tol) break;
            if (new_rtr > old_rtr) force = true;
            if (degenerate_denominator(old_rtr)) break;
            s->beta = (float)(new_rtr / old_rtr);
            k_cg_update_p<<<g, 256, 0, g_stream>>>(p, r, s->beta, n);
            old_rtr = new_rtr;
        }
    }
    FLUX_LAUNCH_CHECK();
    FLUX_HIP(hipStreamSynchronize(g_stream));
    *converged = new_rtr < tol ? 1 : 0;
    if (iterations) *iterations = it;
    return FLUX_OK;
}

int flux_cg_fields(flux_cg* s, flux_field** p, flux_field** r, flux_field** ap, flux_field** ax) {
    FLUX_REQUIRE(s, "conjugateGradient: NULL solver");
    if (p) *p = s->p;
    if (r) *r = s->r;
    if (ap) *ap = s->ap;
    if (ax) *ax = s->ax;
    return FLUX_OK;
}

int flux_cg_destroy(flux_cg* s) {
    if (!s) return FLUX_OK;
    if (g_init) (void)hipStreamSynchronize(g_stream);
    flux_dot_destroy(s->dot_rr);
    flux_dot_destroy(s->dot_pap);
    for (flux_field* f : {s->p, s->r, s->ap, s->ax})
        if (f) flux_field_destroy(f);
    delete s;
    return FLUX_OK;
}

}  // extern "C"
