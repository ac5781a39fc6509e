// createSemiLagrangianAdvectionSolver3 for gfx950 (packages/fluxflow/src/grid/
// advection_solver3.js:67-362): a velocity grid bound at creation,
// advectFaceCentered3 and advectScalar3 of any face / cell grid through it,
// order 1 (semi-Lagrangian, monotone tricubic sample at the midpoint
// back-trace) or order 2 (MacCormack with the min/max clamp).  With a static
// collider (its SDF cell-centred on the velocity grid) the back-trace stops
// at the collider, points inside it keep the output's value, and the order-2
// clamp ranges over the corners outside it.
//
// Order 2 runs in two launches instead of the reference's three: the forward
// launch also gathers the clamp range (min/max of the input over the 8
// trilinear corners of the forward trace, advection_solver3.js:90-128), since
// it already holds that trace; the second launch traces backward (direction
// -1), samples the forward result there and applies the correction and the
// clamp -- the backward value of a sample point is only ever read by that
// same point's correction, so it never goes to memory.
// Built with -ffp-contract=off: the arithmetic rounds like the oracle's.
#include <cmath>

#include "flux_grid3.h"
#include "flux_internal.h"
#include "flux_solver.h"

struct flux_advection3 {
    flux_grid3 g{};
    const flux_field* vel[3] = {nullptr, nullptr, nullptr};
    float dt = 0.0f;
    int max_sub = 32;
    int order = 1;
    const flux_field* collider = nullptr;  // the collider's SDF on the cells, or none
    float* fwd = nullptr;      // order 2: forward result, one slot per sample point
    float2* range = nullptr;   // order 2: clamp range per sample point
    int64_t scratch = 0;       // slots in fwd / range
};

namespace flux {

static constexpr int ATPB = 256;

// Up to three sample grids advected in one launch: the components of a MAC
// grid (shape n + e_a at that component's data origin) or one cell-centred
// grid.  Points are numbered grid after grid, first index fastest.
struct SampleSet {
    int count;
    int start[4];  // start[count] = total points
    int s[3][3];
    float o[3][3];
    const float* data[3];
    const float* fwd[3];
    float* out[3];
    float2* range[3];
};

// per-grid members chosen by selects, so the kernel-argument arrays are never
// indexed by a lane-varying value (which would spill them to scratch)
template <typename T>
__device__ __forceinline__ T pick(const T (&v)[3], int c) {
    return c == 0 ? v[0] : (c == 1 ? v[1] : v[2]);
}

__device__ __forceinline__ int locate(const SampleSet& S, int t, int s[3], float o[3], const G3& g, float pos[3],
                                      int& f) {
    int c = 0;
    if (S.count > 1 && t >= S.start[1]) c = 1;
    if (S.count > 2 && t >= S.start[2]) c = 2;
    f = t - (c == 0 ? S.start[0] : (c == 1 ? S.start[1] : S.start[2]));
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        s[a] = c == 0 ? S.s[0][a] : (c == 1 ? S.s[1][a] : S.s[2][a]);
        o[a] = c == 0 ? S.o[0][a] : (c == 1 ? S.o[1][a] : S.o[2][a]);
    }
    const int row = f / s[0];
    const int i = f - row * s[0];
    const int k = row / s[1];
    const int j = row - k * s[1];
    pos[0] = o[0] + (float)i * g.h0;
    pos[1] = o[1] + (float)j * g.h1;
    pos[2] = o[2] + (float)k * g.h2;
    return c;
}

// order 1 (advection_solver3.js:142-176, 287-298): the monotone tricubic
// sample of the data at the back-trace through the velocity grid
__global__ void __launch_bounds__(ATPB) k_adv_order1(G3 g, Mac vel, Sdf3 col, SampleSet S, float dt, int max_sub) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= S.start[S.count]) return;
    int s[3], f;
    float o[3], pos[3], tr[3];
    const int c = locate(S, t, s, o, g, pos, f);
    if (!(sdf_sample(col, pos) > 0.0f)) return;  // inside the collider: the output keeps its value
    const float h[3] = {g.h0, g.h1, g.h2};
    back_trace_c(g, vel, col, pos, dt, fminf(g.h0, fminf(g.h1, g.h2)), max_sub, 1.0f, tr);
    pick(S.out, c)[f] = sample_cubic(pick(S.data, c), s, o, h, tr);
}

// order 2, forward (makeForward, advection_solver3.js:195-222; the scalar
// form 320-334) and the clamp range of the same trace (gatherLocalMinMax,
// 95-120): min / max of the data over the trilinear corners
// With the collider a point inside it forwards its own input value (and has
// no range: its correction is skipped), and only corners whose own positions
// lie outside the collider enter the range; with none the range stays
// empty (1e30, -1e30), so the correction falls back to fwd, as
// clampCorrection's haveValid does.
__global__ void __launch_bounds__(ATPB) k_mc_forward(G3 g, Mac vel, Sdf3 col, SampleSet S, float dt, int max_sub) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= S.start[S.count]) return;
    int s[3], f;
    float o[3], pos[3], tr[3];
    const int c = locate(S, t, s, o, g, pos, f);
    const float h[3] = {g.h0, g.h1, g.h2};
    const float* __restrict__ d = pick(S.data, c);
    if (!(sdf_sample(col, pos) > 0.0f)) {
        pick(S.out, c)[f] = d[f];
        return;
    }
    back_trace_c(g, vel, col, pos, dt, fminf(g.h0, fminf(g.h1, g.h2)), max_sub, 1.0f, tr);
    pick(S.out, c)[f] = sample_cubic(d, s, o, h, tr);
    Tri T;
    tri_weights(tr, o, h, s, T);
    float lo = 1e30f, hi = -1e30f;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int i = (q & 1) ? T.i1 : T.i0, j = (q & 2) ? T.j1 : T.j0, k = (q & 4) ? T.k1 : T.k0;
        if (col.d) {
            const float cp[3] = {o[0] + (float)i * h[0], o[1] + (float)j * h[1], o[2] + (float)k * h[2]};
            if (!(sdf_sample(col, cp) > 0.0f)) continue;
        }
        const float v = d[i + s[0] * (j + s[1] * k)];
        lo = fminf(lo, v);
        hi = fmaxf(hi, v);
    }
    pick(S.range, c)[f] = make_float2(lo, hi);
}

// order 2, backward + correct + clamp (makeBackward 224-247 and
// makeCorrectAndClamp 256-274; clampCorrection 122-128): the forward result
// sampled at the trace with direction -1, fwd + (in - bwd) / 2 kept where it
// lies within the range, else fwd
__global__ void __launch_bounds__(ATPB) k_mc_correct(G3 g, Mac vel, Sdf3 col, SampleSet S, float dt, int max_sub) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= S.start[S.count]) return;
    int s[3], f;
    float o[3], pos[3], tr[3];
    const int c = locate(S, t, s, o, g, pos, f);
    if (!(sdf_sample(col, pos) > 0.0f)) return;  // inside the collider: the output keeps its value
    const float h[3] = {g.h0, g.h1, g.h2};
    const float* __restrict__ fw = pick(S.fwd, c);
    back_trace_c(g, vel, col, pos, dt, fminf(g.h0, fminf(g.h1, g.h2)), max_sub, -1.0f, tr);
    const float bwd = sample_cubic(fw, s, o, h, tr);
    const float fv = fw[f];
    const float corrected = fv + (pick(S.data, c)[f] - bwd) * 0.5f;
    const float2 r = pick(S.range, c)[f];
    pick(S.out, c)[f] = (corrected >= r.x && corrected <= r.y) ? corrected : fv;
}

static Mac vel_of(const flux_advection3* a) {
    return Mac{(const float*)a->vel[0]->d, (const float*)a->vel[1]->d, (const float*)a->vel[2]->d};
}

static bool aliases(const flux_advection3* a, const void* p, const flux_field* const* in, int n_in) {
    for (int q = 0; q < 3; ++q)
        if (a->vel[q]->d == p) return true;
    for (int q = 0; q < n_in; ++q)
        if (in[q]->d == p) return true;
    return false;
}

static int run(flux_advection3* a, SampleSet S) {
    const int total = S.start[S.count];
    if (total == 0) return FLUX_OK;
    const G3 g = make_g3(&a->g);
    const Mac m = vel_of(a);
    const Sdf3 col = sdf3_on(g, a->collider ? (const float*)a->collider->d : nullptr);
    const unsigned nb = blocks_for(total, ATPB);
    if (a->order == 1) {
        k_adv_order1<<<nb, ATPB, 0, g_stream>>>(g, m, col, S, a->dt, a->max_sub);
    } else {
        FLUX_REQUIRE(total <= a->scratch, "advection: %d points exceed the scratch of %lld", total,
                     (long long)a->scratch);
        float* out[3] = {S.out[0], S.out[1], S.out[2]};
        for (int c = 0; c < S.count; ++c) {
            S.out[c] = a->fwd + S.start[c];
            S.fwd[c] = a->fwd + S.start[c];
            S.range[c] = a->range + S.start[c];
        }
        k_mc_forward<<<nb, ATPB, 0, g_stream>>>(g, m, col, S, a->dt, a->max_sub);
        for (int c = 0; c < S.count; ++c) S.out[c] = out[c];
        k_mc_correct<<<nb, ATPB, 0, g_stream>>>(g, m, col, S, a->dt, a->max_sub);
    }
    FLUX_LAUNCH_CHECK();
    return FLUX_OK;
}

int advection3_faces_async(flux_advection3* a, const flux_field* const* in, flux_field* const* out) {
    FLUX_REQUIRE(a && in && out, "advectFaceCentered3: NULL argument");
    SampleSet S{};
    S.count = 3;
    S.start[0] = 0;
    const G3 g = make_g3(&a->g);
    for (int c = 0; c < 3; ++c) {
        FLUX_TRY(check_field(in[c], FLUX_F32, faces_of(a->g.n, c), "advectFaceCentered3: input"));
        FLUX_TRY(check_field(out[c], FLUX_F32, faces_of(a->g.n, c), "advectFaceCentered3: output"));
    }
    for (int c = 0; c < 3; ++c) {
        const int64_t nf = faces_of(a->g.n, c);
        FLUX_REQUIRE(!aliases(a, out[c]->d, in, 3), "advectFaceCentered3: output must not alias the input or velocity");
        for (int q = 0; q < 3; ++q) S.s[c][q] = a->g.n[q] + (q == c);
        float o[3];
        o[0] = (c != 0) ? g.o0 + g.h0 * 0.5f : g.o0;
        o[1] = (c != 1) ? g.o1 + g.h1 * 0.5f : g.o1;
        o[2] = (c != 2) ? g.o2 + g.h2 * 0.5f : g.o2;
        for (int q = 0; q < 3; ++q) S.o[c][q] = o[q];
        S.data[c] = (const float*)in[c]->d;
        S.out[c] = (float*)out[c]->d;
        S.start[c + 1] = S.start[c] + (int)nf;
    }
    return run(a, S);
}

int advection3_scalar_async(flux_advection3* a, const flux_field* in, flux_field* out) {
    FLUX_REQUIRE(a && in && out, "advectScalar3: NULL argument");
    const int64_t nc = cells_of(a->g.n);
    FLUX_TRY(check_field(in, FLUX_F32, nc, "advectScalar3: input"));
    FLUX_TRY(check_field(out, FLUX_F32, nc, "advectScalar3: output"));
    FLUX_REQUIRE(!aliases(a, out->d, &in, 1), "advectScalar3: output must not alias the input or velocity");
    const G3 g = make_g3(&a->g);
    SampleSet S{};
    S.count = 1;
    S.start[0] = 0;
    S.start[1] = (int)nc;
    for (int q = 0; q < 3; ++q) S.s[0][q] = a->g.n[q];
    S.o[0][0] = g.o0 + g.h0 * 0.5f;
    S.o[0][1] = g.o1 + g.h1 * 0.5f;
    S.o[0][2] = g.o2 + g.h2 * 0.5f;
    S.data[0] = (const float*)in->d;
    S.out[0] = (float*)out->d;
    return run(a, S);
}

int advection3_create(const flux_grid3* g, const flux_field* const* vel, float dt, int max_substeps, int order,
                      flux_advection3** out) {
    FLUX_REQUIRE(g && vel && out, "createSemiLagrangianAdvectionSolver3: NULL argument");
    FLUX_REQUIRE(order == 1 || order == 2, "createSemiLagrangianAdvectionSolver3: order must be 1 or 2");
    FLUX_REQUIRE(max_substeps >= 1 && max_substeps <= 1024,
                 "createSemiLagrangianAdvectionSolver3: maxSubsteps out of range");
    for (int q = 0; q < 3; ++q) {
        FLUX_REQUIRE(g->n[q] >= 1 && g->h[q] > 0.0f, "createSemiLagrangianAdvectionSolver3: bad grid");
        FLUX_TRY(check_field(vel[q], FLUX_F32, faces_of(g->n, q), "createSemiLagrangianAdvectionSolver3: velocity"));
    }
    const int64_t slots = faces_of(g->n, 0) + faces_of(g->n, 1) + faces_of(g->n, 2);
    FLUX_REQUIRE(slots < (1ll << 31), "createSemiLagrangianAdvectionSolver3: grid too large");
    flux_advection3* a = new flux_advection3();
    a->g = *g;
    for (int q = 0; q < 3; ++q) a->vel[q] = vel[q];
    a->dt = dt;
    a->max_sub = max_substeps;
    a->order = order;
    if (order == 2) {
        a->fwd = (float*)dmalloc(sizeof(float) * slots);
        a->range = (float2*)dmalloc(sizeof(float2) * slots);
        a->scratch = slots;
        if (!a->fwd || !a->range) {
            if (a->fwd) dfree(a->fwd);
            if (a->range) dfree(a->range);
            delete a;
            return set_error(FLUX_ERR_HIP, "createSemiLagrangianAdvectionSolver3: out of device memory");
        }
    }
    *out = a;
    return FLUX_OK;
}

void advection3_free(flux_advection3* a) {
    if (!a) return;
    if (g_init) (void)hipStreamSynchronize(g_stream);
    if (a->fwd) dfree(a->fwd);
    if (a->range) dfree(a->range);
    delete a;
}

}  // namespace flux

using namespace flux;

extern "C" {

int flux_advection3_create(const flux_grid3* g, const flux_field* const* velocity, float dt, int max_substeps,
                           int order, flux_advection3** out) {
    FLUX_REQUIRE_INIT();
    return advection3_create(g, velocity, dt, max_substeps, order, out);
}

int flux_advection3_set_dt(flux_advection3* a, float dt) {
    FLUX_REQUIRE(a, "advection: NULL solver");
    a->dt = dt;
    return FLUX_OK;
}

int flux_advection3_set_collider(flux_advection3* a, const flux_field* collider_sdf) {
    FLUX_REQUIRE(a, "advection: NULL solver");
    if (collider_sdf)
        FLUX_TRY(check_field(collider_sdf, FLUX_F32, cells_of(a->g.n), "createSemiLagrangianAdvectionSolver3: collider"));
    a->collider = collider_sdf;
    return FLUX_OK;
}

int flux_advection3_faces(flux_advection3* a, const flux_field* const* in, flux_field* const* out) {
    FLUX_REQUIRE_INIT();
    FLUX_TRY(advection3_faces_async(a, in, out));
    FLUX_HIP(hipStreamSynchronize(g_stream));
    return FLUX_OK;
}

int flux_advection3_scalar(flux_advection3* a, const flux_field* in, flux_field* out) {
    FLUX_REQUIRE_INIT();
    FLUX_TRY(advection3_scalar_async(a, in, out));
    FLUX_HIP(hipStreamSynchronize(g_stream));
    return FLUX_OK;
}

int flux_advection3_destroy(flux_advection3* a) {
    advection3_free(a);
    return FLUX_OK;
}

}  // extern "C"
