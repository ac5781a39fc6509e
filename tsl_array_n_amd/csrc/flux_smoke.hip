// createGridSmokeSolver3 for gfx950: the Eulerian smoke step around the
// pressure projection (packages/fluxflow/src/grid/grid_smoke_solver3.js:25-154
// over createGridSolver3, grid_solver3.js:45-149): buoyancy as an external
// force (external_force_solver3.js:19-47), constrainVelocity, projection,
// semi-Lagrangian velocity advection (order 1 or 2, through velocityPrev), then
// density and temperature advected through the new velocity (advectScalar3,
// advection_solver3.js:285-357) and decayed into the other half of their
// ping-pong pairs.
// Built with -ffp-contract=off.
#include <cmath>
#include <cstring>
#include <string>

#include "flux_grid3.h"
#include "flux_internal.h"
#include "flux_solver.h"

namespace flux {

static constexpr int KTPB = 256;

// force(pos) * dt added to one face component (dispatchU/V/W of
// createExternalForceSolver3), force = up * (a rho(pos) + b (T(pos) - T_amb))
// with rho, T sampled trilinearly at the face position (buoyancyForce,
// grid_smoke_solver3.js:66-81)
__global__ void k_buoyancy(G3 g, int axis, const float* __restrict__ dens, const float* __restrict__ temp, float a,
                           float bt, float tamb, float up, float dt, float* __restrict__ vel) {
    const int s[3] = {g.n0 + (axis == 0), g.n1 + (axis == 1), g.n2 + (axis == 2)};
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= s[0] * s[1] * s[2]) return;
    const int I[3] = {f % s[0], (f / s[0]) % s[1], f / (s[0] * s[1])};
    const float h[3] = {g.h0, g.h1, g.h2};
    float fo[3];
    mac_origin(g, axis, fo);
    float pos[3];
    for (int q = 0; q < 3; ++q) pos[q] = fo[q] + (float)I[q] * h[q];
    const int n[3] = {g.n0, g.n1, g.n2};
    const float co[3] = {g.o0 + g.h0 * 0.5f, g.o1 + g.h1 * 0.5f, g.o2 + g.h2 * 0.5f};
    const float rho = sample_tri(dens, n, co, h, pos);
    const float t = sample_tri(temp, n, co, h, pos);
    const float fb = a * rho + bt * (t - tamb);
    vel[f] = vel[f] + (up * fb) * dt;
}

// buildDecayKernel (grid_smoke_solver3.js:100-108): state = raw * (1 - decay)
__global__ void k_decay(const float* __restrict__ raw, float* __restrict__ state, int n, float decay) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) state[c] = raw[c] * (1.0f - decay);
}

// a ball source re-injected with max (examples/34-smoke-plume-3d/main.js:146-158)
__global__ void k_source_sphere(G3 g, float* __restrict__ d, float cx, float cy, float cz, float r, float value) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.n0 * g.n1 * g.n2) return;
    const int i = c % g.n0, j = (c / g.n0) % g.n1, k = c / (g.n0 * g.n1);
    const float px = (g.o0 + g.h0 * 0.5f) + (float)i * g.h0;
    const float py = (g.o1 + g.h1 * 0.5f) + (float)j * g.h1;
    const float pz = (g.o2 + g.h2 * 0.5f) + (float)k * g.h2;
    const float dx = px - cx, dy = py - cy, dz = pz - cz;
    const bool in = (dx * dx + dy * dy) + dz * dz < r * r;
    d[c] = fmaxf(d[c], in ? value : 0.0f);
}

// the same max re-injection with a per-cell source value
__global__ void k_source_max(float* __restrict__ d, const float* __restrict__ src, int n) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) d[c] = fmaxf(d[c], src[c]);
}

}  // namespace flux

using namespace flux;

struct flux_smoke3 {
    flux_smoke3_desc d;
    flux_grid3 g;
    int64_t cells = 0;
    flux_field* vel[3] = {nullptr, nullptr, nullptr};
    flux_field* prev[3] = {nullptr, nullptr, nullptr};
    flux_field* dens[2] = {nullptr, nullptr};
    flux_field* temp[2] = {nullptr, nullptr};
    flux_field* raw = nullptr;
    flux_boundary3* boundary = nullptr;
    flux_pressure3* pressure = nullptr;
    flux_advection3* adv_vel = nullptr;     // through velocityPrev (grid_solver3.js:90-91)
    flux_advection3* adv_scalar = nullptr;  // through the new velocity (grid_smoke_solver3.js:92)
    int active = 0;  // index of the current density/temperature state
};

extern "C" {

void flux_smoke3_desc_default(flux_smoke3_desc* d) {
    memset(d, 0, sizeof(*d));
    d->resolution[0] = d->resolution[1] = d->resolution[2] = 1;
    d->spacing[0] = d->spacing[1] = d->spacing[2] = 1.0;
    d->buoyancy_smoke_density_factor = -0.000625f;
    d->buoyancy_temperature_factor = 5.0f;
    d->ambient_temperature = 0.0f;
    d->smoke_decay = 0.001f;
    d->temperature_decay = 0.001f;
    d->up[1] = 1.0f;
    d->closed_domain_flags = 63;
    d->max_substeps = 32;
    d->extrapolation_depth = 5;
    flux_pressure3_desc_default(&d->pressure);
    d->advection_order = 1;
}

int flux_smoke3_create(const flux_smoke3_desc* d, flux_smoke3** out) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(d && out, "createGridSmokeSolver3: NULL argument");
    for (int a = 0; a < 3; ++a)
        FLUX_REQUIRE(d->resolution[a] > 0 && d->spacing[a] > 0, "createGridSmokeSolver3: bad grid");
    flux_smoke3* s = new flux_smoke3();
    s->d = *d;
    for (int a = 0; a < 3; ++a) {
        s->g.n[a] = d->resolution[a];
        s->g.h[a] = (float)d->spacing[a];
        s->g.origin[a] = (float)d->origin[a];
    }
    s->cells = (int64_t)d->resolution[0] * d->resolution[1] * d->resolution[2];
    int rc = FLUX_OK;
    auto ok = [&](int r) {
        if (rc == FLUX_OK && r != FLUX_OK) rc = r;
        return rc == FLUX_OK;
    };
    int64_t cs[3] = {d->resolution[0], d->resolution[1], d->resolution[2]};
    for (int a = 0; a < 3 && rc == FLUX_OK; ++a) {
        int64_t fs[3] = {cs[0] + (a == 0), cs[1] + (a == 1), cs[2] + (a == 2)};
        ok(flux_field_create(FLUX_F32, 3, fs, &s->vel[a]));
        ok(flux_field_create(FLUX_F32, 3, fs, &s->prev[a]));
    }
    for (int q = 0; q < 2 && rc == FLUX_OK; ++q) {
        ok(flux_field_create(FLUX_F32, 3, cs, &s->dens[q]));
        ok(flux_field_create(FLUX_F32, 3, cs, &s->temp[q]));
    }
    ok(flux_field_create(FLUX_F32, 3, cs, &s->raw));
    if (rc == FLUX_OK) {
        const int order = d->advection_order == 0 ? 1 : d->advection_order;
        ok(advection3_create(&s->g, s->prev, d->dt, d->max_substeps, order, &s->adv_vel));
        if (rc == FLUX_OK) ok(advection3_create(&s->g, s->vel, d->dt, d->max_substeps, order, &s->adv_scalar));
    }
    if (rc == FLUX_OK) {
        ok(flux_boundary3_create(&s->g, s->vel, nullptr, nullptr, 0.5f, &s->boundary));
        if (rc == FLUX_OK) ok(flux_boundary3_set_closed(s->boundary, d->closed_domain_flags));
    }
    if (rc == FLUX_OK) {
        flux_pressure3_desc pd = d->pressure;
        for (int a = 0; a < 3; ++a) {
            pd.resolution[a] = d->resolution[a];
            pd.spacing[a] = d->spacing[a];
            pd.origin[a] = d->origin[a];
        }
        ok(flux_pressure3_create(&pd, &s->pressure));
    }
    if (rc != FLUX_OK) {
        std::string msg = flux_last_error();
        flux_smoke3_destroy(s);
        return set_error(rc, "%s", msg.c_str());
    }
    *out = s;
    return FLUX_OK;
}

// onAdvanceTimeStep (grid_smoke_solver3.js:110-130 over grid_solver3.js:128-134)
int flux_smoke3_step(flux_smoke3* s, flux_pressure_diag* diag) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(s, "onAdvanceTimeStep: NULL solver");
    const flux_smoke3_desc& d = s->d;
    const G3 g = make_g3(&s->g);
    const int cur = s->active, nxt = 1 - s->active;
    // external forces (buoyancy), then constrainVelocity
    for (int a = 0; a < 3; ++a) {
        const int64_t nf = s->vel[a]->count;
        k_buoyancy<<<blocks_for(nf, KTPB), KTPB, 0, g_stream>>>(
            g, a, (const float*)s->dens[cur]->d, (const float*)s->temp[cur]->d, d.buoyancy_smoke_density_factor,
            d.buoyancy_temperature_factor, d.ambient_temperature, d.up[a], d.dt, (float*)s->vel[a]->d);
    }
    FLUX_LAUNCH_CHECK();
    FLUX_TRY(flux_boundary3_constrain(s->boundary, d.extrapolation_depth));
    // pressure
    flux_pressure_diag pdg;
    FLUX_TRY(flux_pressure3_project(s->pressure, s->vel[0], s->vel[1], s->vel[2], s->vel[0], s->vel[1], s->vel[2],
                                    &pdg));
    if (diag) *diag = pdg;
    FLUX_TRY(flux_boundary3_constrain(s->boundary, d.extrapolation_depth));
    // advection: velocityPrev = velocity; velocity = advect(velocityPrev)
    for (int a = 0; a < 3; ++a) FLUX_TRY(flux_field_copy(s->prev[a], s->vel[a]));
    FLUX_TRY(advection3_faces_async(s->adv_vel, s->prev, s->vel));
    FLUX_TRY(flux_boundary3_constrain(s->boundary, d.extrapolation_depth));
    // scalars through the new velocity, decayed into the other state
    const unsigned gc = blocks_for(s->cells, KTPB);
    FLUX_TRY(advection3_scalar_async(s->adv_scalar, s->dens[cur], s->raw));
    k_decay<<<gc, KTPB, 0, g_stream>>>((const float*)s->raw->d, (float*)s->dens[nxt]->d, (int)s->cells, d.smoke_decay);
    FLUX_TRY(advection3_scalar_async(s->adv_scalar, s->temp[cur], s->raw));
    k_decay<<<gc, KTPB, 0, g_stream>>>((const float*)s->raw->d, (float*)s->temp[nxt]->d, (int)s->cells,
                                       d.temperature_decay);
    FLUX_LAUNCH_CHECK();
    s->active = nxt;
    return FLUX_OK;
}

int flux_smoke3_source_sphere(flux_smoke3* s, const float* center, float radius, float density, float temperature) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(s && center, "smoke source: NULL argument");
    const G3 g = make_g3(&s->g);
    const unsigned gc = blocks_for(s->cells, KTPB);
    k_source_sphere<<<gc, KTPB, 0, g_stream>>>(g, (float*)s->dens[s->active]->d, center[0], center[1], center[2], radius,
                                               density);
    k_source_sphere<<<gc, KTPB, 0, g_stream>>>(g, (float*)s->temp[s->active]->d, center[0], center[1], center[2], radius,
                                               temperature);
    FLUX_LAUNCH_CHECK();
    return FLUX_OK;
}

int flux_smoke3_source_field(flux_smoke3* s, const flux_field* density, const flux_field* temperature) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(s, "smoke source: NULL solver");
    const flux_field* src[2] = {density, temperature};
    flux_field* dst[2] = {s->dens[s->active], s->temp[s->active]};
    for (int q = 0; q < 2; ++q) {
        if (!src[q]) continue;
        FLUX_REQUIRE(src[q]->dtype == FLUX_F32 && src[q]->count == s->cells,
                     "smoke source: the %s source must be a float field of %lld cells",
                     q ? "temperature" : "density", (long long)s->cells);
    }
    const unsigned gc = blocks_for(s->cells, KTPB);
    for (int q = 0; q < 2; ++q)
        if (src[q]) k_source_max<<<gc, KTPB, 0, g_stream>>>((float*)dst[q]->d, (const float*)src[q]->d, (int)s->cells);
    FLUX_LAUNCH_CHECK();
    return FLUX_OK;
}

flux_field* flux_smoke3_field(flux_smoke3* s, int which) {
    if (!s) return nullptr;
    switch (which) {
        case FLUX_SMOKE3_U: return s->vel[0];
        case FLUX_SMOKE3_V: return s->vel[1];
        case FLUX_SMOKE3_W: return s->vel[2];
        case FLUX_SMOKE3_DENSITY: return s->dens[s->active];
        case FLUX_SMOKE3_TEMPERATURE: return s->temp[s->active];
        case FLUX_SMOKE3_PRESSURE: return s->pressure ? flux_pressure3_field(s->pressure, FLUX_P3_PRESSURE) : nullptr;
        default: return nullptr;
    }
}

int flux_smoke3_destroy(flux_smoke3* s) {
    if (!s) return FLUX_OK;
    if (g_init) (void)hipStreamSynchronize(g_stream);
    if (s->pressure) flux_pressure3_destroy(s->pressure);
    if (s->boundary) flux_boundary3_destroy(s->boundary);
    advection3_free(s->adv_vel);
    advection3_free(s->adv_scalar);
    for (int a = 0; a < 3; ++a) {
        if (s->vel[a]) flux_field_destroy(s->vel[a]);
        if (s->prev[a]) flux_field_destroy(s->prev[a]);
    }
    for (int q = 0; q < 2; ++q) {
        if (s->dens[q]) flux_field_destroy(s->dens[q]);
        if (s->temp[q]) flux_field_destroy(s->temp[q]);
    }
    if (s->raw) flux_field_destroy(s->raw);
    delete s;
    return FLUX_OK;
}

}  // extern "C"
