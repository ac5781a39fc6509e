// createGridFlipSolver3 for gfx950: the FLIP time step around the pressure
// projection (packages/fluxflow/src/grid/grid_flip_solver3.js:135-1155),
// free-surface form with reduced pressure, static collider, deterministic
// particle resampling (640-793).  Stages are composed from the library's
// own transfer, boundary, extrapolation and projection objects; this unit
// adds the resampling, the Dirichlet fields of the free surface and gravity.
// Built with -ffp-contract=off.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>

#include "flux_grid3.h"
#include "flux_internal.h"
#include "flux_solver.h"

namespace flux {

static constexpr int STPB = 256;

__device__ __forceinline__ int cell_of(const G3& g, const float* __restrict__ pos, int p) {
    const float x = (pos[3 * p + 0] - g.o0) / g.h0;
    const float y = (pos[3 * p + 1] - g.o1) / g.h1;
    const float z = (pos[3 * p + 2] - g.o2) / g.h2;
    const int i = clampi((int)floorf(x), 0, g.n0 - 1);
    const int j = clampi((int)floorf(y), 0, g.n1 - 1);
    const int k = clampi((int)floorf(z), 0, g.n2 - 1);
    return i + g.n0 * (j + g.n1 * k);
}

// countPerCellKernel (grid_flip_solver3.js:675-680): integer atomics, exact
__global__ void k_count_per_cell(G3 g, const float* __restrict__ pos, int count, int* __restrict__ cnt) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < count) atomicAdd(&cnt[cell_of(g, pos, p)], 1);
}

// markDonorsKernel (684-690)
__global__ void k_mark_donors(G3 g, const float* __restrict__ pos, int count, const int* __restrict__ cnt, int maxpc,
                              int* __restrict__ flag) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < count) flag[p] = cnt[cell_of(g, pos, p)] > maxpc ? 1 : 0;
}

// scatterDonorsKernel (693-698): rank among the donors before it
__global__ void k_scatter_donors(const int* __restrict__ flag, const int* __restrict__ scan, int count,
                                 int* __restrict__ pool) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < count && flag[p] == 1) pool[scan[p] - 1] = p;
}

// neighborFluidCount + markNeedKernel (703-727)
__global__ void k_mark_need(G3 g, const int* __restrict__ cnt, int minpc, int* __restrict__ need) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int n0 = g.n0, n1 = g.n1, n2 = g.n2;
    if (c >= n0 * n1 * n2) return;
    const int i = c % n0, j = (c / n0) % n1, k = c / (n0 * n1);
    auto at = [&](int ii, int jj, int kk) { return cnt[ii + n0 * (jj + n1 * kk)]; };
    const int nb = (at(max(0, i - 1), j, k) >= minpc) + (at(min(n0 - 1, i + 1), j, k) >= minpc) +
                   (at(i, max(0, j - 1), k) >= minpc) + (at(i, min(n1 - 1, j + 1), k) >= minpc) +
                   (at(i, j, max(0, k - 1)) >= minpc) + (at(i, j, min(n2 - 1, k + 1)) >= minpc);
    const int count = cnt[c];
    const bool wants = count < minpc && nb >= 3;  // MIN_FLUID_NEIGHBORS
    need[c] = wants ? minpc - count : 0;
}

// claimDonorsKernel (728-760): a cell's slice of the donor pool starts after
// everything earlier cells asked for; a claimed donor moves to the cell
// centre and takes the grid velocity there.
__global__ void k_claim_donors(G3 g, Mac m, const int* __restrict__ need, const int* __restrict__ need_scan,
                               const int* __restrict__ donor_scan, int count, const int* __restrict__ pool, int minpc,
                               float* __restrict__ pos, float* __restrict__ vel) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.n0 * g.n1 * g.n2) return;
    const int needed = need[c];
    if (needed <= 0) return;
    const int i = c % g.n0, j = (c / g.n0) % g.n1, k = c / (g.n0 * g.n1);
    const int base = need_scan[c] - needed;
    const int total = count > 0 ? donor_scan[count - 1] : 0;
    const float co[3] = {g.o0 + g.h0 * 0.5f, g.o1 + g.h1 * 0.5f, g.o2 + g.h2 * 0.5f};
    const float np[3] = {co[0] + (float)i * g.h0, co[1] + (float)j * g.h1, co[2] + (float)k * g.h2};
    int claimed = 0;
    for (int q = 0; q < minpc; ++q) {
        if (claimed >= needed) break;
        const int slot = base + claimed;
        if (slot >= total) break;
        const int d = pool[slot];
        pos[3 * d + 0] = np[0];
        pos[3 * d + 1] = np[1];
        pos[3 * d + 2] = np[2];
        float v[3];
        sample_mac(g, m, np, v);
        vel[3 * d + 0] = v[0];
        vel[3 * d + 1] = v[1];
        vel[3 * d + 2] = v[2];
        ++claimed;
    }
}

// dirichlet(pos) of the FLIP solver (grid_flip_solver3.js:252-259) evaluated
// by updateDirichletFields at cell centres (grid_pressure_solver3.js:171-206):
// air cells (fluid mask < 0.5) are Dirichlet with the reduced-pressure target
// -dt (g . pos); the mask sampled exactly at a cell centre is the cell's own.
__global__ void k_free_surface(G3 g, const float* __restrict__ mask, float gx, float gy, float gz, float dt,
                               int reduced, uint8_t* __restrict__ active, float* __restrict__ target) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.n0 * g.n1 * g.n2) return;
    const int i = c % g.n0, j = (c / g.n0) % g.n1, k = c / (g.n0 * g.n1);
    const float px = (g.o0 + g.h0 * 0.5f) + (float)i * g.h0;
    const float py = (g.o1 + g.h1 * 0.5f) + (float)j * g.h1;
    const float pz = (g.o2 + g.h2 * 0.5f) + (float)k * g.h2;
    active[c] = mask[c] < 0.5f ? 1 : 0;
    target[c] = reduced ? -(((gx * px + gy * py) + gz * pz) * dt) : 0.0f;
}

// applyGravityU/V/W (grid_flip_solver3.js:798-808), the non-reduced form
__global__ void k_add_const(float* __restrict__ d, int n, float v) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f < n) d[f] = d[f] + v;
}

}  // namespace flux

using namespace flux;

struct flux_flip3 {
    flux_flip3_desc d;
    flux_grid3 g;
    int64_t np = 0;
    int64_t cells = 0;
    flux_field* positions = nullptr;
    flux_field* velocities = nullptr;
    flux_field* fluid_mask = nullptr;
    flux_field* vel[3] = {nullptr, nullptr, nullptr};
    flux_field* old[3] = {nullptr, nullptr, nullptr};
    flux_field* numer[3] = {nullptr, nullptr, nullptr};
    flux_field* denom[3] = {nullptr, nullptr, nullptr};
    flux_field* weight_valid[3] = {nullptr, nullptr, nullptr};
    flux_field* fluid_adj[3] = {nullptr, nullptr, nullptr};
    flux_field* active = nullptr;
    flux_field* target = nullptr;
    flux_extrapolator* ex_weight[3] = {nullptr, nullptr, nullptr};
    flux_extrapolator* ex_pressure[3] = {nullptr, nullptr, nullptr};
    flux_boundary3* boundary = nullptr;
    flux_pressure3* pressure = nullptr;
    // resampling
    int* cell_count = nullptr;
    int* donor_flag = nullptr;
    int* donor_scan = nullptr;
    int* donor_pool = nullptr;
    int* need = nullptr;
    int* need_scan = nullptr;
    void* scan_temp = nullptr;
    size_t scan_temp_bytes = 0;
};

namespace {

int make_face(const int n[3], int axis, int dtype, flux_field** f) {
    int64_t s[3] = {n[0] + (axis == 0), n[1] + (axis == 1), n[2] + (axis == 2)};
    return flux_field_create(dtype, 3, s, f);
}

// the resample pass (grid_flip_solver3.js:762-793)
int resample(flux_flip3* s) {
    const G3 g = make_g3(&s->g);
    const int np = (int)s->np, nc = (int)s->cells;
    const float* pos = (const float*)s->positions->d;
    FLUX_HIP(hipMemsetAsync(s->cell_count, 0, sizeof(int) * nc, g_stream));
    if (np > 0) {
        k_count_per_cell<<<blocks_for(np, STPB), STPB, 0, g_stream>>>(g, pos, np, s->cell_count);
        k_mark_donors<<<blocks_for(np, STPB), STPB, 0, g_stream>>>(g, pos, np, s->cell_count,
                                                                   s->d.max_particles_per_cell, s->donor_flag);
        FLUX_HIP(hipcub::DeviceScan::InclusiveSum(s->scan_temp, s->scan_temp_bytes, s->donor_flag, s->donor_scan, np,
                                                  g_stream));
        k_scatter_donors<<<blocks_for(np, STPB), STPB, 0, g_stream>>>(s->donor_flag, s->donor_scan, np, s->donor_pool);
    }
    k_mark_need<<<blocks_for(nc, STPB), STPB, 0, g_stream>>>(g, s->cell_count, s->d.min_particles_per_cell, s->need);
    FLUX_HIP(hipcub::DeviceScan::InclusiveSum(s->scan_temp, s->scan_temp_bytes, s->need, s->need_scan, nc, g_stream));
    Mac m{(const float*)s->vel[0]->d, (const float*)s->vel[1]->d, (const float*)s->vel[2]->d};
    k_claim_donors<<<blocks_for(nc, STPB), STPB, 0, g_stream>>>(g, m, s->need, s->need_scan, s->donor_scan, np,
                                                                s->donor_pool, s->d.min_particles_per_cell,
                                                                (float*)s->positions->d, (float*)s->velocities->d);
    FLUX_LAUNCH_CHECK();
    return FLUX_OK;
}

int mark_fluid_and_adjacency(flux_flip3* s) {
    FLUX_TRY(flux_mark_fluid_cells(&s->g, s->positions, s->np, s->fluid_mask));
    for (int a = 0; a < 3; ++a) FLUX_TRY(flux_fluid_adjacent(&s->g, s->fluid_mask, a, s->fluid_adj[a]));
    return FLUX_OK;
}

}  // namespace

extern "C" {

void flux_flip3_desc_default(flux_flip3_desc* d) {
    memset(d, 0, sizeof(*d));
    d->resolution[0] = d->resolution[1] = d->resolution[2] = 1;
    d->spacing[0] = d->spacing[1] = d->spacing[2] = 1.0;
    d->gravity[0] = 0.0f;
    d->gravity[1] = -9.81f;
    d->gravity[2] = 0.0f;
    d->reduced_pressure = 1;
    d->flip_ratio = 0.97f;
    d->velocity_damping = 0.02f;
    d->p2g_atomic_scale = 65536.0f;
    d->weight_epsilon = 1e-4f;
    d->closed_domain_flags = 63;
    d->resample = 1;
    d->min_particles_per_cell = 4;
    d->max_particles_per_cell = 8;
    d->friction = 0.5f;
    d->extrapolation_depth = 5;
    d->max_substeps = 32;
    flux_pressure3_desc_default(&d->pressure);
    d->pressure.max_plausible_pressure = 0.0;  // derive from the scene
}

int flux_flip3_create(const flux_flip3_desc* d, flux_flip3** out) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(d && out, "createGridFlipSolver3: NULL argument");
    FLUX_REQUIRE(d->max_particles > 0 && d->max_particles < (1ll << 31),
                 "createGridFlipSolver3: options.maxParticles is required.");
    for (int a = 0; a < 3; ++a)
        FLUX_REQUIRE(d->resolution[a] > 0 && d->spacing[a] > 0, "createGridFlipSolver3: bad grid");
    FLUX_REQUIRE(d->min_particles_per_cell >= 0 && d->max_particles_per_cell >= 0, "createGridFlipSolver3: bad resample counts");
    flux_flip3* s = new flux_flip3();
    s->d = *d;
    for (int a = 0; a < 3; ++a) {
        s->g.n[a] = d->resolution[a];
        s->g.h[a] = (float)d->spacing[a];
        s->g.origin[a] = (float)d->origin[a];
    }
    s->np = d->max_particles;
    s->cells = (int64_t)d->resolution[0] * d->resolution[1] * d->resolution[2];
    int rc = FLUX_OK;
    auto ok = [&](int r) {
        if (rc == FLUX_OK && r != FLUX_OK) rc = r;
        return rc == FLUX_OK;
    };
    int64_t pshape[1] = {s->np};
    int64_t cshape[3] = {d->resolution[0], d->resolution[1], d->resolution[2]};
    ok(flux_field_create(FLUX_VEC3, 1, pshape, &s->positions));
    ok(flux_field_create(FLUX_VEC3, 1, pshape, &s->velocities));
    ok(flux_field_create(FLUX_F32, 3, cshape, &s->fluid_mask));
    ok(flux_field_create(FLUX_U8, 3, cshape, &s->active));
    ok(flux_field_create(FLUX_F32, 3, cshape, &s->target));
    for (int a = 0; a < 3 && rc == FLUX_OK; ++a) {
        ok(make_face(d->resolution, a, FLUX_F32, &s->vel[a]));
        ok(make_face(d->resolution, a, FLUX_F32, &s->old[a]));
        ok(make_face(d->resolution, a, FLUX_I32, &s->numer[a]));
        ok(make_face(d->resolution, a, FLUX_I32, &s->denom[a]));
        ok(make_face(d->resolution, a, FLUX_I32, &s->weight_valid[a]));
        ok(make_face(d->resolution, a, FLUX_I32, &s->fluid_adj[a]));
    }
    for (int a = 0; a < 3 && rc == FLUX_OK; ++a) {
        ok(flux_extrapolator3_create(s->vel[a], s->weight_valid[a], s->vel[a], &s->ex_weight[a]));
        ok(flux_extrapolator3_create(s->vel[a], s->fluid_adj[a], s->vel[a], &s->ex_pressure[a]));
    }
    if (rc == FLUX_OK) {
        ok(flux_boundary3_create(&s->g, s->vel, d->collider_sdf, d->collider_weights[0] ? d->collider_weights : nullptr,
                                 d->friction, &s->boundary));
        if (rc == FLUX_OK) ok(flux_boundary3_set_closed(s->boundary, d->closed_domain_flags));
    }
    if (rc == FLUX_OK) {
        flux_pressure3_desc pd = d->pressure;
        for (int a = 0; a < 3; ++a) {
            pd.resolution[a] = d->resolution[a];
            pd.spacing[a] = d->spacing[a];
            pd.origin[a] = d->origin[a];
        }
        pd.dirichlet_active = s->active;
        pd.dirichlet_target = s->target;
        if (!(pd.max_plausible_pressure > 0)) {
            // derivedMaxPlausiblePressure (grid_flip_solver3.js:425-432)
            const double extent = d->resolution[0] * d->spacing[0] + d->resolution[1] * d->spacing[1] +
                                  d->resolution[2] * d->spacing[2];
            const double gmag = std::hypot(std::hypot((double)d->gravity[0], (double)d->gravity[1]),
                                           (double)d->gravity[2]);
            pd.max_plausible_pressure = (d->reduced_pressure && gmag > 0) ? 8.0 * d->dt * gmag * extent : 1e6;
        }
        ok(flux_pressure3_create(&pd, &s->pressure));
    }
    if (rc == FLUX_OK) {
        const int np = (int)s->np, nc = (int)s->cells;
        s->cell_count = (int*)dmalloc(sizeof(int) * nc);
        s->donor_flag = (int*)dmalloc(sizeof(int) * np);
        s->donor_scan = (int*)dmalloc(sizeof(int) * np);
        s->donor_pool = (int*)dmalloc(sizeof(int) * np);
        s->need = (int*)dmalloc(sizeof(int) * nc);
        s->need_scan = (int*)dmalloc(sizeof(int) * nc);
        size_t t1 = 0, t2 = 0;
        (void)hipcub::DeviceScan::InclusiveSum(nullptr, t1, (const int*)nullptr, (int*)nullptr, np, g_stream);
        (void)hipcub::DeviceScan::InclusiveSum(nullptr, t2, (const int*)nullptr, (int*)nullptr, nc, g_stream);
        s->scan_temp_bytes = std::max<size_t>(std::max(t1, t2), 16);
        s->scan_temp = dmalloc(s->scan_temp_bytes);
        if (!s->cell_count || !s->donor_flag || !s->donor_scan || !s->donor_pool || !s->need || !s->need_scan ||
            !s->scan_temp)
            rc = set_error(FLUX_ERR_NOMEM, "createGridFlipSolver3: allocation failed");
    }
    if (rc != FLUX_OK) {
        std::string msg = flux_last_error();
        flux_flip3_destroy(s);
        return set_error(rc, "%s", msg.c_str());
    }
    *out = s;
    return FLUX_OK;
}

// onAdvanceTimeStep (grid_flip_solver3.js:1120-1132)
int flux_flip3_step(flux_flip3* s, flux_pressure_diag* diag) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(s, "onAdvanceTimeStep: NULL solver");
    const flux_flip3_desc& d = s->d;
    // preP2G: advectParticles (933-940)
    FLUX_TRY(flux_advect_particles(&s->g, s->positions, s->np, s->vel, d.dt, d.max_substeps));
    // afterP2G (1083-1114): scatter + finalize, velOld snapshot, weight
    // extrapolation, constrainVelocity, fluid mask, adjacency, resampling
    for (int a = 0; a < 3; ++a)
        FLUX_TRY(flux_p2g_fixed(&s->g, s->positions, s->velocities, s->np, a, d.p2g_atomic_scale, d.weight_epsilon,
                                s->numer[a], s->denom[a], s->vel[a], s->weight_valid[a]));
    for (int a = 0; a < 3; ++a) FLUX_TRY(flux_field_copy(s->old[a], s->vel[a]));
    for (int a = 0; a < 3; ++a) FLUX_TRY(flux_extrapolator3_run(s->ex_weight[a], d.extrapolation_depth));
    FLUX_TRY(flux_boundary3_constrain(s->boundary, d.extrapolation_depth));
    FLUX_TRY(mark_fluid_and_adjacency(s));
    if (d.resample) {
        FLUX_TRY(resample(s));
        FLUX_TRY(mark_fluid_and_adjacency(s));
    }
    if (!d.reduced_pressure) {
        // applyGravityU/V/W; with reduced pressure and no density coupling the
        // reference adds no body force here (gravity enters via the air target)
        for (int a = 0; a < 3; ++a) {
            const int64_t nf = s->vel[a]->count;
            k_add_const<<<blocks_for(nf, STPB), STPB, 0, g_stream>>>((float*)s->vel[a]->d, (int)nf,
                                                                     d.gravity[a] * d.dt);
        }
    }
    FLUX_TRY(flux_boundary3_constrain(s->boundary, d.extrapolation_depth));
    // the free-surface Dirichlet fields, then the projection in place
    k_free_surface<<<blocks_for(s->cells, STPB), STPB, 0, g_stream>>>(
        make_g3(&s->g), (const float*)s->fluid_mask->d, d.gravity[0], d.gravity[1], d.gravity[2], d.dt,
        d.reduced_pressure, (uint8_t*)s->active->d, (float*)s->target->d);
    FLUX_LAUNCH_CHECK();
    flux_pressure_diag pdg;
    FLUX_TRY(flux_pressure3_project(s->pressure, s->vel[0], s->vel[1], s->vel[2], s->vel[0], s->vel[1], s->vel[2],
                                    &pdg));
    if (diag) *diag = pdg;
    // postProject (1115-1119): constrain, extrapolate into the fluid's
    // neighbourhood, G2P
    FLUX_TRY(flux_boundary3_constrain(s->boundary, d.extrapolation_depth));
    for (int a = 0; a < 3; ++a) FLUX_TRY(flux_extrapolator3_run(s->ex_pressure[a], d.extrapolation_depth));
    FLUX_TRY(flux_g2p(&s->g, s->positions, s->velocities, s->np, s->vel, s->old, d.flip_ratio, d.velocity_damping));
    return FLUX_OK;
}

flux_field* flux_flip3_field(flux_flip3* s, int which) {
    if (!s) return nullptr;
    switch (which) {
        case FLUX_FLIP3_POSITIONS: return s->positions;
        case FLUX_FLIP3_VELOCITIES: return s->velocities;
        case FLUX_FLIP3_FLUID_MASK: return s->fluid_mask;
        case FLUX_FLIP3_U: return s->vel[0];
        case FLUX_FLIP3_V: return s->vel[1];
        case FLUX_FLIP3_W: return s->vel[2];
        case FLUX_FLIP3_PRESSURE: return s->pressure ? flux_pressure3_field(s->pressure, FLUX_P3_PRESSURE) : nullptr;
        default: return nullptr;
    }
}

int flux_flip3_destroy(flux_flip3* s) {
    if (!s) return FLUX_OK;
    if (g_init) (void)hipStreamSynchronize(g_stream);
    if (s->pressure) flux_pressure3_destroy(s->pressure);
    if (s->boundary) flux_boundary3_destroy(s->boundary);
    for (int a = 0; a < 3; ++a) {
        flux_extrapolator3_destroy(s->ex_weight[a]);
        flux_extrapolator3_destroy(s->ex_pressure[a]);
        flux_field* fs[] = {s->vel[a], s->old[a], s->numer[a], s->denom[a], s->weight_valid[a], s->fluid_adj[a]};
        for (flux_field* f : fs)
            if (f) flux_field_destroy(f);
    }
    flux_field* fs[] = {s->positions, s->velocities, s->fluid_mask, s->active, s->target};
    for (flux_field* f : fs)
        if (f) flux_field_destroy(f);
    dfree(s->cell_count);
    dfree(s->donor_flag);
    dfree(s->donor_scan);
    dfree(s->donor_pool);
    dfree(s->need);
    dfree(s->need_scan);
    dfree(s->scan_temp);
    delete s;
    return FLUX_OK;
}

}  // extern "C"
