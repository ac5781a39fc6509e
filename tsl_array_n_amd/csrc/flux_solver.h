// Internal entry points of the solver translation unit used by the others.
#pragma once
#include "flux_internal.h"

namespace flux {
int fill_device(void* p, int dtype, int64_t n, double v);
int check_shape(int ndim, const int* shape, const double* spacing, int spacing_len, const char* who);
int mg_refresh(flux_multigrid* mg);
// Optional r.z fused into the V-cycle's last finest-level launch: when the
// V-cycle ends with such a launch it stores one partial per block in
// `partials` and sets `blocks`; otherwise `blocks` stays 0.
struct DotOut {
    double* partials = nullptr;
    unsigned capacity = 0;  // partials available; a launch with more blocks is not fused
    unsigned blocks = 0;
};
int mg_apply_raw(flux_multigrid* mg, const float* r, float* z, const int* stop, const double* rr, DotOut* dot);
int64_t mg_cells(const flux_multigrid* mg);
int mg_storage(const flux_multigrid* mg);
int pcg_solve_raw(flux_pcg* s, double tol, int maxiter, int relative, flux_pcg_state* st);
// P2G with the reference's fixed-point atomics on the stream; with `conc`
// each particle's weights are scaled by its density over the heavier one
int p2g_fixed_async(const flux_grid3* g, const flux_field* positions, const flux_field* velocities, int64_t count,
                    int component, float scale, float eps, flux_field* numer, flux_field* denom, flux_field* out,
                    flux_field* valid, const flux_field* conc, float d0, float d1);
// semi-Lagrangian advection without the final stream synchronisation
int advection3_create(const flux_grid3* g, const flux_field* const* vel, float dt, int max_substeps, int order,
                      flux_advection3** out);
int advection3_faces_async(flux_advection3* a, const flux_field* const* in, flux_field* const* out);
int advection3_scalar_async(flux_advection3* a, const flux_field* in, flux_field* out);
void advection3_free(flux_advection3* a);
}  // namespace flux
