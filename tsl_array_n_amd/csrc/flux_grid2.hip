// The 2-D Eulerian step around the pressure projection, for gfx950:
// createGridSolver2 (packages/fluxflow/src/grid/grid_solver2.js:131-255) and
// its stages -- polygon SDF colliders / inflows / outflows
// (sdf_collider2.js, sdf_inflow_outflow2.js, polygon_sdf.js),
// semi-Lagrangian face advection stopped at the collider
// (advection_solver2.js), the blocked boundary-condition solver
// (grid_blocked_boundary_condition_solver2.js), the outflow velocity
// condition (grid_outflow_solver2.js), the 2-D pressure solver
// (grid_pressure_solver2.js) and the max-|u| reduction of the CFL sub-step
// count (reduction.js, time/cfl.js).
//
// Built with -ffp-contract=off: every product and sum rounds like the
// reference's f32 shader arithmetic.  A face grid of n cells is U [n0+1, n1]
// (origin (o.x, o.y + h.y/2)) and V [n0, n1+1] (origin (o.x + h.x/2, o.y)),
// first index fastest (grid_data2.js:161-178).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "flux_grid3.h"
#include "flux_internal.h"
#include "flux_solver.h"

namespace flux {

static constexpr int GTPB = 256;

struct G2 {
    int n0, n1;
    float h0, h1, o0, o1;
};

inline G2 make_g2(const flux_grid2* g) { return G2{g->n[0], g->n[1], g->h[0], g->h[1], g->origin[0], g->origin[1]}; }

// bilinearCoordsAndWeights2 + collocatedValueAtPosition2 (grid_math.js:35-76)
__device__ __forceinline__ float bilinear2(const float* __restrict__ d, int s0, int s1, float o0, float o1, float h0,
                                           float h1, float px, float py) {
    const float gx = (px - o0) / h0, gy = (py - o1) / h1;
    const int i0 = (int)floorf(gx), j0 = (int)floorf(gy);
    const float fx = gx - (float)i0, fy = gy - (float)j0;
    const int i0c = clampi(i0, 0, s0 - 1), i1c = clampi(i0 + 1, 0, s0 - 1);
    const int j0c = clampi(j0, 0, s1 - 1), j1c = clampi(j0 + 1, 0, s1 - 1);
    const float ax = 1.0f - fx, ay = 1.0f - fy;
    const float w00 = ax * ay, w10 = fx * ay, w01 = ax * fy, w11 = fx * fy;
    const int r0 = s0 * j0c, r1 = s0 * j1c;
    return ((d[i0c + r0] * w00 + d[i1c + r0] * w10) + d[i0c + r1] * w01) + d[i1c + r1] * w11;
}

// the cell-centred SDF of the grid (sdf_collider2.js:35-45)
__device__ __forceinline__ float sdf_at(const G2& g, const float* __restrict__ sdf, float px, float py) {
    return bilinear2(sdf, g.n0, g.n1, g.o0 + g.h0 * 0.5f, g.o1 + g.h1 * 0.5f, g.h0, g.h1, px, py);
}

// faceCenteredValueAtPosition2 (grid_math.js:78-88)
__device__ __forceinline__ void mac2(const G2& g, const float* __restrict__ u, const float* __restrict__ v, float px,
                                     float py, float out[2]) {
    out[0] = bilinear2(u, g.n0 + 1, g.n1, g.o0 + 0.0f, g.o1 + g.h1 * 0.5f, g.h0, g.h1, px, py);
    out[1] = bilinear2(v, g.n0, g.n1 + 1, g.o0 + g.h0 * 0.5f, g.o1 + 0.0f, g.h0, g.h1, px, py);
}

// face position (dataPosition, grid_data2.js:9-15)
__device__ __forceinline__ void face_pos2(const G2& g, int axis, int i, int j, float& px, float& py) {
    const float ox = axis == 0 ? g.o0 + 0.0f : g.o0 + g.h0 * 0.5f;
    const float oy = axis == 0 ? g.o1 + g.h1 * 0.5f : g.o1 + 0.0f;
    px = ox + g.h0 * (float)i;
    py = oy + g.h1 * (float)j;
}

// bilinearGradientAtPosition2 (grid_math.js:285-328)
__device__ __forceinline__ void grad_axis2(float coord, int n, int& lo, int& hi, float& t) {
    const int idx = (int)floorf(coord);
    int l = idx < n - 2 ? idx : n - 2;
    l = l > 0 ? l : 0;
    hi = l + 1 < n - 1 ? l + 1 : n - 1;
    lo = l;
    const float tt = coord - (float)l;
    t = tt < 0.0f ? 0.0f : (tt > 1.0f ? 1.0f : tt);
}

__device__ __forceinline__ void sdf_grad2(const G2& g, const float* __restrict__ d, float px, float py, float out[2]) {
    int xl, xh, yl, yh;
    float tx, ty;
    grad_axis2((px - (g.o0 + g.h0 * 0.5f)) / g.h0, g.n0, xl, xh, tx);
    grad_axis2((py - (g.o1 + g.h1 * 0.5f)) / g.h1, g.n1, yl, yh, ty);
    const float v00 = d[xl + g.n0 * yl], v10 = d[xh + g.n0 * yl];
    const float v01 = d[xl + g.n0 * yh], v11 = d[xh + g.n0 * yh];
    out[0] = ((1.0f - ty) * (v10 - v00) + ty * (v11 - v01)) / g.h0;
    out[1] = ((1.0f - tx) * (v01 - v00) + tx * (v11 - v10)) / g.h1;
}

// collocatedCubicValueAtPosition2 (grid_math.js:129-167), monotone rows then the column
__device__ __forceinline__ float cubic2(const float* __restrict__ d, int s0, int s1, float o0, float o1, float h0,
                                        float h1, float px, float py) {
    const float cx = (px - o0) / h0, cy = (py - o1) / h1;
    const int ix = (int)floorf(cx), iy = (int)floorf(cy);
    const float fx = cx - (float)ix, fy = cy - (float)iy;
    const int xi[4] = {clampi(ix - 1, 0, s0 - 1), clampi(ix, 0, s0 - 1), clampi(ix + 1, 0, s0 - 1),
                       clampi(ix + 2, 0, s0 - 1)};
    float rv[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const int row = s0 * clampi(iy - 1 + b, 0, s1 - 1);
        rv[b] = mono_cubic(d[xi[0] + row], d[xi[1] + row], d[xi[2] + row], d[xi[3] + row], fx);
    }
    return mono_cubic(rv[0], rv[1], rv[2], rv[3], fy);
}

// backTrace with the collider (advection_solver2.js:81-127)
__device__ void back_trace2(const G2& g, const float* __restrict__ u, const float* __restrict__ v,
                            const float* __restrict__ sdf, float sx, float sy, float dt, int max_sub, float& ox,
                            float& oy, float direction = 1.0f) {
    const float hmin = fminf(g.h0, g.h1);
    float px = sx, py = sy, rem = dt;
    for (int s = 0; s < max_sub; ++s) {
        if (rem <= 1e-6f) break;
        float v0[2], mv[2];
        mac2(g, u, v, px, py, v0);
        const float len = sqrtf(v0[0] * v0[0] + v0[1] * v0[1]);
        float ns = ceilf((len * rem) / hmin);
        if (ns < 1.0f) ns = 1.0f;
        const float sub = rem / ns;
        const float half = sub * 0.5f;
        const float mx = px - (v0[0] * half) * direction, my = py - (v0[1] * half) * direction;
        mac2(g, u, v, mx, my, mv);
        const float nx = px - (mv[0] * sub) * direction, ny = py - (mv[1] * sub) * direction;
        const float phi0 = sdf ? sdf_at(g, sdf, px, py) : 1.0f;
        const float phi1 = sdf ? sdf_at(g, sdf, nx, ny) : 1.0f;
        if (phi1 <= 0.0f) {
            const float w = fabsf(phi1) / (fabsf(phi0) + fabsf(phi1));
            px = px * w + nx * (1.0f - w);
            py = py * w + ny * (1.0f - w);
            break;
        }
        rem = rem - sub;
        px = nx;
        py = ny;
    }
    ox = px;
    oy = py;
}

// fractionInsideSdf (level_set_utils.js:14-30)
__device__ __forceinline__ float fraction_inside_sdf(float phi0, float phi1) {
    const bool in0 = phi0 < 0.0f, in1 = phi1 < 0.0f;
    if (in0 && in1) return 1.0f;
    if (in0) return phi0 / (phi0 - phi1);
    if (in1) return phi1 / (phi1 - phi0);
    return 0.0f;
}

// ------------------------------------------------------------------ kernels
// addPolygons (sdf_collider2.js:82-100) with polygonsSignedDistance
// (polygon_sdf.js:24-97), in FP64 per cell like the reference's host loop.
__global__ void k_polygon_sdf2(int n0, int n1, double h0, double h1, double o0, double o1,
                               const double* __restrict__ xy, const int* __restrict__ starts, int rings, int invert,
                               float* __restrict__ sdf) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n0 * n1) return;
    const int i = c % n0, j = c / n0;
    const double x = (o0 + 0.5 * h0) + i * h0, y = (o1 + 0.5 * h1) + j * h1;
    double best = INFINITY;
    for (int r = 0; r < rings; ++r) {
        const double* ring = xy + 2 * starts[r];
        const int count = starts[r + 1] - starts[r];
        double dist = INFINITY;
        bool inside = false;
        for (int a = 0, b = count - 1; a < count; b = a++) {
            const double xa = ring[2 * a], ya = ring[2 * a + 1], xb = ring[2 * b], yb = ring[2 * b + 1];
            if (((ya > y) != (yb > y)) && (x < (xb - xa) * (y - ya) / (yb - ya) + xa)) inside = !inside;
            const double dx = xa - xb, dy = ya - yb;
            const double len2 = dx * dx + dy * dy;
            double t = len2 == 0 ? 0 : ((x - xb) * dx + (y - yb) * dy) / len2;
            t = t > 1 ? 1 : t;
            t = t < 0 ? 0 : t;
            const double d = hypot(x - (xb + t * dx), y - (yb + t * dy));
            if (d < dist) dist = d;
        }
        const double s = inside ? -dist : dist;
        if (s < best) best = s;
    }
    sdf[c] = (float)(invert ? -best : best);
}

// advectFaceCentered2, order 1 (advection_solver2.js:263-320): U faces then
// V faces in one grid; the trace follows the velocity (u, v), the value is
// the bicubic sample of the advected grid (du, dv); a face inside the
// collider keeps the output's value.
__global__ void k_advect_face2(G2 g, const float* __restrict__ u, const float* __restrict__ v,
                               const float* __restrict__ du, const float* __restrict__ dv,
                               const float* __restrict__ sdf, float dt, int max_sub, float* __restrict__ uo,
                               float* __restrict__ vo) {
    const int nu = (g.n0 + 1) * g.n1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nu + g.n0 * (g.n1 + 1)) return;
    const int axis = t < nu ? 0 : 1;
    const int f = axis == 0 ? t : t - nu;
    const int s0 = g.n0 + (axis == 0);
    const int i = f % s0, j = f / s0;
    float px, py;
    face_pos2(g, axis, i, j, px, py);
    if (sdf && !(sdf_at(g, sdf, px, py) > 0.0f)) return;
    float tx, ty;
    back_trace2(g, u, v, sdf, px, py, dt, max_sub, tx, ty);
    if (axis == 0)
        uo[f] = cubic2(du, g.n0 + 1, g.n1, g.o0 + 0.0f, g.o1 + g.h1 * 0.5f, g.h0, g.h1, tx, ty);
    else
        vo[f] = cubic2(dv, g.n0, g.n1 + 1, g.o0 + g.h0 * 0.5f, g.o1 + 0.0f, g.h0, g.h1, tx, ty);
}

// the data grid of face component `axis`: shape and data origin
__device__ __forceinline__ void face_grid2(const G2& g, int axis, int& s0, int& s1, float& ox, float& oy) {
    s0 = g.n0 + (axis == 0);
    s1 = g.n1 + (axis == 1);
    ox = axis == 0 ? g.o0 + 0.0f : g.o0 + g.h0 * 0.5f;
    oy = axis == 0 ? g.o1 + g.h1 * 0.5f : g.o1 + 0.0f;
}

// advectFaceCentered2, order 2 (MacCormack), forward (advection_solver2.js:
// 322-352): order 1, the input inside the collider; with the clamp range of
// the same trace (gatherLocalMinMax, 221-248): min / max of the input over
// the 4 bilinear corners whose face positions lie outside the collider
// (lo > hi when none does, so the correction is never kept)
__global__ void k_mc_forward2(G2 g, const float* __restrict__ u, const float* __restrict__ v,
                              const float* __restrict__ du, const float* __restrict__ dv,
                              const float* __restrict__ sdf, float dt, int max_sub, float* __restrict__ fu,
                              float* __restrict__ fv, float2* __restrict__ ru, float2* __restrict__ rv) {
    const int nu = (g.n0 + 1) * g.n1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nu + g.n0 * (g.n1 + 1)) return;
    const int axis = t < nu ? 0 : 1;
    const int f = axis == 0 ? t : t - nu;
    int s0, s1;
    float ox, oy;
    face_grid2(g, axis, s0, s1, ox, oy);
    const float* __restrict__ d = axis == 0 ? du : dv;
    float* __restrict__ fo = axis == 0 ? fu : fv;
    const int i = f % s0, j = f / s0;
    float px, py;
    face_pos2(g, axis, i, j, px, py);
    if (sdf && !(sdf_at(g, sdf, px, py) > 0.0f)) {
        fo[f] = d[f];
        return;
    }
    float tx, ty;
    back_trace2(g, u, v, sdf, px, py, dt, max_sub, tx, ty);
    fo[f] = cubic2(d, s0, s1, ox, oy, g.h0, g.h1, tx, ty);
    const float gx = (tx - ox) / g.h0, gy = (ty - oy) / g.h1;
    const int i0 = (int)floorf(gx), j0 = (int)floorf(gy);
    const int ci[2] = {clampi(i0, 0, s0 - 1), clampi(i0 + 1, 0, s0 - 1)};
    const int cj[2] = {clampi(j0, 0, s1 - 1), clampi(j0 + 1, 0, s1 - 1)};
    float lo = 1e30f, hi = -1e30f;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int ni = ci[c & 1], nj = cj[c >> 1];
        const bool valid = sdf ? sdf_at(g, sdf, ox + g.h0 * (float)ni, oy + g.h1 * (float)nj) > 0.0f : true;
        const float val = d[ni + s0 * nj];
        if (valid) {
            lo = fminf(lo, val);
            hi = fmaxf(hi, val);
        }
    }
    (axis == 0 ? ru : rv)[f] = make_float2(lo, hi);
}

// order 2, backward + correct + clamp (354-428): the forward result sampled
// at the trace with direction -1; fwd + (in - bwd) / 2 where it lies in the
// range, else fwd; a face inside the collider keeps the output's value
__global__ void k_mc_correct2(G2 g, const float* __restrict__ u, const float* __restrict__ v,
                              const float* __restrict__ du, const float* __restrict__ dv,
                              const float* __restrict__ sdf, float dt, int max_sub, const float* __restrict__ fu,
                              const float* __restrict__ fv, const float2* __restrict__ ru,
                              const float2* __restrict__ rv, float* __restrict__ uo, float* __restrict__ vo) {
    const int nu = (g.n0 + 1) * g.n1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nu + g.n0 * (g.n1 + 1)) return;
    const int axis = t < nu ? 0 : 1;
    const int f = axis == 0 ? t : t - nu;
    int s0, s1;
    float ox, oy;
    face_grid2(g, axis, s0, s1, ox, oy);
    const int i = f % s0, j = f / s0;
    float px, py;
    face_pos2(g, axis, i, j, px, py);
    if (sdf && !(sdf_at(g, sdf, px, py) > 0.0f)) return;
    const float* __restrict__ fw = axis == 0 ? fu : fv;
    float tx, ty;
    back_trace2(g, u, v, sdf, px, py, dt, max_sub, tx, ty, -1.0f);
    const float bwd = cubic2(fw, s0, s1, ox, oy, g.h0, g.h1, tx, ty);
    const float fval = fw[f];
    const float corrected = fval + ((axis == 0 ? du : dv)[f] - bwd) * 0.5f;
    const float2 r = (axis == 0 ? ru : rv)[f];
    (axis == 0 ? uo : vo)[f] = (corrected >= r.x && corrected <= r.y) ? corrected : fval;
}

// fillUMarker/fillVMarker + markAndProjectU/V
// (grid_blocked_boundary_condition_solver2.js:249-303): a face whose
// open fraction 1 - clamp(fractionInsideSdf(phi(-h/2), phi(+h/2))) is 0
// takes the (static) collider velocity 0 and is marked for extrapolation.
__global__ void k_mark_project2(G2 g, const float* __restrict__ sdf, float* __restrict__ u, float* __restrict__ v,
                                int* __restrict__ mu, int* __restrict__ mv) {
    const int nu = (g.n0 + 1) * g.n1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nu + g.n0 * (g.n1 + 1)) return;
    const int axis = t < nu ? 0 : 1;
    const int f = axis == 0 ? t : t - nu;
    const int s0 = g.n0 + (axis == 0);
    const int i = f % s0, j = f / s0;
    float px, py;
    face_pos2(g, axis, i, j, px, py);
    float s0v, s1v;
    if (axis == 0) {
        s0v = sdf_at(g, sdf, px - g.h0 * 0.5f, py - 0.0f);
        s1v = sdf_at(g, sdf, px + g.h0 * 0.5f, py + 0.0f);
    } else {
        s0v = sdf_at(g, sdf, px - 0.0f, py - g.h1 * 0.5f);
        s1v = sdf_at(g, sdf, px + 0.0f, py + g.h1 * 0.5f);
    }
    const float fr = fminf(fmaxf(fraction_inside_sdf(s0v, s1v), 0.0f), 1.0f);
    const float open = 1.0f - fr;
    int* m = axis == 0 ? mu : mv;
    if (open > 0.0f) {
        m[f] = 1;
    } else {
        (axis == 0 ? u : v)[f] = 0.0f;
        m[f] = 0;
    }
}

// noFluxProjectionU/V (305-370) with projectAndApplyFriction (75-96), for a
// static collider; reads the extrapolated velocity, writes the temporaries.
__global__ void k_no_flux2(G2 g, const float* __restrict__ sdf, float friction, const float* __restrict__ u,
                           const float* __restrict__ v, float* __restrict__ tu, float* __restrict__ tv) {
    const int nu = (g.n0 + 1) * g.n1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nu + g.n0 * (g.n1 + 1)) return;
    const int axis = t < nu ? 0 : 1;
    const int f = axis == 0 ? t : t - nu;
    const int s0 = g.n0 + (axis == 0);
    const int i = f % s0, j = f / s0;
    float px, py;
    face_pos2(g, axis, i, j, px, py);
    float* out = axis == 0 ? tu : tv;
    if (!(sdf_at(g, sdf, px, py) < 0.0f)) {
        out[f] = (axis == 0 ? u : v)[f];
        return;
    }
    float vel[2], gr[2];
    mac2(g, u, v, px, py, vel);
    sdf_grad2(g, sdf, px, py, gr);
    const float gl = sqrtf(gr[0] * gr[0] + gr[1] * gr[1]);
    if (!(gl > 0.0f)) {
        out[f] = 0.0f;
        return;
    }
    const float nn[2] = {gr[0] / gl, gr[1] / gl};
    const float vr[2] = {vel[0] - 0.0f, vel[1] - 0.0f};
    const float dn = vr[0] * nn[0] + vr[1] * nn[1];
    float vt[2] = {vr[0] - nn[0] * dn, vr[1] - nn[1] * dn};
    const float tl = sqrtf(vt[0] * vt[0] + vt[1] * vt[1]);
    if (tl > 0.0f) {
        const float veln = fmaxf(-dn, 0.0f);
        const float scale = fmaxf(1.0f - (veln * friction) / tl, 0.0f);
        vt[0] = vt[0] * scale;
        vt[1] = vt[1] * scale;
    }
    out[f] = vt[axis] + 0.0f;
}

__global__ void k_copy_faces2(int nu, int nv, const float* __restrict__ tu, const float* __restrict__ tv,
                              float* __restrict__ u, float* __restrict__ v) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nu) u[t] = tu[t];
    else if (t < nu + nv) v[t - nu] = tv[t - nu];
}

// blockedBoundary (372-402): a collider cell (sdf < 0) sets each face it
// shares with a fluid neighbour to the collider velocity 0.
__global__ void k_blocked2(G2 g, const float* __restrict__ sdf, float* __restrict__ u, float* __restrict__ v) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.n0 * g.n1) return;
    const int i = c % g.n0, j = c / g.n0;
    if (!(sdf[c] < 0.0f)) return;
    const int su = g.n0 + 1;
    if (i > 0 && !(sdf[c - 1] < 0.0f)) u[i + su * j] = 0.0f;
    if (i < g.n0 - 1 && !(sdf[c + 1] < 0.0f)) u[i + 1 + su * j] = 0.0f;
    if (j > 0 && !(sdf[c - g.n0] < 0.0f)) v[i + g.n0 * j] = 0.0f;
    if (j < g.n1 - 1 && !(sdf[c + g.n0] < 0.0f)) v[i + g.n0 * (j + 1)] = 0.0f;
}

struct Inflows2 {
    const float* sdf[4];
    float vel[4][2];
    int mode[4];  // 0 'set', 1 'add'
    int count;
};

// The closed walls (zeroULeft .. zeroVUp, 154-157), the inflows (167-200)
// and the clamp (124-140), each touching only its own face, in that order.
__global__ void k_walls_inflow_clamp2(G2 g, int flags, Inflows2 in, float* __restrict__ u, float* __restrict__ v) {
    const int nu = (g.n0 + 1) * g.n1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nu + g.n0 * (g.n1 + 1)) return;
    const int axis = t < nu ? 0 : 1;
    const int f = axis == 0 ? t : t - nu;
    const int s0 = g.n0 + (axis == 0);
    const int i = f % s0, j = f / s0;
    float* d = axis == 0 ? u : v;
    float val = d[f];
    if (axis == 0) {
        if ((flags & 1) && i == 0) val = 0.0f;
        if ((flags & 2) && i == g.n0) val = 0.0f;
    } else {
        if ((flags & 4) && j == 0) val = 0.0f;
        if ((flags & 8) && j == g.n1) val = 0.0f;
    }
    if (in.count > 0) {
        float px, py;
        face_pos2(g, axis, i, j, px, py);
        for (int q = 0; q < in.count; ++q) {
            if (!(sdf_at(g, in.sdf[q], px, py) < 0.0f)) continue;
            val = in.mode[q] == 1 ? val + in.vel[q][axis] : in.vel[q][axis];
        }
    }
    d[f] = !isfinite(val) ? 0.0f : fminf(fmaxf(val, -100.0f), 100.0f);
}

// extrapolateU/V of the outflow (grid_outflow_solver2.js:247-334) for one
// axis: faces inside the outflow relax toward the velocity one cell
// upstream by the outward Courant number, into dst.
__global__ void k_outflow2(G2 g, int axis, const float* __restrict__ osdf, float dt, const float* __restrict__ u,
                           const float* __restrict__ v, float* __restrict__ dst) {
    const int s0 = g.n0 + (axis == 0), s1 = g.n1 + (axis == 1);
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= s0 * s1) return;
    const int i = f % s0, j = f / s0;
    float px, py;
    face_pos2(g, axis, i, j, px, py);
    if (!(sdf_at(g, osdf, px, py) < 0.0f)) return;
    const float* cur_field = axis == 0 ? u : v;
    float gr[2];
    sdf_grad2(g, osdf, px, py, gr);
    const float gl = sqrtf(gr[0] * gr[0] + gr[1] * gr[1]);
    if (!(gl > 0.0f)) {
        dst[f] = cur_field[f];
        return;
    }
    const float nn[2] = {gr[0] / gl, gr[1] / gl};
    const float ha = axis == 0 ? g.h0 : g.h1;
    float bulk[2], up[2];
    mac2(g, u, v, px, py, bulk);
    const float dn = bulk[0] * -nn[0] + bulk[1] * -nn[1];
    const float outward = fmaxf(0.0f, dn);
    const float courant = fminf(fmaxf((outward * dt) / ha, 0.0f), 1.0f);
    mac2(g, u, v, px + nn[0] * ha, py + nn[1] * ha, up);
    const float cur = cur_field[f];
    const float relaxed = cur + (up[axis] - cur) * courant;
    dst[f] = fminf(fmaxf(relaxed, -1e3f), 1e3f);
}

// copyU/copyV of the outflow (286-297, 335-346)
__global__ void k_outflow_copy2(G2 g, int axis, const float* __restrict__ osdf, const float* __restrict__ dst,
                                float* __restrict__ d) {
    const int s0 = g.n0 + (axis == 0), s1 = g.n1 + (axis == 1);
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= s0 * s1) return;
    const int i = f % s0, j = f / s0;
    float px, py;
    face_pos2(g, axis, i, j, px, py);
    if (sdf_at(g, osdf, px, py) < 0.0f) d[f] = dst[f];
}

// createOutflowPressureDirichlet2 + combineDirichlet
// (sdf_inflow_outflow2.js:152-199) evaluated at the pressure grid's cell
// centres: active = any outflow's SDF < 0 (target 0) OR the caller's mask;
// the caller's target wins where the caller's mask is set.
__global__ void k_outflow_dirichlet2(G2 g, const float* __restrict__ o0, const float* __restrict__ o1,
                                     const float* __restrict__ o2, const float* __restrict__ o3, int count,
                                     const uint8_t* __restrict__ user_active, const float* __restrict__ user_target,
                                     uint8_t* __restrict__ active, float* __restrict__ target) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.n0 * g.n1) return;
    const int i = c % g.n0, j = c / g.n0;
    const float px = (g.o0 + g.h0 * 0.5f) + g.h0 * (float)i, py = (g.o1 + g.h1 * 0.5f) + g.h1 * (float)j;
    const float* os[4] = {o0, o1, o2, o3};
    bool act = false;
    for (int q = 0; q < count; ++q) act = act || sdf_at(g, os[q], px, py) < 0.0f;
    float tg = 0.0f;
    if (user_active) {
        const bool ua = user_active[c] != 0;
        act = act || ua;
        tg = ua ? (user_target ? user_target[c] : 0.0f) : 0.0f;
    }
    active[c] = (uint8_t)act;
    target[c] = tg;
}

// createMaxAbsReducer (reduction.js:43-77): the largest |x| by bit pattern;
// one atomic per block.
__global__ void k_max_abs_bits(const float* __restrict__ a, int n, unsigned* __restrict__ out) {
    __shared__ unsigned sm[GTPB / 64];
    unsigned m = 0;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const unsigned bits = __float_as_uint(fabsf(a[t]));
        m = bits > m ? bits : m;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned w = __shfl_down(m, o, 64);
        m = w > m ? w : m;
    }
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < GTPB / 64; ++q) m = sm[q] > m ? sm[q] : m;
        atomicMax(out, m);
    }
}

// ------------------------------------------------------ 2-D pressure system
// dispatchBuildSystem (grid_pressure_solver2.js:677-785)
__global__ void k_p2_build(G2 g, float ih2x, float ih2y, const float* __restrict__ u, const float* __restrict__ v,
                           const uint8_t* __restrict__ mask, const float* __restrict__ target,
                           const float* __restrict__ wu, const float* __restrict__ wv, float* __restrict__ b,
                           float* __restrict__ p) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.n0 * g.n1) return;
    const int i = c % g.n0, j = c / g.n0;
    const int su = g.n0 + 1;
    const float div = (u[i + 1 + su * j] - u[i + su * j]) / g.h0 + (v[i + g.n0 * (j + 1)] - v[i + g.n0 * j]) / g.h1;
    if (!mask) {
        b[c] = div;
        return;
    }
    if (mask[c]) {
        const float tg = target[c];
        b[c] = -tg;
        p[c] = tg;
        return;
    }
    float t0 = 0.0f, t1 = 0.0f, t2 = 0.0f, t3 = 0.0f;
    if (i > 0 && mask[c - 1]) t0 = ((wu ? wu[i + su * j] : 1.0f) * target[c - 1]) * ih2x;
    if (i < g.n0 - 1 && mask[c + 1]) t1 = ((wu ? wu[i + 1 + su * j] : 1.0f) * target[c + 1]) * ih2x;
    if (j > 0 && mask[c - g.n0]) t2 = ((wv ? wv[i + g.n0 * j] : 1.0f) * target[c - g.n0]) * ih2y;
    if (j < g.n1 - 1 && mask[c + g.n0]) t3 = ((wv ? wv[i + g.n0 * (j + 1)] : 1.0f) * target[c + g.n0]) * ih2y;
    b[c] = div - (((t0 + t1) + t2) + t3);
}

// dispatchCorrectU/V (796-834): interior faces only
__global__ void k_p2_correct(G2 g, const float* __restrict__ p, const float* __restrict__ wu,
                             const float* __restrict__ wv, const float* __restrict__ ui, const float* __restrict__ vi,
                             float* __restrict__ uo, float* __restrict__ vo) {
    const int nu = (g.n0 + 1) * g.n1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nu + g.n0 * (g.n1 + 1)) return;
    if (t < nu) {
        const int k = t % (g.n0 + 1), j = t / (g.n0 + 1);
        if (!(k > 0 && k < g.n0)) return;
        const float grad = (p[k + g.n0 * j] - p[k - 1 + g.n0 * j]) / g.h0;
        uo[t] = ui[t] - (wu ? grad * wu[t] : grad);
    } else {
        const int f = t - nu;
        const int i = f % g.n0, k = f / g.n0;
        if (!(k > 0 && k < g.n1)) return;
        const float grad = (p[i + g.n0 * k] - p[i + g.n0 * (k - 1)]) / g.h1;
        vo[f] = vi[f] - (wv ? grad * wv[f] : grad);
    }
}

// countBadPressureCells (605-611)
__global__ void k_p2_count_bad(const float* __restrict__ p, int n, float maxp, unsigned long long* count) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = false;
    if (c < n) bad = !(fabsf(p[c]) <= maxp);
    const unsigned long long m = __ballot(bad);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(count, (unsigned long long)__popcll(m));
}

__global__ void k_copy_f(const float* __restrict__ a, float* __restrict__ b, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) b[t] = a[t];
}

inline int64_t faces2(const int* n, int axis) { return (int64_t)(n[0] + (axis == 0)) * (n[1] + (axis == 1)); }
inline int64_t all_faces2(const int* n) { return faces2(n, 0) + faces2(n, 1); }

int check_grid2(const flux_grid2* g, const char* who) {
    FLUX_REQUIRE(g, "%s: NULL grid", who);
    FLUX_REQUIRE(g->n[0] > 0 && g->n[1] > 0 && (int64_t)g->n[0] * g->n[1] < (1ll << 28),
                 "%s: bad resolution [%d, %d]", who, g->n[0], g->n[1]);
    FLUX_REQUIRE(g->h[0] > 0 && g->h[1] > 0, "%s: grid spacing must be positive", who);
    return FLUX_OK;
}

int check_faces2(const flux_grid2* g, const flux_field* u, const flux_field* v, const char* who) {
    FLUX_TRY(check_field(u, FLUX_F32, faces2(g->n, 0), who));
    FLUX_TRY(check_field(v, FLUX_F32, faces2(g->n, 1), who));
    return FLUX_OK;
}

}  // namespace flux

using namespace flux;

// --------------------------------------------------------------- objects
struct flux_boundary2 {
    flux_grid2 g;
    flux_field* vel[2] = {nullptr, nullptr};
    const flux_field* sdf = nullptr;
    float friction = 0.5f;
    int flags = 15;
    Inflows2 inflows{};
    float* tmp[2] = {nullptr, nullptr};
    flux_field* marker[2] = {nullptr, nullptr};
    flux_extrapolator* ex[2] = {nullptr, nullptr};
};

struct flux_outflow2 {
    flux_grid2 g;
    const flux_field* sdf = nullptr;
    float* dst[2] = {nullptr, nullptr};
};

struct flux_pressure2 {
    flux_pressure2_desc d;
    int n[3] = {1, 1, 1};
    double h[2] = {1.0, 1.0};
    flux_field* pressure = nullptr;
    flux_field* b = nullptr;
    flux_field* snapshot = nullptr;
    const flux_field* wop[2] = {nullptr, nullptr};
    flux_operator* op = nullptr;
    flux_multigrid* mg = nullptr;
    flux_pcg* pcg = nullptr;
    unsigned long long* bad = nullptr;
    int maxiter = 0;
};

struct flux_advection2 {
    flux_grid2 g{};
    const flux_field* u = nullptr;  // the velocity grid bound at creation
    const flux_field* v = nullptr;
    const flux_field* sdf = nullptr;
    float dt = 0.0f;
    int max_sub = 32;
    int order = 1;
    float* fwd = nullptr;     // order 2: forward result of the U then V faces
    float2* range = nullptr;  // order 2: clamp ranges
};

struct flux_solver2 {
    flux_solver2_desc d;
    flux_field* vel[2] = {nullptr, nullptr};
    flux_field* prev[2] = {nullptr, nullptr};
    flux_field* active = nullptr;
    flux_field* target = nullptr;
    flux_boundary2* bnd = nullptr;
    flux_outflow2* out[4] = {nullptr, nullptr, nullptr, nullptr};
    flux_pressure2* pres = nullptr;
    flux_advection2* adv = nullptr;  // through velocityPrev (grid_solver2.js:198-199)
};

extern "C" {

int flux_polygon_sdf2(flux_field* sdf, const flux_grid2* g, const double* xy, const int* ring_start, int rings,
                      int invert) {
    FLUX_REQUIRE_INIT();
    FLUX_TRY(check_grid2(g, "addPolygons"));
    FLUX_REQUIRE(sdf && xy && ring_start && rings > 0, "addPolygons: NULL argument or no polygon");
    const int64_t cells = (int64_t)g->n[0] * g->n[1];
    FLUX_TRY(check_field(sdf, FLUX_F32, cells, "addPolygons: sdf"));
    FLUX_REQUIRE(ring_start[0] == 0, "addPolygons: ring_start[0] must be 0");
    for (int r = 0; r < rings; ++r)
        FLUX_REQUIRE(ring_start[r + 1] - ring_start[r] >= 1, "addPolygons: empty polygon %d", r);
    const int points = ring_start[rings];
    double* dxy = (double*)dmalloc(sizeof(double) * 2 * points);
    int* dst = (int*)dmalloc(sizeof(int) * (rings + 1));
    if (!dxy || !dst) {
        dfree(dxy);
        dfree(dst);
        return set_error(FLUX_ERR_NOMEM, "addPolygons: allocation failed");
    }
    int rc = FLUX_OK;
    if (hipMemcpyAsync(dxy, xy, sizeof(double) * 2 * points, hipMemcpyHostToDevice, g_stream) != hipSuccess ||
        hipMemcpyAsync(dst, ring_start, sizeof(int) * (rings + 1), hipMemcpyHostToDevice, g_stream) != hipSuccess)
        rc = set_error(FLUX_ERR_HIP, "addPolygons: copy failed");
    if (!rc) {
        k_polygon_sdf2<<<blocks_for(cells, GTPB), GTPB, 0, g_stream>>>(
            g->n[0], g->n[1], (double)g->h[0], (double)g->h[1], (double)g->origin[0], (double)g->origin[1], dxy, dst,
            rings, invert ? 1 : 0, (float*)sdf->d);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(g_stream) != hipSuccess)
            rc = set_error(FLUX_ERR_HIP, "addPolygons: kernel failed");
    }
    dfree(dxy);
    dfree(dst);
    return rc;
}

int flux_advect_faces2(const flux_grid2* g, const flux_field* u, const flux_field* v, const flux_field* collider_sdf,
                       float dt, int max_substeps, flux_field* u_out, flux_field* v_out) {
    FLUX_REQUIRE_INIT();
    FLUX_TRY(check_grid2(g, "advectFaceCentered2"));
    FLUX_TRY(check_faces2(g, u, v, "advectFaceCentered2: input"));
    FLUX_TRY(check_faces2(g, u_out, v_out, "advectFaceCentered2: output"));
    FLUX_REQUIRE(u_out != u && v_out != v, "advectFaceCentered2: output must not alias the input");
    FLUX_REQUIRE(max_substeps > 0, "advectFaceCentered2: maxSubsteps must be positive");
    if (collider_sdf) FLUX_TRY(check_field(collider_sdf, FLUX_F32, (int64_t)g->n[0] * g->n[1], "collider sdf"));
    const int64_t nf = all_faces2(g->n);
    k_advect_face2<<<blocks_for(nf, GTPB), GTPB, 0, g_stream>>>(
        make_g2(g), (const float*)u->d, (const float*)v->d, (const float*)u->d, (const float*)v->d,
        collider_sdf ? (const float*)collider_sdf->d : nullptr, dt, max_substeps, (float*)u_out->d,
        (float*)v_out->d);
    FLUX_LAUNCH_CHECK();
    return FLUX_OK;
}

int flux_advection2_create(const flux_grid2* g, const flux_field* u, const flux_field* v,
                           const flux_field* collider_sdf, float dt, int max_substeps, int order,
                           flux_advection2** out) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(out, "createSemiLagrangianAdvectionSolver2: NULL argument");
    FLUX_TRY(check_grid2(g, "createSemiLagrangianAdvectionSolver2"));
    FLUX_TRY(check_faces2(g, u, v, "createSemiLagrangianAdvectionSolver2: velocity"));
    FLUX_REQUIRE(order == 1 || order == 2, "createSemiLagrangianAdvectionSolver2: order must be 1 or 2");
    FLUX_REQUIRE(max_substeps >= 1 && max_substeps <= 1024,
                 "createSemiLagrangianAdvectionSolver2: maxSubsteps out of range");
    if (collider_sdf) FLUX_TRY(check_field(collider_sdf, FLUX_F32, (int64_t)g->n[0] * g->n[1], "collider sdf"));
    const int64_t nf = all_faces2(g->n);
    flux_advection2* a = new flux_advection2();
    a->g = *g;
    a->u = u;
    a->v = v;
    a->sdf = collider_sdf;
    a->dt = dt;
    a->max_sub = max_substeps;
    a->order = order;
    if (order == 2) {
        a->fwd = (float*)dmalloc(sizeof(float) * nf);
        a->range = (float2*)dmalloc(sizeof(float2) * nf);
        if (!a->fwd || !a->range) {
            flux_advection2_destroy(a);
            return set_error(FLUX_ERR_HIP, "createSemiLagrangianAdvectionSolver2: out of device memory");
        }
    }
    *out = a;
    return FLUX_OK;
}

int flux_advection2_set_dt(flux_advection2* a, float dt) {
    FLUX_REQUIRE(a, "advection: NULL solver");
    a->dt = dt;
    return FLUX_OK;
}

// advectFaceCentered2(input, output) of the solver (advection_solver2.js:
// 263-430) on the stream, no synchronisation
static int advection2_faces_async(flux_advection2* a, const flux_field* in_u, const flux_field* in_v,
                                  flux_field* out_u, flux_field* out_v) {
    FLUX_REQUIRE(a, "advectFaceCentered2: NULL solver");
    FLUX_TRY(check_faces2(&a->g, in_u, in_v, "advectFaceCentered2: input"));
    FLUX_TRY(check_faces2(&a->g, out_u, out_v, "advectFaceCentered2: output"));
    for (const flux_field* o : {out_u, out_v})
        FLUX_REQUIRE(o->d != in_u->d && o->d != in_v->d && o->d != a->u->d && o->d != a->v->d,
                     "advectFaceCentered2: output must not alias the input or velocity");
    const int64_t nf = all_faces2(a->g.n);
    const int nu = (int)faces2(a->g.n, 0);
    const G2 g = make_g2(&a->g);
    const float* sdf = a->sdf ? (const float*)a->sdf->d : nullptr;
    const float *u = (const float*)a->u->d, *v = (const float*)a->v->d;
    const float *du = (const float*)in_u->d, *dv = (const float*)in_v->d;
    if (a->order == 1) {
        k_advect_face2<<<blocks_for(nf, GTPB), GTPB, 0, g_stream>>>(g, u, v, du, dv, sdf, a->dt, a->max_sub,
                                                                     (float*)out_u->d, (float*)out_v->d);
    } else {
        float *fu = a->fwd, *fv = a->fwd + nu;
        float2 *ru = a->range, *rv = a->range + nu;
        k_mc_forward2<<<blocks_for(nf, GTPB), GTPB, 0, g_stream>>>(g, u, v, du, dv, sdf, a->dt, a->max_sub, fu, fv,
                                                                    ru, rv);
        k_mc_correct2<<<blocks_for(nf, GTPB), GTPB, 0, g_stream>>>(g, u, v, du, dv, sdf, a->dt, a->max_sub, fu, fv,
                                                                    ru, rv, (float*)out_u->d, (float*)out_v->d);
    }
    FLUX_LAUNCH_CHECK();
    return FLUX_OK;
}

int flux_advection2_faces(flux_advection2* a, const flux_field* in_u, const flux_field* in_v, flux_field* out_u,
                          flux_field* out_v) {
    FLUX_REQUIRE_INIT();
    FLUX_TRY(advection2_faces_async(a, in_u, in_v, out_u, out_v));
    FLUX_HIP(hipStreamSynchronize(g_stream));
    return FLUX_OK;
}

int flux_advection2_destroy(flux_advection2* a) {
    if (!a) return FLUX_OK;
    if (g_init) (void)hipStreamSynchronize(g_stream);
    if (a->fwd) dfree(a->fwd);
    if (a->range) dfree(a->range);
    delete a;
    return FLUX_OK;
}

int flux_max_abs(const flux_field* const* fields, int count, float* out) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(fields && out && count > 0, "createMaxAbsReducer: bad arguments");
    unsigned* acc = (unsigned*)dmalloc(sizeof(unsigned));
    FLUX_REQUIRE(acc, "createMaxAbsReducer: allocation failed");
    int rc = FLUX_OK;
    if (hipMemsetAsync(acc, 0, sizeof(unsigned), g_stream) != hipSuccess) rc = set_error(FLUX_ERR_HIP, "memset");
    for (int q = 0; q < count && !rc; ++q) {
        const flux_field* f = fields[q];
        if (!f || f->dtype != FLUX_F32 || f->count >= (1ll << 31)) {
            rc = set_error(FLUX_ERR_ARG, "createMaxAbsReducer: fields must be float");
            break;
        }
        const unsigned nb = std::min<unsigned>(blocks_for(f->count, GTPB), 1024u);
        if (nb) k_max_abs_bits<<<nb, GTPB, 0, g_stream>>>((const float*)f->d, (int)f->count, acc);
    }
    unsigned bits = 0;
    if (!rc && (hipGetLastError() != hipSuccess ||
                hipMemcpyAsync(&bits, acc, sizeof bits, hipMemcpyDeviceToHost, g_stream) != hipSuccess ||
                hipStreamSynchronize(g_stream) != hipSuccess))
        rc = set_error(FLUX_ERR_HIP, "createMaxAbsReducer: reduction failed");
    dfree(acc);
    if (!rc) std::memcpy(out, &bits, sizeof bits);
    return rc;
}

int flux_adaptive_substeps(double max_velocity, double min_spacing, double frame_dt, double courant, int max_sub,
                           int* steps, double* sub_dt) {
    FLUX_REQUIRE(min_spacing > 0 && courant > 0 && max_sub >= 1 && steps && sub_dt,
                 "computeAdaptiveSubSteps: bad arguments");
    const double cfl = max_velocity * frame_dt / min_spacing;
    double k = std::ceil(cfl / courant);
    if (!(k >= 1)) k = 1;  // also a non-finite cfl
    if (k > max_sub) k = max_sub;
    *steps = (int)k;
    *sub_dt = frame_dt / *steps;
    return FLUX_OK;
}

// ------------------------------------------------------------- boundary2
int flux_boundary2_destroy(flux_boundary2* b) {
    if (!b) return FLUX_OK;
    if (g_init) (void)hipStreamSynchronize(g_stream);
    for (int a = 0; a < 2; ++a) {
        flux_extrapolator3_destroy(b->ex[a]);
        if (b->marker[a]) flux_field_destroy(b->marker[a]);
        dfree(b->tmp[a]);
    }
    delete b;
    return FLUX_OK;
}

int flux_boundary2_create(const flux_grid2* g, flux_field* const* vel, const flux_field* sdf, float friction,
                          flux_boundary2** out) {
    FLUX_REQUIRE_INIT();
    FLUX_TRY(check_grid2(g, "createGridBlockedBoundaryConditionSolver2"));
    FLUX_REQUIRE(vel && out, "createGridBlockedBoundaryConditionSolver2: NULL argument");
    FLUX_TRY(check_faces2(g, vel[0], vel[1], "createGridBlockedBoundaryConditionSolver2: velocity"));
    if (sdf) FLUX_TRY(check_field(sdf, FLUX_F32, (int64_t)g->n[0] * g->n[1], "collider sdf"));
    flux_boundary2* b = new flux_boundary2();
    b->g = *g;
    b->vel[0] = vel[0];
    b->vel[1] = vel[1];
    b->sdf = sdf;
    b->friction = friction;
    if (sdf) {
        for (int a = 0; a < 2; ++a) {
            const int64_t shape[2] = {g->n[0] + (a == 0), g->n[1] + (a == 1)};
            int rc = flux_field_create(FLUX_I32, 2, shape, &b->marker[a]);
            if (!rc) rc = flux_extrapolator3_create(vel[a], b->marker[a], vel[a], &b->ex[a]);
            b->tmp[a] = (float*)dmalloc(sizeof(float) * faces2(g->n, a));
            if (!rc && !b->tmp[a]) rc = set_error(FLUX_ERR_NOMEM, "boundary: allocation failed");
            if (rc) {
                flux_boundary2_destroy(b);
                return rc;
            }
        }
    }
    *out = b;
    return FLUX_OK;
}

int flux_boundary2_set_closed(flux_boundary2* b, int flags) {
    FLUX_REQUIRE(b, "closedDomainBoundaryFlag: NULL solver");
    b->flags = flags;
    return FLUX_OK;
}

int flux_boundary2_add_inflow(flux_boundary2* b, const flux_field* sdf, float vx, float vy, int mode) {
    FLUX_REQUIRE(b && sdf, "setInflows: NULL argument");
    FLUX_REQUIRE(b->inflows.count < 4, "setInflows: at most 4 inflows");
    FLUX_REQUIRE(mode == 0 || mode == 1, "setInflows: mode must be 'set' (0) or 'add' (1)");
    FLUX_TRY(check_field(sdf, FLUX_F32, (int64_t)b->g.n[0] * b->g.n[1], "inflow sdf"));
    const int q = b->inflows.count++;
    b->inflows.sdf[q] = (const float*)sdf->d;
    b->inflows.vel[q][0] = vx;
    b->inflows.vel[q][1] = vy;
    b->inflows.mode[q] = mode;
    return FLUX_OK;
}

int flux_boundary2_constrain(flux_boundary2* b, int depth) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(b && depth >= 0, "constrainVelocity: bad arguments");
    const G2 g = make_g2(&b->g);
    const int64_t nf = all_faces2(b->g.n);
    float* u = (float*)b->vel[0]->d;
    float* v = (float*)b->vel[1]->d;
    if (b->sdf) {
        const float* sdf = (const float*)b->sdf->d;
        k_mark_project2<<<blocks_for(nf, GTPB), GTPB, 0, g_stream>>>(g, sdf, u, v, (int*)b->marker[0]->d,
                                                                     (int*)b->marker[1]->d);
        FLUX_LAUNCH_CHECK();
        for (int a = 0; a < 2; ++a) FLUX_TRY(flux_extrapolator3_run(b->ex[a], depth));
        k_no_flux2<<<blocks_for(nf, GTPB), GTPB, 0, g_stream>>>(g, sdf, b->friction, u, v, b->tmp[0], b->tmp[1]);
        k_copy_faces2<<<blocks_for(nf, GTPB), GTPB, 0, g_stream>>>((int)faces2(b->g.n, 0), (int)faces2(b->g.n, 1),
                                                                   b->tmp[0], b->tmp[1], u, v);
        k_blocked2<<<blocks_for((int64_t)g.n0 * g.n1, GTPB), GTPB, 0, g_stream>>>(g, sdf, u, v);
    }
    k_walls_inflow_clamp2<<<blocks_for(nf, GTPB), GTPB, 0, g_stream>>>(g, b->flags, b->inflows, u, v);
    FLUX_LAUNCH_CHECK();
    return FLUX_OK;
}

// -------------------------------------------------------------- outflow2
int flux_outflow2_destroy(flux_outflow2* o) {
    if (!o) return FLUX_OK;
    if (g_init) (void)hipStreamSynchronize(g_stream);
    dfree(o->dst[0]);
    dfree(o->dst[1]);
    delete o;
    return FLUX_OK;
}

int flux_outflow2_create(const flux_grid2* g, const flux_field* sdf, flux_outflow2** out) {
    FLUX_REQUIRE_INIT();
    FLUX_TRY(check_grid2(g, "createGridOutflowSolver2"));
    FLUX_REQUIRE(sdf && out, "createGridOutflowSolver2: NULL argument");
    FLUX_TRY(check_field(sdf, FLUX_F32, (int64_t)g->n[0] * g->n[1], "outflow sdf"));
    flux_outflow2* o = new flux_outflow2();
    o->g = *g;
    o->sdf = sdf;
    for (int a = 0; a < 2; ++a) o->dst[a] = (float*)dmalloc(sizeof(float) * faces2(g->n, a));
    if (!o->dst[0] || !o->dst[1]) {
        flux_outflow2_destroy(o);
        return set_error(FLUX_ERR_NOMEM, "createGridOutflowSolver2: allocation failed");
    }
    *out = o;
    return FLUX_OK;
}

int flux_outflow2_apply(flux_outflow2* o, flux_field* u, flux_field* v, float dt) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(o, "applyOutflowVelocityBC: NULL solver");
    FLUX_TRY(check_faces2(&o->g, u, v, "applyOutflowVelocityBC: velocity"));
    const G2 g = make_g2(&o->g);
    float* d[2] = {(float*)u->d, (float*)v->d};
    for (int a = 0; a < 2; ++a) {
        const int64_t nf = faces2(o->g.n, a);
        k_outflow2<<<blocks_for(nf, GTPB), GTPB, 0, g_stream>>>(g, a, (const float*)o->sdf->d, dt, d[0], d[1],
                                                                o->dst[a]);
        k_outflow_copy2<<<blocks_for(nf, GTPB), GTPB, 0, g_stream>>>(g, a, (const float*)o->sdf->d, o->dst[a], d[a]);
    }
    FLUX_LAUNCH_CHECK();
    return FLUX_OK;
}

int flux_outflow2_dirichlet(const flux_grid2* g, const flux_field* const* outflow_sdf, int count,
                            const flux_field* user_active, const flux_field* user_target, flux_field* active,
                            flux_field* target) {
    FLUX_REQUIRE_INIT();
    FLUX_TRY(check_grid2(g, "createOutflowPressureDirichlet2"));
    FLUX_REQUIRE(count >= 0 && count <= 4 && active && target, "createOutflowPressureDirichlet2: bad arguments");
    const int64_t cells = (int64_t)g->n[0] * g->n[1];
    FLUX_TRY(check_field(active, FLUX_U8, cells, "dirichlet active"));
    FLUX_TRY(check_field(target, FLUX_F32, cells, "dirichlet target"));
    const float* os[4] = {nullptr, nullptr, nullptr, nullptr};
    for (int q = 0; q < count; ++q) {
        FLUX_TRY(check_field(outflow_sdf[q], FLUX_F32, cells, "outflow sdf"));
        os[q] = (const float*)outflow_sdf[q]->d;
    }
    if (user_active) FLUX_TRY(check_field(user_active, FLUX_U8, cells, "dirichlet active"));
    if (user_target) FLUX_TRY(check_field(user_target, FLUX_F32, cells, "dirichlet target"));
    k_outflow_dirichlet2<<<blocks_for(cells, GTPB), GTPB, 0, g_stream>>>(
        make_g2(g), os[0], os[1], os[2], os[3], count, user_active ? (const uint8_t*)user_active->d : nullptr,
        user_target ? (const float*)user_target->d : nullptr, (uint8_t*)active->d, (float*)target->d);
    FLUX_LAUNCH_CHECK();
    return FLUX_OK;
}

// ------------------------------------------------------------- pressure2
void flux_pressure2_desc_default(flux_pressure2_desc* d) {
    std::memset(d, 0, sizeof *d);
    d->spacing[0] = d->spacing[1] = 1.0;
    flux_mg_options_default(&d->multigrid);
    d->preconditioner = FLUX_PREC_MULTIGRID;
    d->tolerance = 1e-6;
    d->max_iterations = 0;
    d->max_iterations_factor = 1.5;
    d->relative_tolerance = 1;
    d->max_plausible_pressure = 1e6;
}

int flux_pressure2_destroy(flux_pressure2* p) {
    if (!p) return FLUX_OK;
    flux_pcg_destroy(p->pcg);
    flux_multigrid_destroy(p->mg);
    flux_operator_destroy(p->op);
    flux_field_destroy(p->pressure);
    flux_field_destroy(p->b);
    flux_field_destroy(p->snapshot);
    dfree(p->bad);
    delete p;
    return FLUX_OK;
}

int flux_pressure2_create(const flux_pressure2_desc* d, flux_pressure2** out) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(d && out, "createGridPressureSolver2: NULL argument");
    FLUX_TRY(check_shape(2, d->resolution, d->spacing, 2, "createGridPressureSolver2"));
    FLUX_REQUIRE(d->preconditioner >= 0 && d->preconditioner <= 2,
                 "grid_pressure_solver2: unknown preconditioner '%d' (expected multigrid, jacobi, none).",
                 d->preconditioner);
    flux_pressure2* p = new flux_pressure2();
    p->d = *d;
    p->n[0] = d->resolution[0];
    p->n[1] = d->resolution[1];
    p->h[0] = d->spacing[0];
    p->h[1] = d->spacing[1];
    const int64_t cells = (int64_t)p->n[0] * p->n[1];
    int rc = FLUX_OK;
#define P2_CHECK(expr)                 \
    do {                               \
        rc = (expr);                   \
        if (rc) {                      \
            flux_pressure2_destroy(p); \
            return rc;                 \
        }                              \
    } while (0)
    P2_CHECK((d->dirichlet_active == nullptr) == (d->dirichlet_target == nullptr)
                 ? FLUX_OK
                 : set_error(FLUX_ERR_ARG, "createGridPressureSolver2: dirichlet needs both active and target"));
    if (d->dirichlet_active) P2_CHECK(check_field(d->dirichlet_active, FLUX_U8, cells, "dirichlet active"));
    if (d->dirichlet_target) P2_CHECK(check_field(d->dirichlet_target, FLUX_F32, cells, "dirichlet target"));
    const bool has_w = d->face_weights[0] != nullptr;
    for (int a = 0; a < 2 && has_w; ++a) {
        P2_CHECK(check_field(d->face_weights[a], FLUX_F32, faces2(p->n, a), "faceWeights"));
        p->wop[a] = d->face_weights[a];
    }
    const int64_t shape[2] = {p->n[0], p->n[1]};
    P2_CHECK(flux_field_create(FLUX_F32, 2, shape, &p->pressure));
    P2_CHECK(flux_field_create(FLUX_F32, 2, shape, &p->b));
    P2_CHECK(flux_field_create(FLUX_F32, 2, shape, &p->snapshot));
    const flux_field* const* wop = has_w ? p->wop : nullptr;
    P2_CHECK(flux_operator_create(2, p->n, p->h, d->dirichlet_active, wop, &p->op));
    if (d->preconditioner == FLUX_PREC_MULTIGRID)
        P2_CHECK(flux_multigrid_create(2, p->n, p->h, &d->multigrid, d->dirichlet_active, nullptr, wop, &p->mg));
    P2_CHECK(flux_pcg_create(p->op, d->preconditioner, p->mg, p->b, p->pressure, &p->pcg));
    p->bad = (unsigned long long*)dmalloc(sizeof(unsigned long long));
    if (!p->bad) P2_CHECK(set_error(FLUX_ERR_NOMEM, "createGridPressureSolver2: allocation failed"));
    // mantaflow's cgMaxIterFac for 2-D (grid_pressure_solver2.js:398)
    const int mr = std::max(p->n[0], p->n[1]);
    p->maxiter = d->max_iterations > 0 ? d->max_iterations : (int)std::ceil(d->max_iterations_factor * mr * 4);
#undef P2_CHECK
    *out = p;
    return FLUX_OK;
}

flux_field* flux_pressure2_field(flux_pressure2* p, int which) {
    if (!p) return nullptr;
    switch (which) {
        case FLUX_P2_PRESSURE: return p->pressure;
        case FLUX_P2_RHS: return p->b;
        default: return nullptr;
    }
}

int flux_pressure2_project(flux_pressure2* p, const flux_field* u_in, const flux_field* v_in, flux_field* u_out,
                           flux_field* v_out, flux_pressure_diag* diag) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(p && diag, "project: NULL argument");
    flux_grid2 gg{{p->n[0], p->n[1]}, {(float)p->h[0], (float)p->h[1]}, {0.0f, 0.0f}};
    if (check_faces2(&gg, u_in, v_in, "project input"))
        return set_error(FLUX_ERR_ARG,
                         "createGridPressureSolver2: project() input resolution does not match constructed resolution [%d, %d].",
                         p->n[0], p->n[1]);
    FLUX_TRY(check_faces2(&gg, u_out, v_out, "project output"));
    const flux_pressure2_desc& d = p->d;
    const G2 g = make_g2(&gg);
    const int64_t cells = (int64_t)p->n[0] * p->n[1];
    const float* wu = p->wop[0] ? (const float*)p->wop[0]->d : nullptr;
    const float* wv = p->wop[1] ? (const float*)p->wop[1]->d : nullptr;
    // buildSystemBatch (839-842) and refreshPreconditionerMask (843-845)
    k_p2_build<<<blocks_for(cells, GTPB), GTPB, 0, g_stream>>>(
        g, (float)(1.0 / (p->h[0] * p->h[0])), (float)(1.0 / (p->h[1] * p->h[1])), (const float*)u_in->d,
        (const float*)v_in->d, d.dirichlet_active ? (const uint8_t*)d.dirichlet_active->d : nullptr,
        d.dirichlet_target ? (const float*)d.dirichlet_target->d : nullptr, wu, wv, (float*)p->b->d,
        (float*)p->pressure->d);
    FLUX_LAUNCH_CHECK();
    if (p->mg) FLUX_TRY(mg_refresh(p->mg));
    flux_pcg_state st;
    FLUX_TRY(pcg_solve_raw(p->pcg, d.tolerance, p->maxiter, d.relative_tolerance, &st));
    // circuit breaker (605-661, 930-954)
    const unsigned gc = blocks_for(cells, GTPB);
    FLUX_HIP(hipMemsetAsync(p->bad, 0, sizeof(unsigned long long), g_stream));
    k_p2_count_bad<<<gc, GTPB, 0, g_stream>>>((const float*)p->pressure->d, (int)cells, (float)d.max_plausible_pressure,
                                               p->bad);
    unsigned long long bad = 0;
    FLUX_HIP(hipMemcpyAsync(&bad, p->bad, sizeof bad, hipMemcpyDeviceToHost, g_stream));
    FLUX_HIP(hipStreamSynchronize(g_stream));
    const bool rejected = bad > 0;
    if (rejected)
        k_copy_f<<<gc, GTPB, 0, g_stream>>>((const float*)p->snapshot->d, (float*)p->pressure->d, (int)cells);
    else
        k_copy_f<<<gc, GTPB, 0, g_stream>>>((const float*)p->pressure->d, (float*)p->snapshot->d, (int)cells);
    const int64_t nf = all_faces2(p->n);
    k_p2_correct<<<blocks_for(nf, GTPB), GTPB, 0, g_stream>>>(g, (const float*)p->pressure->d, wu, wv,
                                                               (const float*)u_in->d, (const float*)v_in->d,
                                                               (float*)u_out->d, (float*)v_out->d);
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

// --------------------------------------------------------------- solver2
void flux_solver2_desc_default(flux_solver2_desc* d) {
    std::memset(d, 0, sizeof *d);
    d->grid.h[0] = d->grid.h[1] = 1.0f;
    d->friction = 0.5f;
    d->closed_flags = 15;
    d->outflow_velocity_bc = 1;
    d->max_substeps = 32;
    d->extrapolation_depth = 5;
    d->advection_order = 1;
    flux_pressure2_desc pd;
    flux_pressure2_desc_default(&pd);
    d->multigrid = pd.multigrid;
    d->preconditioner = pd.preconditioner;
    d->tolerance = pd.tolerance;
    d->max_iterations = 0;
    d->max_iterations_factor = pd.max_iterations_factor;
    d->relative_tolerance = 1;
    d->max_plausible_pressure = pd.max_plausible_pressure;
}

int flux_solver2_destroy(flux_solver2* s) {
    if (!s) return FLUX_OK;
    flux_pressure2_destroy(s->pres);
    flux_advection2_destroy(s->adv);
    for (int q = 0; q < 4; ++q) flux_outflow2_destroy(s->out[q]);
    flux_boundary2_destroy(s->bnd);
    for (int a = 0; a < 2; ++a) {
        flux_field_destroy(s->vel[a]);
        flux_field_destroy(s->prev[a]);
    }
    flux_field_destroy(s->active);
    flux_field_destroy(s->target);
    delete s;
    return FLUX_OK;
}

int flux_solver2_create(const flux_solver2_desc* d, flux_solver2** out) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(d && out, "createGridSolver2: NULL argument");
    FLUX_TRY(check_grid2(&d->grid, "createGridSolver2"));
    FLUX_REQUIRE(d->n_inflows >= 0 && d->n_inflows <= 4 && d->n_outflows >= 0 && d->n_outflows <= 4,
                 "createGridSolver2: at most 4 inflows and 4 outflows");
    FLUX_REQUIRE(d->max_substeps > 0, "createGridSolver2: advection maxSubsteps must be positive");
    FLUX_REQUIRE((d->dirichlet_active == nullptr) == (d->dirichlet_target == nullptr),
                 "createGridSolver2: dirichlet needs both active and target");
    flux_solver2* s = new flux_solver2();
    s->d = *d;
    const flux_grid2& g = d->grid;
    int rc = FLUX_OK;
#define S2_CHECK(expr)               \
    do {                             \
        rc = (expr);                 \
        if (rc) {                    \
            flux_solver2_destroy(s); \
            return rc;               \
        }                            \
    } while (0)
    for (int a = 0; a < 2; ++a) {
        const int64_t shape[2] = {g.n[0] + (a == 0), g.n[1] + (a == 1)};
        S2_CHECK(flux_field_create(FLUX_F32, 2, shape, &s->vel[a]));
        S2_CHECK(flux_field_create(FLUX_F32, 2, shape, &s->prev[a]));
        S2_CHECK(flux_field_upload(s->vel[a], std::vector<float>(shape[0] * shape[1], 0.0f).data(),
                                   sizeof(float) * shape[0] * shape[1]));
    }
    S2_CHECK(flux_advection2_create(&g, s->prev[0], s->prev[1], d->advect_collider ? d->collider_sdf : nullptr, 0.0f,
                                    d->max_substeps, d->advection_order == 0 ? 1 : d->advection_order, &s->adv));
    S2_CHECK(flux_boundary2_create(&g, s->vel, d->collider_sdf, d->friction, &s->bnd));
    S2_CHECK(flux_boundary2_set_closed(s->bnd, d->closed_flags));
    for (int q = 0; q < d->n_inflows; ++q)
        S2_CHECK(flux_boundary2_add_inflow(s->bnd, d->inflow_sdf[q], d->inflow_velocity[q][0],
                                           d->inflow_velocity[q][1], d->inflow_mode[q]));
    for (int q = 0; q < d->n_outflows; ++q) S2_CHECK(flux_outflow2_create(&g, d->outflow_sdf[q], &s->out[q]));
    flux_pressure2_desc pd;
    flux_pressure2_desc_default(&pd);
    pd.resolution[0] = g.n[0];
    pd.resolution[1] = g.n[1];
    pd.spacing[0] = g.h[0];
    pd.spacing[1] = g.h[1];
    pd.origin[0] = g.origin[0];
    pd.origin[1] = g.origin[1];
    if (d->n_outflows > 0 || d->dirichlet_active) {
        const int64_t shape[2] = {g.n[0], g.n[1]};
        S2_CHECK(flux_field_create(FLUX_U8, 2, shape, &s->active));
        S2_CHECK(flux_field_create(FLUX_F32, 2, shape, &s->target));
        S2_CHECK(flux_outflow2_dirichlet(&g, d->outflow_sdf, d->n_outflows, d->dirichlet_active, d->dirichlet_target,
                                         s->active, s->target));
        pd.dirichlet_active = s->active;
        pd.dirichlet_target = s->target;
    }
    pd.multigrid = d->multigrid;
    pd.preconditioner = d->preconditioner;
    pd.tolerance = d->tolerance;
    pd.max_iterations = d->max_iterations;
    pd.max_iterations_factor = d->max_iterations_factor;
    pd.relative_tolerance = d->relative_tolerance;
    pd.max_plausible_pressure = d->max_plausible_pressure;
    S2_CHECK(flux_pressure2_create(&pd, &s->pres));
    S2_CHECK(hipStreamSynchronize(g_stream) == hipSuccess ? FLUX_OK
                                                          : set_error(FLUX_ERR_HIP, "createGridSolver2: sync failed"));
#undef S2_CHECK
    *out = s;
    return FLUX_OK;
}

int flux_solver2_step(flux_solver2* s, float dt, flux_pressure_diag* diag) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(s && diag, "onAdvanceTimeStep: NULL argument");
    const flux_solver2_desc& d = s->d;
    // computePressure (grid_solver2.js:213-218): project in place, constrain
    FLUX_TRY(flux_pressure2_project(s->pres, s->vel[0], s->vel[1], s->vel[0], s->vel[1], diag));
    FLUX_TRY(flux_boundary2_constrain(s->bnd, d.extrapolation_depth));
    // computeAdvection (220-233): clone, advect the clone into velocity,
    // the outflow velocity condition, constrain
    const int64_t nf = all_faces2(d.grid.n);
    const int nu = (int)faces2(d.grid.n, 0), nv = (int)faces2(d.grid.n, 1);
    k_copy_faces2<<<blocks_for(nf, GTPB), GTPB, 0, g_stream>>>(nu, nv, (const float*)s->vel[0]->d,
                                                               (const float*)s->vel[1]->d, (float*)s->prev[0]->d,
                                                               (float*)s->prev[1]->d);
    FLUX_LAUNCH_CHECK();
    FLUX_TRY(flux_advection2_set_dt(s->adv, dt));
    FLUX_TRY(advection2_faces_async(s->adv, s->prev[0], s->prev[1], s->vel[0], s->vel[1]));
    if (d.outflow_velocity_bc)
        for (int q = 0; q < d.n_outflows; ++q) FLUX_TRY(flux_outflow2_apply(s->out[q], s->vel[0], s->vel[1], dt));
    FLUX_TRY(flux_boundary2_constrain(s->bnd, d.extrapolation_depth));
    FLUX_HIP(hipStreamSynchronize(g_stream));
    return FLUX_OK;
}

flux_field* flux_solver2_field(flux_solver2* s, int which) {
    if (!s) return nullptr;
    switch (which) {
        case FLUX_S2_U: return s->vel[0];
        case FLUX_S2_V: return s->vel[1];
        case FLUX_S2_U_PREV: return s->prev[0];
        case FLUX_S2_V_PREV: return s->prev[1];
        case FLUX_S2_PRESSURE: return s->pres ? s->pres->pressure : nullptr;
        case FLUX_S2_DIRICHLET: return s->active;
        default: return nullptr;
    }
}

}  // extern "C"
