/*
 * Oracle for the 2-D Eulerian step around the projection (TEST
 * INFRASTRUCTURE ONLY): the polygon SDF of a static collider / inflow /
 * outflow, semi-Lagrangian face advection that stops at the collider, the
 * blocked boundary-condition solver with inflows, the outflow velocity
 * condition, the 2-D pressure system and correction, and the max-|u|
 * reduction behind the CFL sub-step count.  f32 arithmetic in the
 * reference's operation order (built with -ffp-contract=off); the polygon
 * SDF in f64 like the reference's host code.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_util.h"

static int clampi2(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* ---- polygon_sdf.js:24-97: even-odd point in ring, distance to the ring's
 *      segments (closed: segment j -> i with j the previous vertex), signed
 *      (negative inside); the minimum over rings. ---- */
static int point_in_ring(double x, double y, const double* xy, int count) {
    int inside = 0;
    for (int i = 0, j = count - 1; i < count; j = i++) {
        const double xi = xy[2 * i], yi = xy[2 * i + 1];
        const double xj = xy[2 * j], yj = xy[2 * j + 1];
        const int intersect = ((yi > y) != (yj > y)) && (x < (xj - xi) * (y - yi) / (yj - yi) + xi);
        if (intersect) inside = !inside;
    }
    return inside;
}

static double segment_distance(double x, double y, double x1, double y1, double x2, double y2) {
    const double dx = x2 - x1, dy = y2 - y1;
    const double len2 = dx * dx + dy * dy;
    double t = len2 == 0 ? 0 : ((x - x1) * dx + (y - y1) * dy) / len2;
    t = t > 1 ? 1 : t;
    t = t < 0 ? 0 : t;
    const double px = x1 + t * dx, py = y1 + t * dy;
    return hypot(x - px, y - py);
}

double orc_polygons_distance2(const double* xy, const int* ring_start, int rings, double x, double y) {
    double best = INFINITY;
    for (int r = 0; r < rings; ++r) {
        const double* ring = xy + 2 * ring_start[r];
        const int count = ring_start[r + 1] - ring_start[r];
        double dist = INFINITY;
        for (int i = 0, j = count - 1; i < count; j = i++) {
            const double d = segment_distance(x, y, ring[2 * j], ring[2 * j + 1], ring[2 * i], ring[2 * i + 1]);
            if (d < dist) dist = d;
        }
        const double s = point_in_ring(x, y, ring, count) ? -dist : dist;
        if (s < best) best = s;
    }
    return best;
}

/* ---- addPolygons: sdf_collider2.js:82-100, the SDF at cell centres ---- */
void orc_polygon_sdf2(const double* xy, const int* ring_start, int rings, int invert, const int n[2],
                      const double h[2], const double origin[2], float* sdf) {
    const double ox = origin[0] + 0.5 * h[0], oy = origin[1] + 0.5 * h[1];
    for (int j = 0; j < n[1]; ++j) {
        const double y = oy + j * h[1];
        for (int i = 0; i < n[0]; ++i) {
            const double x = ox + i * h[0];
            const double d = orc_polygons_distance2(xy, ring_start, rings, x, y);
            sdf[i + (int64_t)n[0] * j] = (float)(invert ? -d : d);
        }
    }
}

/* ---- bilinearCoordsAndWeights2 + collocatedValueAtPosition2:
 *      grid_math.js:35-76 ---- */
float orc_sample_bilinear2(const float* d, const int s[2], const float o[2], const float h[2], float px, float py) {
    const float gx = (px - o[0]) / h[0], gy = (py - o[1]) / h[1];
    const int i0 = (int)floorf(gx), j0 = (int)floorf(gy);
    const float fx = gx - (float)i0, fy = gy - (float)j0;
    const int i0c = clampi2(i0, 0, s[0] - 1), i1c = clampi2(i0 + 1, 0, s[0] - 1);
    const int j0c = clampi2(j0, 0, s[1] - 1), j1c = clampi2(j0 + 1, 0, s[1] - 1);
    const float ax = 1.0f - fx, ay = 1.0f - fy;
    const float w00 = ax * ay, w10 = fx * ay, w01 = ax * fy, w11 = fx * fy;
    const int64_t r0 = (int64_t)s[0] * j0c, r1 = (int64_t)s[0] * j1c;
    return ((d[i0c + r0] * w00 + d[i1c + r0] * w10) + d[i0c + r1] * w01) + d[i1c + r1] * w11;
}

/* faceCenteredValueAtPosition2 (grid_math.js:78-88) on the MAC grid of n
 * cells: u at (o.x, o.y + h.y/2), v at (o.x + h.x/2, o.y) (grid_data2.js:166-178) */
static void sample_mac2(const float* u, const float* v, const int n[2], const float h[2], const float origin[2],
                        float px, float py, float out[2]) {
    const int su[2] = {n[0] + 1, n[1]}, sv[2] = {n[0], n[1] + 1};
    const float ou[2] = {origin[0] + 0.0f, origin[1] + h[1] * 0.5f};
    const float ov[2] = {origin[0] + h[0] * 0.5f, origin[1] + 0.0f};
    out[0] = orc_sample_bilinear2(u, su, ou, h, px, py);
    out[1] = orc_sample_bilinear2(v, sv, ov, h, px, py);
}

/* ---- bilinearGradientAtPosition2: grid_math.js:285-328 ---- */
static void grad_axis2(float coord, int n, int* lo, int* hi, float* t) {
    const int idx = (int)floorf(coord);
    int l = idx < n - 2 ? idx : n - 2;
    l = l > 0 ? l : 0;
    const int hh = l + 1 < n - 1 ? l + 1 : n - 1;
    const float tt = coord - (float)l;
    *lo = l;
    *hi = hh;
    *t = tt < 0 ? 0 : (tt > 1 ? 1 : tt);
}

static void sdf_gradient2(const float* d, const int s[2], const float o[2], const float h[2], float px, float py,
                          float g[2]) {
    int xl, xh, yl, yh;
    float tx, ty;
    grad_axis2((px - o[0]) / h[0], s[0], &xl, &xh, &tx);
    grad_axis2((py - o[1]) / h[1], s[1], &yl, &yh, &ty);
    const float v00 = d[xl + (int64_t)s[0] * yl], v10 = d[xh + (int64_t)s[0] * yl];
    const float v01 = d[xl + (int64_t)s[0] * yh], v11 = d[xh + (int64_t)s[0] * yh];
    g[0] = ((1.0f - ty) * (v10 - v00) + ty * (v11 - v01)) / h[0];
    g[1] = ((1.0f - tx) * (v01 - v00) + tx * (v11 - v10)) / h[1];
}

/* ---- monotonicCubic1d + collocatedCubicValueAtPosition2:
 *      grid_math.js:102-167 ---- */
static float mono_cubic2(float f0, float f1, float f2, float f3, float f) {
    const float D1 = f2 - f1;
    const float rawD1 = (f2 - f0) * 0.5f, rawD2 = (f3 - f1) * 0.5f;
    const int flat = fabsf(D1) < 1e-12f;
    const float d1 = (flat || rawD1 * D1 < 0.0f) ? 0.0f : rawD1;
    const float d2 = (flat || rawD2 * D1 < 0.0f) ? 0.0f : rawD2;
    const float a3 = (d1 + d2) - D1 * 2.0f;
    const float a2 = (D1 * 3.0f - d1 * 2.0f) - d2;
    return ((((a3 * f) * f) * f + (a2 * f) * f) + d1 * f) + f1;
}

float orc_sample_cubic2(const float* d, const int s[2], const float o[2], const float h[2], float px, float py) {
    int xi[4], yi[4];
    float f[2];
    const float coord[2] = {(px - o[0]) / h[0], (py - o[1]) / h[1]};
    int* idx[2] = {xi, yi};
    for (int a = 0; a < 2; ++a) {
        const int i0 = (int)floorf(coord[a]);
        f[a] = coord[a] - (float)i0;
        idx[a][0] = clampi2(i0 - 1, 0, s[a] - 1);
        idx[a][1] = clampi2(i0, 0, s[a] - 1);
        idx[a][2] = clampi2(i0 + 1, 0, s[a] - 1);
        idx[a][3] = clampi2(i0 + 2, 0, s[a] - 1);
    }
    float rv[4];
    for (int b = 0; b < 4; ++b) {
        const int64_t base = (int64_t)s[0] * yi[b];
        rv[b] = mono_cubic2(d[xi[0] + base], d[xi[1] + base], d[xi[2] + base], d[xi[3] + base], f[0]);
    }
    return mono_cubic2(rv[0], rv[1], rv[2], rv[3], f[1]);
}

/* ---- backTrace with the collider: advection_solver2.js:81-127.  Midpoint
 *      sub-steps of at most one cell; a step whose end lies inside the
 *      collider (phi <= 0) ends at the linear estimate of the surface. ---- */
static void back_trace2(const float* u, const float* v, const int n[2], const float h[2], const float origin[2],
                        const float* sdf, float sx, float sy, float dt, int max_sub, float direction, float out[2]) {
    const int sc[2] = {n[0], n[1]};
    const float so[2] = {origin[0] + h[0] * 0.5f, origin[1] + h[1] * 0.5f};
    const float hmin = h[0] < h[1] ? h[0] : h[1];
    float px = sx, py = sy, rem = dt;
    for (int s = 0; s < max_sub; ++s) {
        if (rem <= 1e-6f) break;
        float v0[2], mv[2];
        sample_mac2(u, v, n, h, origin, px, py, v0);
        const float len = sqrtf(v0[0] * v0[0] + v0[1] * v0[1]);
        float ns = ceilf((len * rem) / hmin);
        if (ns < 1.0f) ns = 1.0f;
        const float sub = rem / ns;
        const float half = sub * 0.5f;
        const float mx = px - (v0[0] * half) * direction, my = py - (v0[1] * half) * direction;
        sample_mac2(u, v, n, h, origin, mx, my, mv);
        const float nx = px - (mv[0] * sub) * direction, ny = py - (mv[1] * sub) * direction;
        const float phi0 = sdf ? orc_sample_bilinear2(sdf, sc, so, h, px, py) : 1.0f;
        const float phi1 = sdf ? orc_sample_bilinear2(sdf, sc, so, h, nx, ny) : 1.0f;
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
    out[0] = px;
    out[1] = py;
}

/* ---- constrainVelocity: grid_blocked_boundary_condition_solver2.js:75-96
 *      (friction), 251-402 (collider kernels), 520-575 (the plan), for a
 *      static collider (velocity 0) given by its cell-centred SDF, the
 *      closed-wall flags (LEFT 1, RIGHT 2, DOWN 4, UP 8) and inflows
 *      (mode 0 'set', 1 'add'), then the clamp to [-100, 100]. ---- */
static float clamp_component2(float v) {
    if (!isfinite(v)) return 0.0f;
    return v < -100.0f ? -100.0f : (v > 100.0f ? 100.0f : v);
}

void orc_constrain_velocity2(float* u, float* v, const int n[2], const float h[2], const float origin[2],
                             int closed_flags, const float* sdf, float friction, int depth, int n_inflows,
                             const float* const* inflow_sdf, const float* inflow_vel, const int* inflow_mode) {
    const int sc[2] = {n[0], n[1]};
    const float so[2] = {origin[0] + h[0] * 0.5f, origin[1] + h[1] * 0.5f};
    const int sz[2][2] = {{n[0] + 1, n[1]}, {n[0], n[1] + 1}};
    const float oo[2][2] = {{origin[0] + 0.0f, origin[1] + h[1] * 0.5f}, {origin[0] + h[0] * 0.5f, origin[1] + 0.0f}};
    float* data[2] = {u, v};
    const int64_t cnt[2] = {(int64_t)sz[0][0] * sz[0][1], (int64_t)sz[1][0] * sz[1][1]};
    if (sdf) {
        int32_t* marker[2];
        float* tmp[2];
        for (int a = 0; a < 2; ++a) {
            marker[a] = (int32_t*)malloc(sizeof(int32_t) * cnt[a]);
            tmp[a] = (float*)malloc(sizeof(float) * cnt[a]);
        }
        /* fillMarker + markAndProject (258-303) */
        for (int a = 0; a < 2; ++a)
            for (int j = 0; j < sz[a][1]; ++j)
                for (int i = 0; i < sz[a][0]; ++i) {
                    const int64_t f = i + (int64_t)sz[a][0] * j;
                    const float px = oo[a][0] + h[0] * (float)i, py = oo[a][1] + h[1] * (float)j;
                    float s0, s1;
                    if (a == 0) {
                        s0 = orc_sample_bilinear2(sdf, sc, so, h, px - h[0] * 0.5f, py - 0.0f);
                        s1 = orc_sample_bilinear2(sdf, sc, so, h, px + h[0] * 0.5f, py + 0.0f);
                    } else {
                        s0 = orc_sample_bilinear2(sdf, sc, so, h, px - 0.0f, py - h[1] * 0.5f);
                        s1 = orc_sample_bilinear2(sdf, sc, so, h, px + 0.0f, py + h[1] * 0.5f);
                    }
                    float fr;
                    const int in0 = s0 < 0, in1 = s1 < 0;
                    if (in0 && in1) fr = 1.0f;
                    else if (in0) fr = s0 / (s0 - s1);
                    else if (in1) fr = s1 / (s1 - s0);
                    else fr = 0.0f;
                    fr = fminf(fmaxf(fr, 0.0f), 1.0f);
                    const float open = 1.0f - fr;
                    if (open > 0) {
                        marker[a][f] = 1;
                    } else {
                        data[a][f] = 0.0f;
                        marker[a][f] = 0;
                    }
                }
        /* createExtrapolateToRegion2 (array_utils.js:143-295), in place */
        for (int a = 0; a < 2; ++a) {
            const int n3[3] = {sz[a][0], sz[a][1], 1};
            orc_extrapolate3(n3, data[a], marker[a], data[a], depth);
        }
        /* noFluxProjection into tmp (305-370), from the extrapolated velocity */
        for (int a = 0; a < 2; ++a)
            for (int j = 0; j < sz[a][1]; ++j)
                for (int i = 0; i < sz[a][0]; ++i) {
                    const int64_t f = i + (int64_t)sz[a][0] * j;
                    const float px = oo[a][0] + h[0] * (float)i, py = oo[a][1] + h[1] * (float)j;
                    if (!(orc_sample_bilinear2(sdf, sc, so, h, px, py) < 0)) {
                        tmp[a][f] = data[a][f];
                        continue;
                    }
                    float vel[2], g[2];
                    sample_mac2(u, v, n, h, origin, px, py, vel);
                    sdf_gradient2(sdf, sc, so, h, px, py, g);
                    const float gl = sqrtf(g[0] * g[0] + g[1] * g[1]);
                    if (!(gl > 0)) {
                        tmp[a][f] = 0.0f;
                        continue;
                    }
                    const float nn[2] = {g[0] / gl, g[1] / gl};
                    const float vr[2] = {vel[0] - 0.0f, vel[1] - 0.0f};
                    const float dn = vr[0] * nn[0] + vr[1] * nn[1];
                    float vt[2] = {vr[0] - nn[0] * dn, vr[1] - nn[1] * dn};
                    const float tl = sqrtf(vt[0] * vt[0] + vt[1] * vt[1]);
                    if (tl > 0) {
                        const float veln = fmaxf(-dn, 0.0f);
                        const float scale = fmaxf(1.0f - (veln * friction) / tl, 0.0f);
                        vt[0] *= scale;
                        vt[1] *= scale;
                    }
                    tmp[a][f] = vt[a] + 0.0f;
                }
        for (int a = 0; a < 2; ++a) memcpy(data[a], tmp[a], sizeof(float) * cnt[a]);
        /* blockedBoundary (372-402): collider cells with a fluid neighbour */
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
#define S2(ii, jj) sdf[(ii) + (int64_t)n[0] * (jj)]
                if (!(S2(i, j) < 0)) continue;
                if (i > 0 && !(S2(i - 1, j) < 0)) u[i + (int64_t)sz[0][0] * j] = 0;
                if (i < n[0] - 1 && !(S2(i + 1, j) < 0)) u[i + 1 + (int64_t)sz[0][0] * j] = 0;
                if (j > 0 && !(S2(i, j - 1) < 0)) v[i + (int64_t)sz[1][0] * j] = 0;
                if (j < n[1] - 1 && !(S2(i, j + 1) < 0)) v[i + (int64_t)sz[1][0] * (j + 1)] = 0;
#undef S2
            }
        for (int a = 0; a < 2; ++a) {
            free(marker[a]);
            free(tmp[a]);
        }
    }
    for (int j = 0; j < n[1]; ++j) {
        if (closed_flags & 1) u[0 + (int64_t)sz[0][0] * j] = 0;
        if (closed_flags & 2) u[n[0] + (int64_t)sz[0][0] * j] = 0;
    }
    for (int i = 0; i < n[0]; ++i) {
        if (closed_flags & 4) v[i] = 0;
        if (closed_flags & 8) v[i + (int64_t)sz[1][0] * n[1]] = 0;
    }
    /* inflows (167-200) */
    for (int q = 0; q < n_inflows; ++q)
        for (int a = 0; a < 2; ++a)
            for (int j = 0; j < sz[a][1]; ++j)
                for (int i = 0; i < sz[a][0]; ++i) {
                    const float px = oo[a][0] + h[0] * (float)i, py = oo[a][1] + h[1] * (float)j;
                    if (!(orc_sample_bilinear2(inflow_sdf[q], sc, so, h, px, py) < 0)) continue;
                    float* d = &data[a][i + (int64_t)sz[a][0] * j];
                    if (inflow_mode[q] == 1) *d = *d + inflow_vel[2 * q + a];
                    else *d = inflow_vel[2 * q + a];
                }
    for (int a = 0; a < 2; ++a)
        for (int64_t f = 0; f < cnt[a]; ++f) data[a][f] = clamp_component2(data[a][f]);
}

/* ---- applyOutflowVelocityBC: grid_outflow_solver2.js:245-365.  Faces
 *      inside the outflow relax toward the velocity one cell upstream
 *      (along the outflow SDF's normal) by the outward Courant number;
 *      U is written back before V is computed. ---- */
void orc_outflow_bc2(float* u, float* v, const int n[2], const float h[2], const float origin[2], float dt,
                     const float* osdf) {
    const int sc[2] = {n[0], n[1]};
    const float so[2] = {origin[0] + h[0] * 0.5f, origin[1] + h[1] * 0.5f};
    const int sz[2][2] = {{n[0] + 1, n[1]}, {n[0], n[1] + 1}};
    const float oo[2][2] = {{origin[0] + 0.0f, origin[1] + h[1] * 0.5f}, {origin[0] + h[0] * 0.5f, origin[1] + 0.0f}};
    float* data[2] = {u, v};
    for (int a = 0; a < 2; ++a) {
        const int64_t cnt = (int64_t)sz[a][0] * sz[a][1];
        float* dst = (float*)malloc(sizeof(float) * cnt);
        uint8_t* inside = (uint8_t*)calloc(cnt, 1);
        const float ha = h[a];
        for (int j = 0; j < sz[a][1]; ++j)
            for (int i = 0; i < sz[a][0]; ++i) {
                const int64_t f = i + (int64_t)sz[a][0] * j;
                const float px = oo[a][0] + h[0] * (float)i, py = oo[a][1] + h[1] * (float)j;
                if (!(orc_sample_bilinear2(osdf, sc, so, h, px, py) < 0)) continue;
                inside[f] = 1;
                float g[2];
                sdf_gradient2(osdf, sc, so, h, px, py, g);
                const float gl = sqrtf(g[0] * g[0] + g[1] * g[1]);
                if (!(gl > 0)) {
                    dst[f] = data[a][f];
                    continue;
                }
                const float nn[2] = {g[0] / gl, g[1] / gl};
                const float ux = px + nn[0] * ha, uy = py + nn[1] * ha;
                float bulk[2], up[2];
                sample_mac2(u, v, n, h, origin, px, py, bulk);
                const float dn = bulk[0] * -nn[0] + bulk[1] * -nn[1];
                const float outward = fmaxf(0.0f, dn);
                const float courant = fminf(fmaxf((outward * dt) / ha, 0.0f), 1.0f);
                sample_mac2(u, v, n, h, origin, ux, uy, up);
                const float cur = data[a][f];
                const float relaxed = cur + (up[a] - cur) * courant;
                dst[f] = fminf(fmaxf(relaxed, -1e3f), 1e3f);
            }
        for (int64_t f = 0; f < cnt; ++f)
            if (inside[f]) data[a][f] = dst[f];
        free(dst);
        free(inside);
    }
}

/* ---- createGridPressureSolver2 system build: grid_pressure_solver2.js:677-785.
 *      Fluid row b = divergence - sum of the Dirichlet neighbours'
 *      target / h^2 (order -x, +x, -y, +y); a Dirichlet row has
 *      b = -target and the pressure seeded with the target.  `active` may be
 *      NULL (no Dirichlet cells). ---- */
void orc_pressure2_assemble(const int n[2], const float h[2], const float* u, const float* v, const uint8_t* active,
                            const float* target, float* b, float* p) {
    const float ih2x = (float)(1.0 / ((double)h[0] * (double)h[0]));
    const float ih2y = (float)(1.0 / ((double)h[1] * (double)h[1]));
    const int su = n[0] + 1;
    for (int j = 0; j < n[1]; ++j)
        for (int i = 0; i < n[0]; ++i) {
            const int64_t c = i + (int64_t)n[0] * j;
            const float div = (u[i + 1 + (int64_t)su * j] - u[i + (int64_t)su * j]) / h[0] +
                              (v[i + (int64_t)n[0] * (j + 1)] - v[i + (int64_t)n[0] * j]) / h[1];
            if (!active) {
                b[c] = div;
                continue;
            }
            if (active[c]) {
                b[c] = -target[c];
                p[c] = target[c];
                continue;
            }
            float t0 = 0.0f, t1 = 0.0f, t2 = 0.0f, t3 = 0.0f;
            if (i > 0 && active[c - 1]) t0 = (1.0f * target[c - 1]) * ih2x;
            if (i < n[0] - 1 && active[c + 1]) t1 = (1.0f * target[c + 1]) * ih2x;
            if (j > 0 && active[c - n[0]]) t2 = (1.0f * target[c - n[0]]) * ih2y;
            if (j < n[1] - 1 && active[c + n[0]]) t3 = (1.0f * target[c + n[0]]) * ih2y;
            b[c] = div - (((t0 + t1) + t2) + t3);
        }
}

/* ---- dispatchCorrectU/V: grid_pressure_solver2.js:796-834, interior
 *      faces only (boundary faces keep their values) ---- */
void orc_pressure2_correct(const int n[2], const float h[2], const float* p, float* u, float* v) {
    const int su = n[0] + 1;
    for (int j = 0; j < n[1]; ++j)
        for (int k = 1; k < n[0]; ++k) {
            const float g = (p[k + (int64_t)n[0] * j] - p[k - 1 + (int64_t)n[0] * j]) / h[0];
            u[k + (int64_t)su * j] = u[k + (int64_t)su * j] - g;
        }
    for (int k = 1; k < n[1]; ++k)
        for (int i = 0; i < n[0]; ++i) {
            const float g = (p[i + (int64_t)n[0] * k] - p[i + (int64_t)n[0] * (k - 1)]) / h[1];
            v[i + (int64_t)n[0] * k] = v[i + (int64_t)n[0] * k] - g;
        }
}

/* ---- createMaxAbsReducer: reduction.js:43-77 -- the largest |x| by its
 *      IEEE bit pattern (the atomicMax of floatBitsToUint) ---- */
float orc_max_abs(const float* const* arrays, const int64_t* counts, int narrays) {
    uint32_t best = 0;
    for (int q = 0; q < narrays; ++q)
        for (int64_t t = 0; t < counts[q]; ++t) {
            float a = fabsf(arrays[q][t]);
            uint32_t bits;
            memcpy(&bits, &a, 4);
            if ((int32_t)bits > (int32_t)best) best = bits;
        }
    float out;
    memcpy(&out, &best, 4);
    return out;
}

/* ---- computeAdaptiveSubSteps: time/cfl.js:38-48 ---- */
int orc_adaptive_substeps(double max_velocity, double min_spacing, double frame_dt, double courant, int max_sub,
                          double* sub_dt) {
    const double cfl = max_velocity * frame_dt / min_spacing;
    double k = ceil(cfl / courant);
    if (!(k >= 1)) k = 1; /* also a non-finite cfl, which the reference would pass on as NaN */
    if (k > max_sub) k = max_sub;
    const int steps = (int)k;
    *sub_dt = frame_dt / steps;
    return steps;
}

/* ---- createOutflowPressureDirichlet2: sdf_inflow_outflow2.js:152-166, the
 *      pressure solver evaluating it at its cell positions
 *      (grid_pressure_solver2.js:403-417): active where the outflow SDF
 *      sampled at the cell centre is negative; the target is 0 ---- */
void orc_outflow_dirichlet2(const float* osdf, const int n[2], const float h[2], const float origin[2],
                            uint8_t* active) {
    const int sc[2] = {n[0], n[1]};
    const float so[2] = {origin[0] + h[0] * 0.5f, origin[1] + h[1] * 0.5f};
    for (int j = 0; j < n[1]; ++j)
        for (int i = 0; i < n[0]; ++i) {
            const float px = so[0] + h[0] * (float)i, py = so[1] + h[1] * (float)j;
            active[i + (int64_t)n[0] * j] = orc_sample_bilinear2(osdf, sc, so, h, px, py) < 0;
        }
}

/* ---- advectFaceCentered2: advection_solver2.js:263-320 (order 1),
 *      221-258 and 310-430 (order 2, MacCormack).  The trace follows the
 *      solver's velocity grid (u, v) and stops at the collider; the value is
 *      the bicubic sample of the advected grid (du, dv).  A face whose
 *      position lies inside the collider keeps the output's value (uo/vo
 *      hold its previous contents).  Order 2: forward = order 1 (the input
 *      inside the collider), backward = the forward result sampled at the
 *      trace with direction -1, then fwd + (in - bwd) / 2 where it lies in
 *      the range of the 4 bilinear corners of the forward trace whose face
 *      positions lie outside the collider, else fwd. ---- */
void orc_advect_faces2_vd(const float* u, const float* v, const float* du, const float* dv, const int n[2],
                          const float h[2], const float origin[2], const float* sdf, float dt, int max_sub, int order,
                          float* uo, float* vo) {
    const int sc[2] = {n[0], n[1]};
    const float so[2] = {origin[0] + h[0] * 0.5f, origin[1] + h[1] * 0.5f};
    const int sz[2][2] = {{n[0] + 1, n[1]}, {n[0], n[1] + 1}};
    const float oo[2][2] = {{origin[0] + 0.0f, origin[1] + h[1] * 0.5f}, {origin[0] + h[0] * 0.5f, origin[1] + 0.0f}};
    const float* src[2] = {du, dv};
    float* dst[2] = {uo, vo};
    float* fwd[2] = {uo, vo};
    if (order == 2)
        for (int a = 0; a < 2; ++a) fwd[a] = (float*)malloc(sizeof(float) * (size_t)sz[a][0] * sz[a][1]);
    for (int a = 0; a < 2; ++a)
        for (int j = 0; j < sz[a][1]; ++j)
            for (int i = 0; i < sz[a][0]; ++i) {
                const int64_t f = i + (int64_t)sz[a][0] * j;
                const float px = oo[a][0] + h[0] * (float)i, py = oo[a][1] + h[1] * (float)j;
                if (sdf && !(orc_sample_bilinear2(sdf, sc, so, h, px, py) > 0.0f)) {
                    if (order == 2) fwd[a][f] = src[a][f];
                    continue;
                }
                float t[2];
                back_trace2(u, v, n, h, origin, sdf, px, py, dt, max_sub, 1.0f, t);
                fwd[a][f] = orc_sample_cubic2(src[a], sz[a], oo[a], h, t[0], t[1]);
            }
    if (order != 2) return;
    for (int a = 0; a < 2; ++a)
        for (int j = 0; j < sz[a][1]; ++j)
            for (int i = 0; i < sz[a][0]; ++i) {
                const int64_t f = i + (int64_t)sz[a][0] * j;
                const float px = oo[a][0] + h[0] * (float)i, py = oo[a][1] + h[1] * (float)j;
                if (sdf && !(orc_sample_bilinear2(sdf, sc, so, h, px, py) > 0.0f)) continue;
                float tb[2], tf[2];
                back_trace2(u, v, n, h, origin, sdf, px, py, dt, max_sub, -1.0f, tb);
                const float bwd = orc_sample_cubic2(fwd[a], sz[a], oo[a], h, tb[0], tb[1]);
                back_trace2(u, v, n, h, origin, sdf, px, py, dt, max_sub, 1.0f, tf);
                const float fv = fwd[a][f];
                const float corrected = fv + (src[a][f] - bwd) * 0.5f;
                const float gx = (tf[0] - oo[a][0]) / h[0], gy = (tf[1] - oo[a][1]) / h[1];
                const int i0 = (int)floorf(gx), j0 = (int)floorf(gy);
                const int ci[2] = {clampi2(i0, 0, sz[a][0] - 1), clampi2(i0 + 1, 0, sz[a][0] - 1)};
                const int cj[2] = {clampi2(j0, 0, sz[a][1] - 1), clampi2(j0 + 1, 0, sz[a][1] - 1)};
                float mn = 1e30f, mx = -1e30f;
                int have = 0;
                for (int c = 0; c < 4; ++c) {
                    const int ni = ci[c & 1], nj = cj[c >> 1];
                    const float qx = oo[a][0] + h[0] * (float)ni, qy = oo[a][1] + h[1] * (float)nj;
                    const int valid = sdf ? orc_sample_bilinear2(sdf, sc, so, h, qx, qy) > 0.0f : 1;
                    const float val = src[a][ni + (int64_t)sz[a][0] * nj];
                    if (valid) {
                        mn = fminf(mn, val);
                        mx = fmaxf(mx, val);
                        have = 1;
                    }
                }
                dst[a][f] = (have && corrected >= mn && corrected <= mx) ? corrected : fv;
            }
    for (int a = 0; a < 2; ++a) free(fwd[a]);
}

void orc_advect_face2(const float* u, const float* v, const int n[2], const float h[2], const float origin[2],
                      const float* sdf, float dt, int max_sub, float* uo, float* vo) {
    orc_advect_faces2_vd(u, v, u, v, n, h, origin, sdf, dt, max_sub, 1, uo, vo);
}

void orc_advect_face2_mc(const float* u, const float* v, const int n[2], const float h[2], const float origin[2],
                         const float* sdf, float dt, int max_sub, float* uo, float* vo) {
    orc_advect_faces2_vd(u, v, u, v, n, h, origin, sdf, dt, max_sub, 2, uo, vo);
}

/* ---- trace (advection_solver2.js:197-201) of each point, with the
 *      collider's crossing clamp; exposed for the same reason as
 *      orc_trace3. ---- */
void orc_trace2(float* pos, int64_t count, const float* u, const float* v, const int n[2], const float h[2],
                const float origin[2], const float* sdf, float dt, int max_sub, float direction) {
    for (int64_t p = 0; p < count; ++p) {
        float t[2];
        back_trace2(u, v, n, h, origin, sdf, pos[2 * p], pos[2 * p + 1], dt, max_sub, direction, t);
        pos[2 * p] = t[0];
        pos[2 * p + 1] = t[1];
    }
}
