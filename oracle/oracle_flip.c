/*
 * Oracle for the FLIP grid stages around the projection (TEST
 * INFRASTRUCTURE ONLY): extrapolation, fluid marking, fluid-adjacent faces,
 * collider face weights, velocity constraint, resampling and the stage
 * order of one FLIP step.  f32 arithmetic in the reference's operation
 * order (built with -ffp-contract=off); the collider face weights follow the
 * reference's host code in f64.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_util.h"

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* ---- createExtrapolateStepKernel3 + createExtrapolateToRegion3:
 *      array_utils3.js:88-162, 166-228.  `iterations` Jacobi steps; an
 *      invalid cell takes the mean of its valid 6-neighbours summed in the
 *      order +x, -x, +y, -y, +z, -z and becomes valid; a cell with none keeps
 *      its value and stays invalid. ---- */
void orc_extrapolate3(const int n[3], const float* in, const int32_t* valid_in, float* out, int iterations) {
    const int64_t cells = (int64_t)n[0] * n[1] * n[2];
    float* va = (float*)malloc(sizeof(float) * cells);
    float* vb = (float*)malloc(sizeof(float) * cells);
    int32_t* ma = (int32_t*)malloc(sizeof(int32_t) * cells);
    int32_t* mb = (int32_t*)malloc(sizeof(int32_t) * cells);
    memcpy(va, in, sizeof(float) * cells);
    memcpy(ma, valid_in, sizeof(int32_t) * cells);
    for (int it = 0; it < iterations; ++it) {
        for (int k = 0; k < n[2]; ++k)
            for (int j = 0; j < n[1]; ++j)
                for (int i = 0; i < n[0]; ++i) {
                    int64_t c = i + (int64_t)n[0] * (j + (int64_t)n[1] * k);
                    if (ma[c] != 0) {
                        vb[c] = va[c];
                        mb[c] = 1;
                        continue;
                    }
                    float total = 0.0f;
                    int count = 0;
                    const int64_t sx = 1, sy = n[0], sz = (int64_t)n[0] * n[1];
                    if (i + 1 < n[0] && ma[c + sx]) { total += va[c + sx]; ++count; }
                    if (i > 0 && ma[c - sx]) { total += va[c - sx]; ++count; }
                    if (j + 1 < n[1] && ma[c + sy]) { total += va[c + sy]; ++count; }
                    if (j > 0 && ma[c - sy]) { total += va[c - sy]; ++count; }
                    if (k + 1 < n[2] && ma[c + sz]) { total += va[c + sz]; ++count; }
                    if (k > 0 && ma[c - sz]) { total += va[c - sz]; ++count; }
                    if (count > 0) {
                        vb[c] = total / (float)count;
                        mb[c] = 1;
                    } else {
                        vb[c] = va[c];
                        mb[c] = 0;
                    }
                }
        float* t = va; va = vb; vb = t;
        int32_t* tm = ma; ma = mb; mb = tm;
    }
    memcpy(out, va, sizeof(float) * cells);
    free(va); free(vb); free(ma); free(mb);
}

/* ---- clearFluidMask + markFluidCellsKernel: grid_flip_solver3.js:216-245 ---- */
void orc_mark_fluid3(const float* pos, int64_t count, const float origin[3], const float h[3], const int n[3],
                     float* mask) {
    const int64_t cells = (int64_t)n[0] * n[1] * n[2];
    memset(mask, 0, sizeof(float) * cells);
    int32_t* cell = (int32_t*)malloc(sizeof(int32_t) * (count > 0 ? count : 1));
    orc_cell_index3(pos, count, origin, h, n, cell);
    for (int64_t p = 0; p < count; ++p) mask[cell[p]] = 1.0f;
    free(cell);
}

/* ---- computeU/V/WFluidAdjacent: grid_flip_solver3.js:595-625 ---- */
void orc_fluid_adjacent3(const float* mask, const int n[3], int axis, int32_t* out) {
    int s[3] = {n[0], n[1], n[2]};
    s[axis] += 1;
    for (int k = 0; k < s[2]; ++k)
        for (int j = 0; j < s[1]; ++j)
            for (int i = 0; i < s[0]; ++i) {
                int I[3] = {i, j, k};
                int lo[3] = {i, j, k}, hi[3] = {i, j, k};
                lo[axis] = clampi(I[axis] - 1, 0, n[axis] - 1);
                hi[axis] = clampi(I[axis], 0, n[axis] - 1);
                float a = mask[lo[0] + (int64_t)n[0] * (lo[1] + (int64_t)n[1] * lo[2])];
                float b = mask[hi[0] + (int64_t)n[0] * (hi[1] + (int64_t)n[1] * hi[2])];
                out[i + (int64_t)s[0] * (j + (int64_t)s[1] * k)] = (a > 0.5f || b > 0.5f) ? 1 : 0;
            }
}

/* ---- primitivesDistanceCpu: primitive_sdf3.js:149-208 ---- */
static double prim_distance(const orc_primitive3* sh, double x, double y, double z) {
    double lx = x - sh->center[0], ly = y - sh->center[1], lz = z - sh->center[2];
    if (sh->type == 0) return hypot(hypot(lx, ly), lz) - sh->radius;
    double q[3] = {lx, ly, lz};
    if (sh->rotated) {
        /* rotateAroundAxisCpu(v, normalizeAxis(axis), -angle) */
        double ax = sh->axis[0], ay = sh->axis[1], az = sh->axis[2];
        double len = hypot(hypot(ax, ay), az);
        double kx = 0, ky = 0, kz = 1;
        if (len > 0) { kx = ax / len; ky = ay / len; kz = az / len; }
        double ang = -sh->angle;
        double c = cos(ang), s = sin(ang);
        double cx = ky * lz - kz * ly, cy = kz * lx - kx * lz, cz = kx * ly - ky * lx;
        double kd = kx * lx + ky * ly + kz * lz;
        q[0] = lx * c + cx * s + kx * kd * (1 - c);
        q[1] = ly * c + cy * s + ky * kd * (1 - c);
        q[2] = lz * c + cz * s + kz * kd * (1 - c);
    }
    double dx = fabs(q[0]) - sh->half[0], dy = fabs(q[1]) - sh->half[1], dz = fabs(q[2]) - sh->half[2];
    double outside = hypot(hypot(fmax(dx, 0), fmax(dy, 0)), fmax(dz, 0));
    double inside = fmin(fmax(dx, fmax(dy, dz)), 0);
    return outside + inside;
}

double orc_primitives_distance3(const orc_primitive3* shapes, int count, double x, double y, double z) {
    double d = INFINITY;
    for (int q = 0; q < count; ++q) d = fmin(d, prim_distance(&shapes[q], x, y, z));
    return d;
}

/* ---- fractionInsideSdfCpu / fractionInsideCpu: level_set_utils.js:51-161 ---- */
static double frac_sdf(double p0, double p1) {
    int in0 = p0 < 0, in1 = p1 < 0;
    if (in0 && in1) return 1;
    if (in0 && !in1) return p0 / (p0 - p1);
    if (!in0 && in1) return p1 / (p1 - p0);
    return 0;
}

static void cycle4(double* a) {
    double t = a[0];
    a[0] = a[1]; a[1] = a[2]; a[2] = a[3]; a[3] = t;
}

double orc_fraction_inside(double bl, double br, double tl, double tr) {
    int inside = (bl < 0) + (tl < 0) + (br < 0) + (tr < 0);
    double l[4] = {bl, br, tr, tl};
    if (inside == 4) return 1;
    if (inside == 3) {
        for (int g = 0; g < 4 && l[0] < 0; ++g) cycle4(l);
        double s0 = 1 - frac_sdf(l[0], l[3]);
        double s1 = 1 - frac_sdf(l[0], l[1]);
        return 1 - 0.5 * s0 * s1;
    }
    if (inside == 2) {
        for (int g = 0; g < 4 && (l[0] >= 0 || !(l[1] < 0 || l[2] < 0)); ++g) cycle4(l);
        if (l[1] < 0) {
            double sl = frac_sdf(l[0], l[3]);
            double sr = frac_sdf(l[1], l[2]);
            return 0.5 * (sl + sr);
        }
        double mid = 0.25 * (l[0] + l[1] + l[2] + l[3]);
        if (mid < 0) {
            double area = 0;
            double s1 = 1 - frac_sdf(l[0], l[3]);
            double s3 = 1 - frac_sdf(l[2], l[3]);
            area += 0.5 * s1 * s3;
            double s2 = 1 - frac_sdf(l[2], l[1]);
            double s0 = 1 - frac_sdf(l[0], l[1]);
            area += 0.5 * s0 * s2;
            return 1 - area;
        }
        double area = 0;
        double s0 = frac_sdf(l[0], l[1]);
        double s1 = frac_sdf(l[0], l[3]);
        area += 0.5 * s0 * s1;
        double s2 = frac_sdf(l[2], l[1]);
        double s3 = frac_sdf(l[2], l[3]);
        area += 0.5 * s2 * s3;
        return area;
    }
    if (inside == 1) {
        for (int g = 0; g < 4 && l[0] >= 0; ++g) cycle4(l);
        double s0 = frac_sdf(l[0], l[3]);
        double s1 = frac_sdf(l[0], l[1]);
        return 0.5 * s0 * s1;
    }
    return 0;
}

/* ---- computeFaceWeights: sdf_collider3.js:131-262 (f64, stored as f32) ---- */
void orc_face_weights3(const orc_primitive3* shapes, int count, int invert, const int n[3], const double h[3],
                       const double origin[3], float* wu, float* wv, float* ww) {
    const double sign = invert ? -1.0 : 1.0;
    const double hx = 0.5 * h[0], hy = 0.5 * h[1], hz = 0.5 * h[2];
    for (int axis = 0; axis < 3; ++axis) {
        int s[3] = {n[0], n[1], n[2]};
        s[axis] += 1;
        double o[3] = {origin[0] + hx, origin[1] + hy, origin[2] + hz};
        o[axis] = origin[axis];
        double off[4][3];
        if (axis == 0) {
            double t[4][3] = {{0, -hy, -hz}, {0, hy, -hz}, {0, -hy, hz}, {0, hy, hz}};
            memcpy(off, t, sizeof(off));
        } else if (axis == 1) {
            double t[4][3] = {{-hx, 0, -hz}, {hx, 0, -hz}, {-hx, 0, hz}, {hx, 0, hz}};
            memcpy(off, t, sizeof(off));
        } else {
            double t[4][3] = {{-hx, -hy, 0}, {hx, -hy, 0}, {-hx, hy, 0}, {hx, hy, 0}};
            memcpy(off, t, sizeof(off));
        }
        float* out = axis == 0 ? wu : (axis == 1 ? wv : ww);
        for (int k = 0; k < s[2]; ++k) {
            double z = o[2] + k * h[2];
            for (int j = 0; j < s[1]; ++j) {
                double y = o[1] + j * h[1];
                for (int i = 0; i < s[0]; ++i) {
                    double x = o[0] + i * h[0];
                    double c[4];
                    for (int q = 0; q < 4; ++q)
                        c[q] = sign * orc_primitives_distance3(shapes, count, x + off[q][0], y + off[q][1],
                                                               z + off[q][2]);
                    double frac = orc_fraction_inside(c[0], c[1], c[2], c[3]);
                    double w = fmin(1, fmax(0, 1 - frac));
                    if (c[0] <= 0 && c[1] <= 0 && c[2] <= 0 && c[3] <= 0) w = 0;
                    else if (w > 0 && w < 0.01) w = 0.01;  /* MIN_FACE_WEIGHT */
                    out[i + (int64_t)s[0] * (j + (int64_t)s[1] * k)] = (float)w;
                }
            }
        }
    }
}

/* ---- addShapes: sdf_collider3.js:71-110, the SDF at cell centres (f64 -> f32) ---- */
void orc_collider_sdf3(const orc_primitive3* shapes, int count, int invert, const int n[3], const double h[3],
                       const double origin[3], float* sdf) {
    const double ox = origin[0] + 0.5 * h[0], oy = origin[1] + 0.5 * h[1], oz = origin[2] + 0.5 * h[2];
    for (int k = 0; k < n[2]; ++k) {
        double z = oz + k * h[2];
        for (int j = 0; j < n[1]; ++j) {
            double y = oy + j * h[1];
            for (int i = 0; i < n[0]; ++i) {
                double x = ox + i * h[0];
                double d = orc_primitives_distance3(shapes, count, x, y, z);
                sdf[i + (int64_t)n[0] * (j + (int64_t)n[1] * k)] = (float)(invert ? -d : d);
            }
        }
    }
}

/* ---- trilinearGradientAtPosition3: grid_math3.js:295-356 ---- */
static void grad_axis(float coord, int n, int* lo, int* hi, float* t) {
    int idx = (int)floorf(coord);
    int l = idx < n - 2 ? idx : n - 2;
    l = l > 0 ? l : 0;
    int hh = l + 1 < n - 1 ? l + 1 : n - 1;
    float tt = coord - (float)l;
    *lo = l;
    *hi = hh;
    *t = tt < 0 ? 0 : (tt > 1 ? 1 : tt);
}

static void sdf_gradient3(const float* d, const int n[3], const float o[3], const float h[3], const float p[3],
                          float g[3]) {
    int xl, xh, yl, yh, zl, zh;
    float tx, ty, tz;
    grad_axis((p[0] - o[0]) / h[0], n[0], &xl, &xh, &tx);
    grad_axis((p[1] - o[1]) / h[1], n[1], &yl, &yh, &ty);
    grad_axis((p[2] - o[2]) / h[2], n[2], &zl, &zh, &tz);
#define D(i, j, k) d[(i) + (int64_t)n[0] * ((j) + (int64_t)n[1] * (k))]
    float v000 = D(xl, yl, zl), v100 = D(xh, yl, zl), v010 = D(xl, yh, zl), v110 = D(xh, yh, zl);
    float v001 = D(xl, yl, zh), v101 = D(xh, yl, zh), v011 = D(xl, yh, zh), v111 = D(xh, yh, zh);
#undef D
    g[0] = ((((1.0f - ty) * (1.0f - tz)) * (v100 - v000) + (ty * (1.0f - tz)) * (v110 - v010)) +
            ((1.0f - ty) * tz) * (v101 - v001) + (ty * tz) * (v111 - v011)) / h[0];
    g[1] = ((((1.0f - tx) * (1.0f - tz)) * (v010 - v000) + (tx * (1.0f - tz)) * (v110 - v100)) +
            ((1.0f - tx) * tz) * (v011 - v001) + (tx * tz) * (v111 - v101)) / h[1];
    g[2] = ((((1.0f - tx) * (1.0f - ty)) * (v001 - v000) + (tx * (1.0f - ty)) * (v101 - v100)) +
            ((1.0f - tx) * ty) * (v011 - v010) + (tx * ty) * (v111 - v110)) / h[2];
}

/* ---- constrainVelocity: grid_blocked_boundary_condition_solver3.js:389-477
 *      with a static collider (velocity 0) given by its cell-centred SDF on
 *      the velocity grid and, optionally, its face weights. ---- */
static float clamp_component(float v) {
    if (!isfinite(v)) return 0.0f;
    return v < -100.0f ? -100.0f : (v > 100.0f ? 100.0f : v);
}

void orc_constrain_velocity3(float* u, float* v, float* w, const int n[3], const float h[3], const float origin[3],
                             int closed_flags, const float* sdf, const float* wu, const float* wv, const float* ww,
                             float friction, int depth) {
    orc_constrain_velocity3_inflows(u, v, w, n, h, origin, closed_flags, sdf, wu, wv, ww, friction, depth, 0, NULL,
                                    NULL, NULL);
}

/* The same with inflow objects (createSDFInflow3, sdf_inflow_outflow3.js:15-24;
 * buildInflowKernels, grid_blocked_boundary_condition_solver3.js:130-176):
 * after the domain walls, inflow q in order sets (mode 0) or adds (mode 1)
 * its velocity component on every face whose position lies inside its SDF
 * (trilinear sample < 0, isInsideSdf), then the +-100 clamp.  Each inflow
 * SDF is cell-centred on the velocity grid; inflow_vel holds 3 floats per
 * inflow. */
void orc_constrain_velocity3_inflows(float* u, float* v, float* w, const int n[3], const float h[3],
                                     const float origin[3], int closed_flags, const float* sdf, const float* wu,
                                     const float* wv, const float* ww, float friction, int depth, int n_inflows,
                                     const float* const* inflow_sdf, const float* inflow_vel,
                                     const int* inflow_mode) {
    int sz[3][3] = {{n[0] + 1, n[1], n[2]}, {n[0], n[1] + 1, n[2]}, {n[0], n[1], n[2] + 1}};
    float* data[3] = {u, v, w};
    const float* wts[3] = {wu, wv, ww};
    int64_t cnt[3];
    for (int a = 0; a < 3; ++a) cnt[a] = (int64_t)sz[a][0] * sz[a][1] * sz[a][2];
    if (sdf) {
        const float so[3] = {origin[0] + 0.5f * h[0], origin[1] + 0.5f * h[1], origin[2] + 0.5f * h[2]};
        int32_t* marker[3];
        float* tmp[3];
        for (int a = 0; a < 3; ++a) {
            marker[a] = (int32_t*)malloc(sizeof(int32_t) * cnt[a]);
            tmp[a] = (float*)malloc(sizeof(float) * cnt[a]);
        }
        /* fillMarker + markAndProject */
        for (int a = 0; a < 3; ++a) {
            float fo[3] = {origin[0] + h[0] * 0.5f, origin[1] + h[1] * 0.5f, origin[2] + h[2] * 0.5f};
            fo[a] = origin[a];
            for (int k = 0; k < sz[a][2]; ++k)
                for (int j = 0; j < sz[a][1]; ++j)
                    for (int i = 0; i < sz[a][0]; ++i) {
                        int64_t f = i + (int64_t)sz[a][0] * (j + (int64_t)sz[a][1] * k);
                        float open;
                        if (wts[a]) {
                            open = wts[a][f];
                        } else {
                            float pt[3] = {fo[0] + (float)i * h[0], fo[1] + (float)j * h[1], fo[2] + (float)k * h[2]};
                            float pa[3] = {pt[0], pt[1], pt[2]}, pb[3] = {pt[0], pt[1], pt[2]};
                            pa[a] = pt[a] - h[a] * 0.5f;
                            pb[a] = pt[a] + h[a] * 0.5f;
                            float s0 = orc_sample_trilinear3(sdf, n, so, h, pa);
                            float s1 = orc_sample_trilinear3(sdf, n, so, h, pb);
                            float fr;
                            int in0 = s0 < 0, in1 = s1 < 0;
                            if (in0 && in1) fr = 1.0f;
                            else if (in0) fr = s0 / (s0 - s1);
                            else if (in1) fr = s1 / (s1 - s0);
                            else fr = 0.0f;
                            fr = fr < 0 ? 0 : (fr > 1 ? 1 : fr);
                            open = 1.0f - fr;
                        }
                        if (open > 0) marker[a][f] = 1;
                        else {
                            data[a][f] = 0.0f;  /* static collider velocity */
                            marker[a][f] = 0;
                        }
                    }
        }
        for (int a = 0; a < 3; ++a) orc_extrapolate3(sz[a], data[a], marker[a], data[a], depth);
        /* noFluxProjection into tmp, from the velocity after extrapolation */
        float ou[3], ov[3], ow[3];
        ou[0] = origin[0]; ou[1] = origin[1] + h[1] * 0.5f; ou[2] = origin[2] + h[2] * 0.5f;
        ov[0] = origin[0] + h[0] * 0.5f; ov[1] = origin[1]; ov[2] = origin[2] + h[2] * 0.5f;
        ow[0] = origin[0] + h[0] * 0.5f; ow[1] = origin[1] + h[1] * 0.5f; ow[2] = origin[2];
        const float* oo[3] = {ou, ov, ow};
        for (int a = 0; a < 3; ++a) {
            for (int k = 0; k < sz[a][2]; ++k)
                for (int j = 0; j < sz[a][1]; ++j)
                    for (int i = 0; i < sz[a][0]; ++i) {
                        int64_t f = i + (int64_t)sz[a][0] * (j + (int64_t)sz[a][1] * k);
                        float pt[3] = {oo[a][0] + (float)i * h[0], oo[a][1] + (float)j * h[1],
                                       oo[a][2] + (float)k * h[2]};
                        int blocked = wts[a] ? (wts[a][f] <= 0) : (orc_sample_trilinear3(sdf, n, so, h, pt) < 0);
                        if (!blocked) {
                            tmp[a][f] = data[a][f];
                            continue;
                        }
                        float vel[3];
                        vel[0] = orc_sample_trilinear3(u, sz[0], ou, h, pt);
                        vel[1] = orc_sample_trilinear3(v, sz[1], ov, h, pt);
                        vel[2] = orc_sample_trilinear3(w, sz[2], ow, h, pt);
                        float g[3];
                        sdf_gradient3(sdf, n, so, h, pt, g);
                        float gl = sqrtf(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
                        if (!(gl > 0)) {
                            tmp[a][f] = 0.0f;
                            continue;
                        }
                        float nn[3] = {g[0] / gl, g[1] / gl, g[2] / gl};
                        /* projectAndApplyFriction (…solver3.js:29-45), collider velocity 0 */
                        float dn = vel[0] * nn[0] + vel[1] * nn[1] + vel[2] * nn[2];
                        float vt[3] = {vel[0] - nn[0] * dn, vel[1] - nn[1] * dn, vel[2] - nn[2] * dn};
                        float tl = sqrtf(vt[0] * vt[0] + vt[1] * vt[1] + vt[2] * vt[2]);
                        if (tl > 0) {
                            float veln = fmaxf(-dn, 0.0f);
                            float scale = fmaxf(1.0f - (veln * friction) / tl, 0.0f);
                            vt[0] *= scale; vt[1] *= scale; vt[2] *= scale;
                        }
                        tmp[a][f] = vt[a] + 0.0f;
                    }
        }
        for (int a = 0; a < 3; ++a) memcpy(data[a], tmp[a], sizeof(float) * cnt[a]);
        /* blockedBoundary: collider cells (sdf < 0) with a fluid neighbour set that face to 0 */
        for (int k = 0; k < n[2]; ++k)
            for (int j = 0; j < n[1]; ++j)
                for (int i = 0; i < n[0]; ++i) {
#define S(ii, jj, kk) sdf[(ii) + (int64_t)n[0] * ((jj) + (int64_t)n[1] * (kk))]
                    if (!(S(i, j, k) < 0)) continue;
                    if (i > 0 && !(S(i - 1, j, k) < 0)) u[i + (int64_t)sz[0][0] * (j + (int64_t)sz[0][1] * k)] = 0;
                    if (i < n[0] - 1 && !(S(i + 1, j, k) < 0))
                        u[i + 1 + (int64_t)sz[0][0] * (j + (int64_t)sz[0][1] * k)] = 0;
                    if (j > 0 && !(S(i, j - 1, k) < 0)) v[i + (int64_t)sz[1][0] * (j + (int64_t)sz[1][1] * k)] = 0;
                    if (j < n[1] - 1 && !(S(i, j + 1, k) < 0))
                        v[i + (int64_t)sz[1][0] * (j + 1 + (int64_t)sz[1][1] * k)] = 0;
                    if (k > 0 && !(S(i, j, k - 1) < 0)) w[i + (int64_t)sz[2][0] * (j + (int64_t)sz[2][1] * k)] = 0;
                    if (k < n[2] - 1 && !(S(i, j, k + 1) < 0))
                        w[i + (int64_t)sz[2][0] * (j + (int64_t)sz[2][1] * (k + 1))] = 0;
#undef S
                }
        for (int a = 0; a < 3; ++a) {
            free(marker[a]);
            free(tmp[a]);
        }
    }
    /* domain walls by flag: LEFT 1, RIGHT 2, DOWN 4, UP 8, BACK 16, FRONT 32 */
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j) {
            if (closed_flags & 1) u[0 + (int64_t)sz[0][0] * (j + (int64_t)sz[0][1] * k)] = 0;
            if (closed_flags & 2) u[n[0] + (int64_t)sz[0][0] * (j + (int64_t)sz[0][1] * k)] = 0;
        }
    for (int k = 0; k < n[2]; ++k)
        for (int i = 0; i < n[0]; ++i) {
            if (closed_flags & 4) v[i + (int64_t)sz[1][0] * (0 + (int64_t)sz[1][1] * k)] = 0;
            if (closed_flags & 8) v[i + (int64_t)sz[1][0] * (n[1] + (int64_t)sz[1][1] * k)] = 0;
        }
    for (int j = 0; j < n[1]; ++j)
        for (int i = 0; i < n[0]; ++i) {
            if (closed_flags & 16) w[i + (int64_t)sz[2][0] * (j + (int64_t)sz[2][1] * 0)] = 0;
            if (closed_flags & 32) w[i + (int64_t)sz[2][0] * (j + (int64_t)sz[2][1] * n[2])] = 0;
        }
    for (int q = 0; q < n_inflows; ++q) {
        const float so[3] = {origin[0] + 0.5f * h[0], origin[1] + 0.5f * h[1], origin[2] + 0.5f * h[2]};
        for (int a = 0; a < 3; ++a) {
            float fo[3] = {origin[0] + h[0] * 0.5f, origin[1] + h[1] * 0.5f, origin[2] + h[2] * 0.5f};
            fo[a] = origin[a];
            const float value = inflow_vel[3 * q + a];
            for (int k = 0; k < sz[a][2]; ++k)
                for (int j = 0; j < sz[a][1]; ++j)
                    for (int i = 0; i < sz[a][0]; ++i) {
                        const int64_t f = i + (int64_t)sz[a][0] * (j + (int64_t)sz[a][1] * k);
                        const float pt[3] = {fo[0] + (float)i * h[0], fo[1] + (float)j * h[1],
                                             fo[2] + (float)k * h[2]};
                        if (!(orc_sample_trilinear3(inflow_sdf[q], n, so, h, pt) < 0)) continue;
                        data[a][f] = inflow_mode[q] == 1 ? data[a][f] + value : value;
                    }
        }
    }
    for (int a = 0; a < 3; ++a)
        for (int64_t f = 0; f < cnt[a]; ++f) data[a][f] = clamp_component(data[a][f]);
}

/* ---- the resample pass: grid_flip_solver3.js:640-793.  Counts per cell,
 *      donors (particles in cells holding more than max_pc) ranked by
 *      particle index, needs of under-filled cells with at least 3 of 6
 *      (clamped) neighbours holding min_pc, and the claim of each cell's
 *      slice of the donor pool in cell order; a claimed donor moves to the
 *      cell centre and takes the MAC velocity sampled there. ---- */
void orc_resample3(float* pos, float* vel, int64_t count, const float* u, const float* v, const float* w,
                   const int n[3], const float h[3], const float origin[3], int min_pc, int max_pc) {
    const int64_t cells = (int64_t)n[0] * n[1] * n[2];
    int32_t* cell = (int32_t*)malloc(sizeof(int32_t) * (count > 0 ? count : 1));
    int32_t* cnt = (int32_t*)calloc(cells, sizeof(int32_t));
    int32_t* pool = (int32_t*)malloc(sizeof(int32_t) * (count > 0 ? count : 1));
    int32_t* need = (int32_t*)calloc(cells, sizeof(int32_t));
    orc_cell_index3(pos, count, origin, h, n, cell);
    for (int64_t p = 0; p < count; ++p) cnt[cell[p]]++;
    int64_t donors = 0;
    for (int64_t p = 0; p < count; ++p)
        if (cnt[cell[p]] > max_pc) pool[donors++] = (int32_t)p;
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
#define CNT(ii, jj, kk) cnt[(ii) + (int64_t)n[0] * ((jj) + (int64_t)n[1] * (kk))]
                int nb = (CNT(i > 0 ? i - 1 : 0, j, k) >= min_pc) + (CNT(i < n[0] - 1 ? i + 1 : n[0] - 1, j, k) >= min_pc) +
                         (CNT(i, j > 0 ? j - 1 : 0, k) >= min_pc) + (CNT(i, j < n[1] - 1 ? j + 1 : n[1] - 1, k) >= min_pc) +
                         (CNT(i, j, k > 0 ? k - 1 : 0) >= min_pc) + (CNT(i, j, k < n[2] - 1 ? k + 1 : n[2] - 1) >= min_pc);
                int c0 = CNT(i, j, k);
#undef CNT
                need[i + (int64_t)n[0] * (j + (int64_t)n[1] * k)] = (c0 < min_pc && nb >= 3) ? min_pc - c0 : 0;
            }
    float ou[3], ov[3], ow[3];
    ou[0] = origin[0]; ou[1] = origin[1] + h[1] * 0.5f; ou[2] = origin[2] + h[2] * 0.5f;
    ov[0] = origin[0] + h[0] * 0.5f; ov[1] = origin[1]; ov[2] = origin[2] + h[2] * 0.5f;
    ow[0] = origin[0] + h[0] * 0.5f; ow[1] = origin[1] + h[1] * 0.5f; ow[2] = origin[2];
    int su[3] = {n[0] + 1, n[1], n[2]}, sv[3] = {n[0], n[1] + 1, n[2]}, sw[3] = {n[0], n[1], n[2] + 1};
    const float co[3] = {origin[0] + h[0] * 0.5f, origin[1] + h[1] * 0.5f, origin[2] + h[2] * 0.5f};
    int64_t base = 0;
    for (int64_t c = 0; c < cells; ++c) {
        int needed = need[c];
        if (needed > 0) {
            int i = (int)(c % n[0]), j = (int)((c / n[0]) % n[1]), k = (int)(c / ((int64_t)n[0] * n[1]));
            float np_[3] = {co[0] + (float)i * h[0], co[1] + (float)j * h[1], co[2] + (float)k * h[2]};
            for (int q = 0; q < needed && q < min_pc; ++q) {
                int64_t slot = base + q;
                if (slot >= donors) break;
                int32_t d = pool[slot];
                pos[3 * d + 0] = np_[0];
                pos[3 * d + 1] = np_[1];
                pos[3 * d + 2] = np_[2];
                vel[3 * d + 0] = orc_sample_trilinear3(u, su, ou, h, np_);
                vel[3 * d + 1] = orc_sample_trilinear3(v, sv, ov, h, np_);
                vel[3 * d + 2] = orc_sample_trilinear3(w, sw, ow, h, np_);
            }
        }
        base += needed;
    }
    free(cell); free(cnt); free(pool); free(need);
}

/* ---- createExternalForceSolver3 with the smoke buoyancy force:
 *      external_force_solver3.js:19-47, grid_smoke_solver3.js:66-81 ---- */
void orc_buoyancy3(float* u, float* v, float* w, const float* dens, const float* temp, const int n[3],
                   const float h[3], const float origin[3], float a, float b, float tamb, const float up[3],
                   float dt) {
    float* d[3] = {u, v, w};
    const float co[3] = {origin[0] + h[0] * 0.5f, origin[1] + h[1] * 0.5f, origin[2] + h[2] * 0.5f};
    for (int ax = 0; ax < 3; ++ax) {
        int s[3] = {n[0], n[1], n[2]};
        s[ax] += 1;
        float fo[3] = {origin[0] + h[0] * 0.5f, origin[1] + h[1] * 0.5f, origin[2] + h[2] * 0.5f};
        fo[ax] = origin[ax];
        for (int k = 0; k < s[2]; ++k)
            for (int j = 0; j < s[1]; ++j)
                for (int i = 0; i < s[0]; ++i) {
                    float pos[3] = {fo[0] + (float)i * h[0], fo[1] + (float)j * h[1], fo[2] + (float)k * h[2]};
                    float rho = orc_sample_trilinear3(dens, n, co, h, pos);
                    float t = orc_sample_trilinear3(temp, n, co, h, pos);
                    float fb = a * rho + b * (t - tamb);
                    int64_t f = i + (int64_t)s[0] * (j + (int64_t)s[1] * k);
                    d[ax][f] = d[ax][f] + (up[ax] * fb) * dt;
                }
    }
}

/* ---- the ball source of examples/34-smoke-plume-3d/main.js:146-158 ---- */
void orc_source_sphere3(float* d, const int n[3], const float h[3], const float origin[3], const float c[3], float r,
                        float value) {
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                float px = (origin[0] + h[0] * 0.5f) + (float)i * h[0];
                float py = (origin[1] + h[1] * 0.5f) + (float)j * h[1];
                float pz = (origin[2] + h[2] * 0.5f) + (float)k * h[2];
                float dx = px - c[0], dy = py - c[1], dz = pz - c[2];
                int in = (dx * dx + dy * dy) + dz * dz < r * r;
                int64_t q = i + (int64_t)n[0] * (j + (int64_t)n[1] * k);
                float s = in ? value : 0.0f;
                d[q] = d[q] > s ? d[q] : s;
            }
}

/* ---- phase-aware resampling (grid_two_phase_flip_solver2.js:1131-1503, in
 *      3-D): over/under detection on the total count, donors split by phase
 *      (concentration > 0.5) into a liquid pool (bottom half, ranked by
 *      particle index) and a gas pool (top half, ranked downward), each capped
 *      at max(1, count / 2); an under-filled cell with at least 3 of its 6
 *      clamped neighbours holding min_pc takes donors of its own majority
 *      phase (its neighbours' when it is empty).  cnt / accum are the
 *      per-cell count and fixed-point concentration sum at `scale`. ---- */
void orc_resample_phase3(float* pos, float* vel, const float* conc, int64_t count, const float* u, const float* v,
                         const float* w, const int n[3], const float h[3], const float origin[3], int min_pc,
                         int max_pc, const int32_t* cnt, const int32_t* accum, float scale) {
    const int64_t cells = (int64_t)n[0] * n[1] * n[2];
    int32_t* cell = (int32_t*)malloc(sizeof(int32_t) * (count > 0 ? count : 1));
    int32_t* pool = (int32_t*)malloc(sizeof(int32_t) * (count > 0 ? count : 1));
    int32_t* need = (int32_t*)calloc(cells, sizeof(int32_t));
    orc_cell_index3(pos, count, origin, h, n, cell);
    const int64_t half = count / 2 > 1 ? count / 2 : 1;
    const int64_t top = count - 1;
    int64_t nl = 0, ng = 0;
    for (int64_t p = 0; p < count; ++p) {
        if (cnt[cell[p]] <= max_pc) continue;
        if (conc[p] > 0.5f) {
            if (nl < half) pool[nl] = (int32_t)p;
            ++nl;
        } else {
            if (ng < half) pool[top - ng] = (int32_t)p;
            ++ng;
        }
    }
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                const int64_t nbr[6] = {orc_cidx(n, i > 0 ? i - 1 : 0, j, k), orc_cidx(n, i < n[0] - 1 ? i + 1 : n[0] - 1, j, k),
                                        orc_cidx(n, i, j > 0 ? j - 1 : 0, k), orc_cidx(n, i, j < n[1] - 1 ? j + 1 : n[1] - 1, k),
                                        orc_cidx(n, i, j, k > 0 ? k - 1 : 0), orc_cidx(n, i, j, k < n[2] - 1 ? k + 1 : n[2] - 1)};
                int well = 0;
                int32_t ntot = 0, nsum = 0;
                for (int q = 0; q < 6; ++q) {
                    well += cnt[nbr[q]] >= min_pc;
                    ntot = (int32_t)((uint32_t)ntot + (uint32_t)cnt[nbr[q]]);
                    nsum = (int32_t)((uint32_t)nsum + (uint32_t)accum[nbr[q]]);
                }
                const int64_t c = orc_cidx(n, i, j, k);
                const float total = (float)cnt[c];
                const float own = ((float)accum[c] / scale) / (total > 1.0f ? total : 1.0f);
                const float ntf = (float)ntot;
                const float nmean = ((float)nsum / scale) / (ntf > 1.0f ? ntf : 1.0f);
                const int liquid = total > 0.5f ? own > 0.5f : nmean > 0.5f;
                const int eligible = cnt[c] < min_pc && well >= 3;
                const int needed = min_pc - cnt[c];
                need[c] = eligible ? (liquid ? needed : -needed) : 0;
            }
    float ou[3], ov[3], ow[3];
    ou[0] = origin[0]; ou[1] = origin[1] + h[1] * 0.5f; ou[2] = origin[2] + h[2] * 0.5f;
    ov[0] = origin[0] + h[0] * 0.5f; ov[1] = origin[1]; ov[2] = origin[2] + h[2] * 0.5f;
    ow[0] = origin[0] + h[0] * 0.5f; ow[1] = origin[1] + h[1] * 0.5f; ow[2] = origin[2];
    int su[3] = {n[0] + 1, n[1], n[2]}, sv[3] = {n[0], n[1] + 1, n[2]}, sw[3] = {n[0], n[1], n[2] + 1};
    const float co[3] = {origin[0] + h[0] * 0.5f, origin[1] + h[1] * 0.5f, origin[2] + h[2] * 0.5f};
    const int64_t avail_l = nl < half ? nl : half, avail_g = ng < half ? ng : half;
    int64_t base_l = 0, base_g = 0;
    for (int64_t c = 0; c < cells; ++c) {
        const int nd = need[c];
        if (nd == 0) continue;
        const int liquid = nd > 0;
        const int needed = liquid ? nd : -nd;
        const int64_t base = liquid ? base_l : base_g;
        const int64_t avail = liquid ? avail_l : avail_g;
        const int i = (int)(c % n[0]), j = (int)((c / n[0]) % n[1]), k = (int)(c / ((int64_t)n[0] * n[1]));
        const float np_[3] = {co[0] + (float)i * h[0], co[1] + (float)j * h[1], co[2] + (float)k * h[2]};
        for (int q = 0; q < needed && q < min_pc; ++q) {
            const int64_t slot = base + q;
            if (slot >= avail) break;
            const int32_t d = pool[liquid ? slot : top - slot];
            pos[3 * d + 0] = np_[0];
            pos[3 * d + 1] = np_[1];
            pos[3 * d + 2] = np_[2];
            vel[3 * d + 0] = orc_sample_trilinear3(u, su, ou, h, np_);
            vel[3 * d + 1] = orc_sample_trilinear3(v, sv, ov, h, np_);
            vel[3 * d + 2] = orc_sample_trilinear3(w, sw, ow, h, np_);
        }
        if (liquid) base_l += needed;
        else base_g += needed;
    }
    free(cell); free(pool); free(need);
}

/* ---- createOutflowPressureDirichlet3 (sdf_inflow_outflow3.js:44-62) for one
 *      outflow whose SDF is cell-centred on the pressure grid: active where
 *      the SDF sampled at the cell centre (dataPosition) is < 0 ---- */
void orc_outflow_dirichlet3(const float* osdf, const int n[3], const float h[3], const float origin[3],
                            uint8_t* active) {
    const float so[3] = {origin[0] + h[0] * 0.5f, origin[1] + h[1] * 0.5f, origin[2] + h[2] * 0.5f};
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                const float p[3] = {so[0] + h[0] * (float)i, so[1] + h[1] * (float)j, so[2] + h[2] * (float)k};
                active[i + (int64_t)n[0] * (j + (int64_t)n[1] * k)] = orc_sample_trilinear3(osdf, n, so, h, p) < 0;
            }
}

/* ---- pushOutOfCollider (grid_flip_solver3.js:945-964): a particle where
 *      the collider's SDF (cell-centred on the grid) is below thresh moves
 *      along the SDF's normalised trilinear gradient by (thresh - sdf) +
 *      shift ---- */
void orc_push_out3(float* pos, int64_t count, const float* sdf, const int n[3], const float h[3],
                   const float origin[3], float thresh, float shift) {
    const float so[3] = {origin[0] + 0.5f * h[0], origin[1] + 0.5f * h[1], origin[2] + 0.5f * h[2]};
    for (int64_t p = 0; p < count; ++p) {
        const float pp[3] = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
        const float v = orc_sample_trilinear3(sdf, n, so, h, pp);
        if (!(v < thresh)) continue;
        float g[3];
        sdf_gradient3(sdf, n, so, h, pp, g);
        const float len = sqrtf((g[0] * g[0] + g[1] * g[1]) + g[2] * g[2]);
        if (!(len > 0.0f)) continue;
        const float dist = (thresh - v) + shift;
        for (int a = 0; a < 3; ++a) pos[3 * p + a] = pp[a] + (g[a] / len) * dist;
    }
}
