/*
 * Oracle for the particle/grid transfers and advection (TEST INFRASTRUCTURE
 * ONLY).  f32 arithmetic in the reference's operation order, built with
 * -ffp-contract=off so that no multiply-add is fused behind our back.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "oracle_util.h"

void orc_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

int orc_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---- mulberry32: packages/fluxflow/src/grid/jitter_random.js:29-42 ---- */
void orc_rng_seed(orc_rng* r, uint32_t seed) { r->state = seed; }

uint32_t orc_rng_next_u32(orc_rng* r) {
    r->state += 0x6D2B79F5u;
    uint32_t t = r->state;
    t = (t ^ (t >> 15)) * (t | 1u);
    t ^= t + (t ^ (t >> 7)) * (t | 61u);
    return t ^ (t >> 14);
}

double orc_rng_next(orc_rng* r) { return (double)orc_rng_next_u32(r) / 4294967296.0; }

void orc_rng_fill(uint32_t seed, double* out, int64_t n) {
    orc_rng r;
    orc_rng_seed(&r, seed);
    for (int64_t q = 0; q < n; ++q) out[q] = orc_rng_next(&r);
}

/* ---- computeFlipBoxSeed3: grid_flip_solver3.js:70-130 ---- */
static int js_round_pos(double v) { return (int)floor(v + 0.5); } /* Math.round */

int64_t orc_flip_box_seed3(const double bmin[3], const double bmax[3], const double h[3],
                           int ppa, double jitter, uint32_t seed, float* out) {
    int cells[3];
    double sub[3];
    for (int a = 0; a < 3; ++a) {
        int c = js_round_pos((bmax[a] - bmin[a]) / h[a]);
        cells[a] = c > 1 ? c : 1;
        sub[a] = h[a] / ppa;
    }
    int64_t count = (int64_t)cells[0] * cells[1] * cells[2] * ppa * ppa * ppa;
    if (!out) return count;
    orc_rng r;
    orc_rng_seed(&r, seed);
    int64_t q = 0;
    for (int ck = 0; ck < cells[2]; ++ck)
        for (int cj = 0; cj < cells[1]; ++cj)
            for (int ci = 0; ci < cells[0]; ++ci)
                for (int sk = 0; sk < ppa; ++sk)
                    for (int sj = 0; sj < ppa; ++sj)
                        for (int si = 0; si < ppa; ++si) {
                            double jx = (orc_rng_next(&r) * 2 - 1) * jitter * sub[0];
                            double jy = (orc_rng_next(&r) * 2 - 1) * jitter * sub[1];
                            double jz = (orc_rng_next(&r) * 2 - 1) * jitter * sub[2];
                            double x = bmin[0] + ci * h[0] + (si + 0.5) * sub[0] + jx;
                            double y = bmin[1] + cj * h[1] + (sj + 0.5) * sub[1] + jy;
                            double z = bmin[2] + ck * h[2] + (sk + 0.5) * sub[2] + jz;
                            out[3 * q + 0] = (float)x;
                            out[3 * q + 1] = (float)y;
                            out[3 * q + 2] = (float)z;
                            ++q;
                        }
    return count;
}

/* ---- cellIndexOf: grid_flip_solver3.js:227-235 ---- */
static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

void orc_cell_index3(const float* pos, int64_t count, const float origin[3], const float h[3],
                     const int n[3], int32_t* cell_out) {
    for (int64_t p = 0; p < count; ++p) {
        int c[3];
        for (int a = 0; a < 3; ++a) {
            float g = (pos[3 * p + a] - origin[a]) / h[a];
            c[a] = clampi((int)floorf(g), 0, n[a] - 1);
        }
        cell_out[p] = (int32_t)orc_cidx(n, c[0], c[1], c[2]);
    }
}

/* stable counting sort: the permutation a stable LSD radix sort produces */
void orc_sort_by_cell(const int32_t* cell, int64_t count, int64_t n_cells, int32_t* perm,
                      int32_t* cell_start) {
    int64_t* cnt = (int64_t*)calloc(n_cells + 1, sizeof(int64_t));
    for (int64_t p = 0; p < count; ++p) cnt[cell[p] + 1]++;
    for (int64_t c = 0; c < n_cells; ++c) cnt[c + 1] += cnt[c];
    for (int64_t c = 0; c <= n_cells; ++c) cell_start[c] = (int32_t)cnt[c];
    for (int64_t p = 0; p < count; ++p) perm[cnt[cell[p]]++] = (int32_t)p;
    free(cnt);
}

/* ---- trilinearCoordsAndWeights3: grid_math3.js:61-95 ---- */
typedef struct {
    int i0c, j0c, k0c, i1c, j1c, k1c;
    float w[8]; /* 000,100,010,110,001,101,011,111 */
} orc_tri;

static inline void tri_weights(const float pos[3], const float o[3], const float h[3],
                               const int shape[3], orc_tri* t) {
    float g[3], f[3];
    int i0[3];
    for (int a = 0; a < 3; ++a) {
        g[a] = (pos[a] - o[a]) / h[a];
        i0[a] = (int)floorf(g[a]);
        f[a] = g[a] - (float)i0[a];
    }
    t->i0c = clampi(i0[0], 0, shape[0] - 1);
    t->i1c = clampi(i0[0] + 1, 0, shape[0] - 1);
    t->j0c = clampi(i0[1], 0, shape[1] - 1);
    t->j1c = clampi(i0[1] + 1, 0, shape[1] - 1);
    t->k0c = clampi(i0[2], 0, shape[2] - 1);
    t->k1c = clampi(i0[2] + 1, 0, shape[2] - 1);
    float gx = 1.0f - f[0], gy = 1.0f - f[1], gz = 1.0f - f[2];
    float fx = f[0], fy = f[1], fz = f[2];
    t->w[0] = (gx * gy) * gz;
    t->w[1] = (fx * gy) * gz;
    t->w[2] = (gx * fy) * gz;
    t->w[3] = (fx * fy) * gz;
    t->w[4] = (gx * gy) * fz;
    t->w[5] = (fx * gy) * fz;
    t->w[6] = (gx * fy) * fz;
    t->w[7] = (fx * fy) * fz;
}

static inline void tri_corner(const orc_tri* t, int c, int* i, int* j, int* k) {
    *i = (c & 1) ? t->i1c : t->i0c;
    *j = (c & 2) ? t->j1c : t->j0c;
    *k = (c & 4) ? t->k1c : t->k0c;
}

/* ---- P2G fixed-point: buildScatter / buildFinalize 508-564 ---- */
static inline int32_t f32_to_i32_sat(float v) {
    if (!(v == v)) return 0;
    if (v >= 2147483520.0f) return 2147483647;
    if (v <= -2147483648.0f) return (int32_t)0x80000000u;
    return (int32_t)v;
}

void orc_p2g_fixed3(const float* pos, const float* vel, int64_t count, int comp,
                    const int face_n[3], const float data_origin[3], const float h[3], float scale,
                    int32_t* numer, int32_t* denom) {
    for (int64_t p = 0; p < count; ++p) {
        orc_tri t;
        tri_weights(&pos[3 * p], data_origin, h, face_n, &t);
        float value = vel[3 * p + comp];
        for (int c = 0; c < 8; ++c) {
            int i, j, k;
            tri_corner(&t, c, &i, &j, &k);
            int64_t f = orc_cidx(face_n, i, j, k);
            float w = t.w[c];
            /* round() in WGSL is round-half-to-even: rintf in the default mode */
            uint32_t nu = (uint32_t)numer[f] + (uint32_t)f32_to_i32_sat(rintf((value * w) * scale));
            uint32_t de = (uint32_t)denom[f] + (uint32_t)f32_to_i32_sat(rintf(w * scale));
            numer[f] = (int32_t)nu; /* atomicAdd wraps */
            denom[f] = (int32_t)de;
        }
    }
}

void orc_p2g_finalize_fixed(const int32_t* numer, const int32_t* denom, int64_t n_faces, float scale,
                            float eps, float* out, int32_t* valid) {
    for (int64_t f = 0; f < n_faces; ++f) {
        float nu = (float)numer[f] / scale;
        float de = (float)denom[f] / scale;
        if (de > eps) {
            out[f] = nu / de;
            valid[f] = 1;
        } else {
            out[f] = 0;
            valid[f] = 0;
        }
    }
}

/* ---- P2G as a deterministic gather over cell-sorted particles ----
 * For each face the contributing cells are visited in (k, j, i) order over
 * the 2x3x3 neighbourhood that can reach it, the particles of a cell in
 * sorted order, and each particle's eight corners in the reference's order;
 * every corner that lands on this face adds v*w and w in f32.  The HIP gather
 * kernel uses exactly this order, so the two agree bit for bit. */
void orc_p2g_gather3(const float* pos, const float* vel, const int32_t* perm,
                     const int32_t* cell_start, int comp, const int n[3],
                     const float data_origin[3], const float h[3], float eps,
                     float* out, int32_t* valid) {
    int fn[3] = {n[0] + (comp == 0), n[1] + (comp == 1), n[2] + (comp == 2)};
#pragma omp parallel for schedule(static)
    for (int fk = 0; fk < fn[2]; ++fk)
        for (int fj = 0; fj < fn[1]; ++fj)
            for (int fi = 0; fi < fn[0]; ++fi) {
                int F[3] = {fi, fj, fk};
                int lo[3], hi[3];
                for (int a = 0; a < 3; ++a) {
                    lo[a] = (a == comp) ? F[a] - 1 : F[a] - 1;
                    hi[a] = (a == comp) ? F[a] : F[a] + 1;
                    if (lo[a] < 0) lo[a] = 0;
                    if (hi[a] > n[a] - 1) hi[a] = n[a] - 1;
                }
                float nu = 0.0f, de = 0.0f;
                for (int ck = lo[2]; ck <= hi[2]; ++ck)
                    for (int cj = lo[1]; cj <= hi[1]; ++cj)
                        for (int ci = lo[0]; ci <= hi[0]; ++ci) {
                            int64_t c = orc_cidx(n, ci, cj, ck);
                            for (int32_t s = cell_start[c]; s < cell_start[c + 1]; ++s) {
                                int32_t p = perm[s];
                                orc_tri t;
                                tri_weights(&pos[3 * p], data_origin, h, fn, &t);
                                float value = vel[3 * p + comp];
                                for (int q = 0; q < 8; ++q) {
                                    int i, j, k;
                                    tri_corner(&t, q, &i, &j, &k);
                                    if (i == fi && j == fj && k == fk) {
                                        nu = nu + value * t.w[q];
                                        de = de + t.w[q];
                                    }
                                }
                            }
                        }
                int64_t f = orc_cidx(fn, fi, fj, fk);
                if (de > eps) {
                    out[f] = nu / de;
                    valid[f] = 1;
                } else {
                    out[f] = 0;
                    valid[f] = 0;
                }
            }
}

void orc_p2g_float3(const float* pos, const float* vel, int64_t count, int comp,
                    const int face_n[3], const float data_origin[3], const float h[3],
                    double* numer, double* denom) {
    for (int64_t p = 0; p < count; ++p) {
        orc_tri t;
        tri_weights(&pos[3 * p], data_origin, h, face_n, &t);
        double value = vel[3 * p + comp];
        for (int c = 0; c < 8; ++c) {
            int i, j, k;
            tri_corner(&t, c, &i, &j, &k);
            int64_t f = orc_cidx(face_n, i, j, k);
            numer[f] += value * (double)t.w[c];
            denom[f] += (double)t.w[c];
        }
    }
}

void orc_p2g_finalize_float(const double* numer, const double* denom, int64_t n_faces, float eps,
                            float* out, int32_t* valid) {
    for (int64_t f = 0; f < n_faces; ++f) {
        if (denom[f] > eps) {
            out[f] = (float)(numer[f] / denom[f]);
            valid[f] = 1;
        } else {
            out[f] = 0;
            valid[f] = 0;
        }
    }
}

/* ---- collocatedValueAtPosition3: grid_math3.js:99-113 ---- */
float orc_sample_trilinear3(const float* data, const int shape[3], const float o[3], const float h[3],
                            const float pos[3]) {
    orc_tri t;
    tri_weights(pos, o, h, shape, &t);
    float s = 0;
    for (int c = 0; c < 8; ++c) {
        int i, j, k;
        tri_corner(&t, c, &i, &j, &k);
        float term = data[orc_cidx(shape, i, j, k)] * t.w[c];
        s = (c == 0) ? term : s + term;
    }
    return s;
}

static void mac_origins(const float origin[3], const float h[3], float ou[3], float ov[3], float ow[3]) {
    /* createFaceCenteredGrid3: grid_data3.js:128-130 */
    ou[0] = origin[0];             ou[1] = origin[1] + h[1] * 0.5f; ou[2] = origin[2] + h[2] * 0.5f;
    ov[0] = origin[0] + h[0] * 0.5f; ov[1] = origin[1];             ov[2] = origin[2] + h[2] * 0.5f;
    ow[0] = origin[0] + h[0] * 0.5f; ow[1] = origin[1] + h[1] * 0.5f; ow[2] = origin[2];
}

static void sample_mac(const float* u, const float* v, const float* w, const int n[3], const float h[3],
                       const float ou[3], const float ov[3], const float ow[3], const float pos[3],
                       float out[3]) {
    int su[3] = {n[0] + 1, n[1], n[2]}, sv[3] = {n[0], n[1] + 1, n[2]}, sw[3] = {n[0], n[1], n[2] + 1};
    out[0] = orc_sample_trilinear3(u, su, ou, h, pos);
    out[1] = orc_sample_trilinear3(v, sv, ov, h, pos);
    out[2] = orc_sample_trilinear3(w, sw, ow, h, pos);
}

/* ---- G2P: grid_flip_solver3.js:903-922 ---- */
void orc_g2p3(const float* pos, float* vel, int64_t count, const float* u, const float* v,
              const float* w, const float* uo, const float* vo, const float* wo, const int n[3],
              const float h[3], const float origin[3], float flip_ratio, float damping) {
    float ou[3], ov[3], ow[3];
    mac_origins(origin, h, ou, ov, ow);
    float dmp = damping < 0 ? 0 : (damping > 1 ? 1 : damping);
    float one_r = 1.0f - flip_ratio, one_d = 1.0f - dmp;
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < count; ++p) {
        float nv[3], ov3[3];
        sample_mac(u, v, w, n, h, ou, ov, ow, &pos[3 * p], nv);
        sample_mac(uo, vo, wo, n, h, ou, ov, ow, &pos[3 * p], ov3);
        float res[3];
        int bad = 0;
        for (int a = 0; a < 3; ++a) {
            float delta = nv[a] - ov3[a];
            float flip = vel[3 * p + a] + delta;
            float blended = flip * flip_ratio + nv[a] * one_r;
            res[a] = blended * one_d;
            if (!isfinite(res[a])) bad = 1;
        }
        for (int a = 0; a < 3; ++a) {
            float r = res[a];
            r = r < -500.0f ? -500.0f : (r > 500.0f ? 500.0f : r);
            vel[3 * p + a] = bad ? 0.0f : r;
        }
    }
}

/* ---- a static collider's SDF, cell-centred on the velocity grid
 *      (createSDFStaticCollider3's sample, sdf_collider3.js:33-37:
 *      collocatedValueAtPosition3); no collider samples as 1
 *      (advection_solver3.js:78-82) ---- */
static float collider_sample(const float* csdf, const int n[3], const float h[3], const float so[3],
                             const float p[3]) {
    return csdf ? orc_sample_trilinear3(csdf, n, so, h, p) : 1.0f;
}

/* ---- backTrace: advection_solver3.js:20-62.  With a collider a sub-step
 *      whose end point lies in it (phi1 <= 0) stops at the SDF-weighted
 *      blend of its two ends, w = |phi1| / (|phi0| + |phi1|), and ends the
 *      trace. ---- */
static void back_trace_c(const float* u, const float* v, const float* w, const int n[3], const float h[3],
                         const float ou[3], const float ov[3], const float ow[3], const float* csdf,
                         const float start[3], float dt, float hmin, int max_sub, float direction, float out[3]) {
    const float so[3] = {ov[0], ou[1], ou[2]};
    float pt[3] = {start[0], start[1], start[2]};
    float rem = dt;
    for (int s = 0; s < max_sub; ++s) {
        if (rem <= 1e-6f) break;
        float v0[3];
        sample_mac(u, v, w, n, h, ou, ov, ow, pt, v0);
        float len = sqrtf((v0[0] * v0[0] + v0[1] * v0[1]) + v0[2] * v0[2]);
        float ns = ceilf((len * rem) / hmin);
        if (ns < 1.0f) ns = 1.0f;
        float sub = rem / ns;
        float half = sub * 0.5f;
        float mid[3];
        for (int a = 0; a < 3; ++a) mid[a] = pt[a] - (v0[a] * half) * direction;
        float mv[3];
        sample_mac(u, v, w, n, h, ou, ov, ow, mid, mv);
        float next[3];
        for (int a = 0; a < 3; ++a) next[a] = pt[a] - (mv[a] * sub) * direction;
        if (csdf) {
            const float phi0 = collider_sample(csdf, n, h, so, pt);
            const float phi1 = collider_sample(csdf, n, h, so, next);
            if (phi1 <= 0.0f) {
                const float wt = fabsf(phi1) / (fabsf(phi0) + fabsf(phi1));
                for (int a = 0; a < 3; ++a) pt[a] = pt[a] * wt + next[a] * (1.0f - wt);
                break;
            }
        }
        for (int a = 0; a < 3; ++a) pt[a] = next[a];
        rem = rem - sub;
    }
    out[0] = pt[0];
    out[1] = pt[1];
    out[2] = pt[2];
}


/* ---- advectParticles: grid_flip_solver3.js:933-940 ---- */
void orc_advect_particles3(float* pos, int64_t count, const float* u, const float* v, const float* w,
                           const int n[3], const float h[3], const float origin[3], float dt,
                           int max_sub) {
    orc_advect_particles3_c(pos, count, u, v, w, n, h, origin, NULL, dt, max_sub);
}

/* the same through the collider (its SDF cell-centred on the grid, or NULL) */
void orc_advect_particles3_c(float* pos, int64_t count, const float* u, const float* v, const float* w,
                             const int n[3], const float h[3], const float origin[3], const float* csdf, float dt,
                             int max_sub) {
    float ou[3], ov[3], ow[3];
    mac_origins(origin, h, ou, ov, ow);
    float hmin = fminf(h[0], fminf(h[1], h[2]));
    float lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = origin[a] + 1e-4f;
        hi[a] = (origin[a] + (float)n[a] * h[a]) - 1e-4f;
    }
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < count; ++p) {
        float tr[3];
        back_trace_c(u, v, w, n, h, ou, ov, ow, csdf, &pos[3 * p], dt, hmin, max_sub, -1.0f, tr);
        int bad = !isfinite(tr[0]) || !isfinite(tr[1]) || !isfinite(tr[2]);
        for (int a = 0; a < 3; ++a) {
            float s = bad ? pos[3 * p + a] : tr[a];
            pos[3 * p + a] = s < lo[a] ? lo[a] : (s > hi[a] ? hi[a] : s);
        }
    }
}

/* ---- monotonicCubic1d3 / collocatedCubicValueAtPosition3: grid_math3.js:137-207 ---- */
static inline float mono_cubic(float f0, float f1, float f2, float f3, float f) {
    float D1 = f2 - f1;
    float rawD1 = (f2 - f0) * 0.5f;
    float rawD2 = (f3 - f1) * 0.5f;
    int flat = fabsf(D1) < 1e-12f;
    float d1 = (flat || rawD1 * D1 < 0) ? 0.0f : rawD1;
    float d2 = (flat || rawD2 * D1 < 0) ? 0.0f : rawD2;
    float a3 = (d1 + d2) - D1 * 2.0f;
    float a2 = (D1 * 3.0f - d1 * 2.0f) - d2;
    float a1 = d1, a0 = f1;
    return (((a3 * f) * f) * f + (a2 * f) * f) + a1 * f + a0;
}

static inline void cubic_idx(float coord, int nn, int idx[4], float* f) {
    int i0 = (int)floorf(coord);
    *f = coord - (float)i0;
    idx[0] = clampi(i0 - 1, 0, nn - 1);
    idx[1] = clampi(i0, 0, nn - 1);
    idx[2] = clampi(i0 + 1, 0, nn - 1);
    idx[3] = clampi(i0 + 2, 0, nn - 1);
}

static float sample_cubic(const float* data, const int s[3], const float o[3], const float h[3],
                          const float pos[3]) {
    int xi[4], yi[4], zi[4];
    float fx, fy, fz;
    cubic_idx((pos[0] - o[0]) / h[0], s[0], xi, &fx);
    cubic_idx((pos[1] - o[1]) / h[1], s[1], yi, &fy);
    cubic_idx((pos[2] - o[2]) / h[2], s[2], zi, &fz);
    float zv[4];
    for (int c = 0; c < 4; ++c) {
        float rv[4];
        for (int b = 0; b < 4; ++b) {
            int j = yi[b], k = zi[c];
            rv[b] = mono_cubic(data[orc_cidx(s, xi[0], j, k)], data[orc_cidx(s, xi[1], j, k)],
                               data[orc_cidx(s, xi[2], j, k)], data[orc_cidx(s, xi[3], j, k)], fx);
        }
        zv[c] = mono_cubic(rv[0], rv[1], rv[2], rv[3], fy);
    }
    return mono_cubic(zv[0], zv[1], zv[2], zv[3], fz);
}

/* ---- advectFaceCentered3 (advection_solver3.js:130-282) through the
 *      collider (csdf: its SDF cell-centred on the grid, or NULL).  The trace
 *      follows the solver's velocity grid (vel); the face value is the
 *      tricubic sample of the advected grid (data) at the trace.  A face
 *      whose position lies in the collider (sample <= 0) keeps out's value.
 *      Order 2 (MacCormack): fwd = order 1 (the input where the face lies in
 *      the collider); bwd = fwd sampled at the trace with direction -1;
 *      out = fwd + (in - bwd) / 2 where that lies within the min/max of the
 *      input over the 8 trilinear corners of the forward trace whose own
 *      face positions lie outside the collider (gatherLocalMinMax,
 *      clampCorrection 95-128), else fwd. ---- */
static float mc_clamp_c(const float* data, const int s[3], const float o[3], const float h[3], const float tr[3],
                        float corrected, float fwd, const float* csdf, const int n[3], const float so[3]) {
    orc_tri t;
    tri_weights(tr, o, h, s, &t);
    float mn = 1e30f, mx = -1e30f;
    int have = 0;
    for (int c = 0; c < 8; ++c) {
        int i, j, k;
        tri_corner(&t, c, &i, &j, &k);
        const float cp[3] = {o[0] + (float)i * h[0], o[1] + (float)j * h[1], o[2] + (float)k * h[2]};
        if (!(collider_sample(csdf, n, h, so, cp) > 0.0f)) continue;
        const float v = data[orc_cidx(s, i, j, k)];
        mn = fminf(mn, v);
        mx = fmaxf(mx, v);
        have = 1;
    }
    return (have && corrected >= mn && corrected <= mx) ? corrected : fwd;
}

void orc_advect_faces3_c(const float* const vel[3], const float* const data[3], float* const out[3], const int n[3],
                         const float h[3], const float origin[3], const float* csdf, float dt, int max_sub,
                         int order) {
    float ou[3], ov[3], ow[3];
    mac_origins(origin, h, ou, ov, ow);
    const float hmin = fminf(h[0], fminf(h[1], h[2]));
    const float so[3] = {ov[0], ou[1], ou[2]};
    const float* org[3] = {ou, ov, ow};
    float* fwd[3] = {NULL, NULL, NULL};
    for (int a = 0; a < 3; ++a) {
        int s[3] = {n[0] + (a == 0), n[1] + (a == 1), n[2] + (a == 2)};
        float* dst = out[a];
        if (order == 2) dst = fwd[a] = (float*)malloc(sizeof(float) * (size_t)orc_faces(n, a));
#pragma omp parallel for schedule(static)
        for (int k = 0; k < s[2]; ++k)
            for (int j = 0; j < s[1]; ++j)
                for (int i = 0; i < s[0]; ++i) {
                    float pos[3] = {org[a][0] + (float)i * h[0], org[a][1] + (float)j * h[1],
                                    org[a][2] + (float)k * h[2]};
                    const int64_t f = orc_cidx(s, i, j, k);
                    if (!(collider_sample(csdf, n, h, so, pos) > 0.0f)) {
                        if (order == 2) dst[f] = data[a][f];
                        continue;
                    }
                    float tr[3];
                    back_trace_c(vel[0], vel[1], vel[2], n, h, ou, ov, ow, csdf, pos, dt, hmin, max_sub, 1.0f, tr);
                    dst[f] = sample_cubic(data[a], s, org[a], h, tr);
                }
    }
    if (order != 2) return;
    for (int a = 0; a < 3; ++a) {
        int s[3] = {n[0] + (a == 0), n[1] + (a == 1), n[2] + (a == 2)};
#pragma omp parallel for schedule(static)
        for (int k = 0; k < s[2]; ++k)
            for (int j = 0; j < s[1]; ++j)
                for (int i = 0; i < s[0]; ++i) {
                    const float pos[3] = {org[a][0] + (float)i * h[0], org[a][1] + (float)j * h[1],
                                          org[a][2] + (float)k * h[2]};
                    if (!(collider_sample(csdf, n, h, so, pos) > 0.0f)) continue;
                    float tb[3], tf[3];
                    back_trace_c(vel[0], vel[1], vel[2], n, h, ou, ov, ow, csdf, pos, dt, hmin, max_sub, -1.0f, tb);
                    const float bwd = sample_cubic(fwd[a], s, org[a], h, tb);
                    back_trace_c(vel[0], vel[1], vel[2], n, h, ou, ov, ow, csdf, pos, dt, hmin, max_sub, 1.0f, tf);
                    const int64_t f = orc_cidx(s, i, j, k);
                    const float fv = fwd[a][f];
                    const float corrected = fv + (data[a][f] - bwd) * 0.5f;
                    out[a][f] = mc_clamp_c(data[a], s, org[a], h, tf, corrected, fv, csdf, n, so);
                }
    }
    for (int a = 0; a < 3; ++a) free(fwd[a]);
}

void orc_advect_faces3_vd(const float* const vel[3], const float* const data[3], float* const out[3], const int n[3],
                          const float h[3], const float origin[3], float dt, int max_sub) {
    orc_advect_faces3_c(vel, data, out, n, h, origin, NULL, dt, max_sub, 1);
}

void orc_advect_face3(const float* u, const float* v, const float* w, float* uo_, float* vo_, float* wo_,
                      const int n[3], const float h[3], const float origin[3], float dt, int max_sub) {
    const float* vel[3] = {u, v, w};
    float* out[3] = {uo_, vo_, wo_};
    orc_advect_faces3_vd(vel, vel, out, n, h, origin, dt, max_sub);
}

/* ---- two-phase face beta (3-D form of grid_two_phase_flip_solver2.js:693-762) ---- */
void orc_two_phase_beta3(const float* frac, const int n[3], float rho_gas, float rho_liq, float* bu,
                         float* bv, float* bw) {
    const float min_d = 1e-4f;
    float r0 = rho_gas > min_d ? rho_gas : min_d;
    float r1 = rho_liq > min_d ? rho_liq : min_d;
    float ref = r0 > r1 ? r0 : r1;
    float* out[3] = {bu, bv, bw};
    for (int a = 0; a < 3; ++a) {
        int s[3] = {n[0] + (a == 0), n[1] + (a == 1), n[2] + (a == 2)};
        for (int k = 0; k < s[2]; ++k)
            for (int j = 0; j < s[1]; ++j)
                for (int i = 0; i < s[0]; ++i) {
                    int A[3] = {i, j, k}, B[3] = {i, j, k};
                    A[a] = clampi(A[a] - 1, 0, n[a] - 1);
                    B[a] = clampi(B[a], 0, n[a] - 1);
                    float ra = r0 + (r1 - r0) * frac[orc_cidx(n, A[0], A[1], A[2])];
                    float rb = r0 + (r1 - r0) * frac[orc_cidx(n, B[0], B[1], B[2])];
                    float rf = (ra + rb) * 0.5f;
                    out[a][orc_cidx(s, i, j, k)] = ref / (rf > min_d ? rf : min_d);
                }
    }
}

/* ---- advectScalar3 (advection_solver3.js:285-357) through the collider
 *      (csdf or NULL): a cell whose centre lies in the collider keeps dst's
 *      value; order 2 as for the faces, corners at the cell centres ---- */
void orc_advect_scalar3_c(const float* src, float* dst, const float* u, const float* v, const float* w,
                          const int n[3], const float h[3], const float origin[3], const float* csdf, float dt,
                          int max_sub, int order) {
    float ou[3], ov[3], ow[3];
    mac_origins(origin, h, ou, ov, ow);
    const float hmin = fminf(h[0], fminf(h[1], h[2]));
    const float co[3] = {origin[0] + h[0] * 0.5f, origin[1] + h[1] * 0.5f, origin[2] + h[2] * 0.5f};
    const int s[3] = {n[0], n[1], n[2]};
    float* fwd = order == 2 ? (float*)malloc(sizeof(float) * (size_t)orc_cells(n)) : dst;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                float pos[3] = {co[0] + (float)i * h[0], co[1] + (float)j * h[1], co[2] + (float)k * h[2]};
                const int64_t c = orc_cidx(s, i, j, k);
                if (!(collider_sample(csdf, n, h, co, pos) > 0.0f)) {
                    if (order == 2) fwd[c] = src[c];
                    continue;
                }
                float tr[3];
                back_trace_c(u, v, w, n, h, ou, ov, ow, csdf, pos, dt, hmin, max_sub, 1.0f, tr);
                fwd[c] = sample_cubic(src, s, co, h, tr);
            }
    if (order != 2) return;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                const float pos[3] = {co[0] + (float)i * h[0], co[1] + (float)j * h[1], co[2] + (float)k * h[2]};
                if (!(collider_sample(csdf, n, h, co, pos) > 0.0f)) continue;
                float tb[3], tf[3];
                back_trace_c(u, v, w, n, h, ou, ov, ow, csdf, pos, dt, hmin, max_sub, -1.0f, tb);
                const float bwd = sample_cubic(fwd, s, co, h, tb);
                back_trace_c(u, v, w, n, h, ou, ov, ow, csdf, pos, dt, hmin, max_sub, 1.0f, tf);
                const int64_t c = orc_cidx(s, i, j, k);
                const float fv = fwd[c];
                const float corrected = fv + (src[c] - bwd) * 0.5f;
                dst[c] = mc_clamp_c(src, s, co, h, tf, corrected, fv, csdf, n, co);
            }
    free(fwd);
}

void orc_advect_scalar3(const float* src, float* dst, const float* u, const float* v, const float* w, const int n[3],
                        const float h[3], const float origin[3], float dt, int max_sub) {
    orc_advect_scalar3_c(src, dst, u, v, w, n, h, origin, NULL, dt, max_sub, 1);
}

void orc_advect_faces3_mc_vd(const float* const vel[3], const float* const data[3], float* const out[3],
                             const int n[3], const float h[3], const float origin[3], float dt, int max_sub) {
    orc_advect_faces3_c(vel, data, out, n, h, origin, NULL, dt, max_sub, 2);
}

void orc_advect_face3_mc(const float* u, const float* v, const float* w, float* uo_, float* vo_, float* wo_,
                         const int n[3], const float h[3], const float origin[3], float dt, int max_sub) {
    const float* vel[3] = {u, v, w};
    float* out[3] = {uo_, vo_, wo_};
    orc_advect_faces3_mc_vd(vel, vel, out, n, h, origin, dt, max_sub);
}

void orc_advect_scalar3_mc(const float* src, float* dst, const float* u, const float* v, const float* w,
                           const int n[3], const float h[3], const float origin[3], float dt, int max_sub) {
    orc_advect_scalar3_c(src, dst, u, v, w, n, h, origin, NULL, dt, max_sub, 2);
}

/* ---- trace (advection_solver3.js:84-88): backTrace of each point through
 *      the MAC velocity with the given direction, no collider.  Exposed so
 *      the MacCormack clamp's corners can be recomputed outside the
 *      oracle. ---- */
void orc_trace3(float* pos, int64_t count, const float* u, const float* v, const float* w, const int n[3],
                const float h[3], const float origin[3], float dt, int max_sub, float direction) {
    orc_trace3_c(pos, count, u, v, w, n, h, origin, NULL, dt, max_sub, direction);
}

/* the same through the collider (csdf or NULL) */
void orc_trace3_c(float* pos, int64_t count, const float* u, const float* v, const float* w, const int n[3],
                  const float h[3], const float origin[3], const float* csdf, float dt, int max_sub,
                  float direction) {
    float ou[3], ov[3], ow[3];
    mac_origins(origin, h, ou, ov, ow);
    const float hmin = fminf(h[0], fminf(h[1], h[2]));
    for (int64_t p = 0; p < count; ++p) {
        float tr[3];
        back_trace_c(u, v, w, n, h, ou, ov, ow, csdf, &pos[3 * p], dt, hmin, max_sub, direction, tr);
        pos[3 * p] = tr[0];
        pos[3 * p + 1] = tr[1];
        pos[3 * p + 2] = tr[2];
    }
}

/* ==== carried concentration and density coupling of createGridFlipSolver3
 *      (grid_flip_solver3.js:265-357, 493-531, 848-880, 967-1049) ==== */
#define ORC_MIN_DENSITY 1e-4f

static inline float clamp01(float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); }

/* accumulateConcentration (992-998): per cell the particle count and the sum
 * of round(clamp(c, 0, 1) * scale) as wrapping int32 atomics */
void orc_conc_accumulate3(const float* pos, const float* conc, int64_t count, const int n[3], const float h[3],
                          const float origin[3], float scale, int32_t* accum, int32_t* cnt) {
    int32_t* cell = (int32_t*)malloc(sizeof(int32_t) * (count > 0 ? count : 1));
    orc_cell_index3(pos, count, origin, h, n, cell);
    for (int64_t p = 0; p < count; ++p) {
        cnt[cell[p]] = (int32_t)((uint32_t)cnt[cell[p]] + 1u);
        accum[cell[p]] = (int32_t)((uint32_t)accum[cell[p]] + (uint32_t)f32_to_i32_sat(rintf(clamp01(conc[p]) * scale)));
    }
    free(cell);
}

/* resolveCellConcentration (1000-1012): mean concentration of a cell's
 * particles (0 for an empty cell), and with density coupling the cell density
 * d0 + (d1 - d0) c with d0, d1 floored at MIN_DENSITY */
void orc_conc_resolve3(const int32_t* accum, const int32_t* cnt, int64_t cells, float scale, int coupling, float d0,
                       float d1, float* cell_conc, float* cell_density) {
    const float a = d0 > ORC_MIN_DENSITY ? d0 : ORC_MIN_DENSITY, b = d1 > ORC_MIN_DENSITY ? d1 : ORC_MIN_DENSITY;
    for (int64_t c = 0; c < cells; ++c) {
        const float total = (float)cnt[c];
        const float sum = (float)accum[c] / scale;
        const float mean = sum / (total > 1.0f ? total : 1.0f);
        const float cc = total > 0.5f ? clamp01(mean) : 0.0f;
        cell_conc[c] = cc;
        if (coupling) cell_density[c] = a + (b - a) * cc;
    }
}

/* computeBetaU/V/W with faceDensity (313-357): beta = rho_ref / max(rho_face,
 * MIN_DENSITY), rho_face the mean of two fluid cells, the fluid one's own
 * density when one side is fluid, else rho_ref; edge faces clamp indices */
void orc_density_beta3(const float* density, const float* fluid_mask, const int n[3], float d0, float d1,
                       float* bu, float* bv, float* bw) {
    const float a = d0 > ORC_MIN_DENSITY ? d0 : ORC_MIN_DENSITY, b = d1 > ORC_MIN_DENSITY ? d1 : ORC_MIN_DENSITY;
    const float ref = a > b ? a : b;
    float* out[3] = {bu, bv, bw};
    for (int ax = 0; ax < 3; ++ax) {
        const int s[3] = {n[0] + (ax == 0), n[1] + (ax == 1), n[2] + (ax == 2)};
        for (int k = 0; k < s[2]; ++k)
            for (int j = 0; j < s[1]; ++j)
                for (int i = 0; i < s[0]; ++i) {
                    int A[3] = {i, j, k}, B[3] = {i, j, k};
                    A[ax] = clampi(A[ax] - 1, 0, n[ax] - 1);
                    B[ax] = clampi(B[ax], 0, n[ax] - 1);
                    const int64_t ca = orc_cidx(n, A[0], A[1], A[2]), cb = orc_cidx(n, B[0], B[1], B[2]);
                    const float ra = density[ca], rb = density[cb];
                    const int fa = fluid_mask[ca] > 0.5f, fb = fluid_mask[cb] > 0.5f;
                    const float rf = (fa && fb) ? (ra + rb) * 0.5f : ((fa || fb) ? (fa ? ra : rb) : ref);
                    out[ax][orc_cidx(s, i, j, k)] = ref / (rf > ORC_MIN_DENSITY ? rf : ORC_MIN_DENSITY);
                }
    }
}

/* applyReducedGravityU/V/W (848-880): u -= beta dt / rho_ref (g . x_face)
 * d(rho)/dx across the face, indices clamped at the walls */
void orc_reduced_gravity3(float* u, float* v, float* w, const float* const beta[3], const float* density,
                          const int n[3], const float h[3], const float origin[3], const float g[3], float dt, float d0,
                          float d1) {
    const float a = d0 > ORC_MIN_DENSITY ? d0 : ORC_MIN_DENSITY, b = d1 > ORC_MIN_DENSITY ? d1 : ORC_MIN_DENSITY;
    const float ref = a > b ? a : b;
    float* vel[3] = {u, v, w};
    for (int ax = 0; ax < 3; ++ax) {
        const int s[3] = {n[0] + (ax == 0), n[1] + (ax == 1), n[2] + (ax == 2)};
        for (int k = 0; k < s[2]; ++k)
            for (int j = 0; j < s[1]; ++j)
                for (int i = 0; i < s[0]; ++i) {
                    const int I[3] = {i, j, k};
                    float pos[3];
                    for (int q = 0; q < 3; ++q)
                        pos[q] = origin[q] + (q == ax ? (float)I[q] * h[q] : ((float)I[q] + 0.5f) * h[q]);
                    int A[3] = {i, j, k}, B[3] = {i, j, k};
                    A[ax] = clampi(A[ax] - 1, 0, n[ax] - 1);
                    B[ax] = clampi(B[ax], 0, n[ax] - 1);
                    const float drho = (density[orc_cidx(n, B[0], B[1], B[2])] - density[orc_cidx(n, A[0], A[1], A[2])]) / h[ax];
                    const float gdot = (g[0] * pos[0] + g[1] * pos[1]) + g[2] * pos[2];
                    const int64_t f = orc_cidx(s, i, j, k);
                    const float term = (((beta[ax][f] * dt) / ref) * gdot) * drho;
                    vel[ax][f] = vel[ax][f] - term;
                }
    }
}

/* mixAndFade (1018-1027): c <- clamp((c + (local - c) mixing) (1 - fade)),
 * local = the cell concentration sampled trilinearly at the particle (c
 * itself where that is not finite); mixing and fade clamped to [0, 1] */
void orc_mix_fade3(const float* pos, float* conc, int64_t count, const float* cell_conc, const int n[3],
                   const float h[3], const float origin[3], float mixing, float fade) {
    const float co[3] = {origin[0] + h[0] * 0.5f, origin[1] + h[1] * 0.5f, origin[2] + h[2] * 0.5f};
    const float mx = clamp01(mixing), fd = clamp01(fade);
    for (int64_t p = 0; p < count; ++p) {
        const float sampled = orc_sample_trilinear3(cell_conc, n, co, h, &pos[3 * p]);
        const float c = clamp01(conc[p]);
        const float local = isfinite(sampled) ? sampled : c;
        const float blended = c + (local - c) * mx;
        conc[p] = clamp01(blended * (1.0f - fd));
    }
}

/* buildScatter with massWeight (493-531): each corner weight times
 * rho(c) / rho_ref, rho(c) = d0 + (d1 - d0) clamp(c, 0, 1) */
void orc_p2g_fixed_mass3(const float* pos, const float* vel, const float* conc, int64_t count, int comp,
                         const int face_n[3], const float data_origin[3], const float h[3], float scale, float d0,
                         float d1, int32_t* numer, int32_t* denom) {
    const float a = d0 > ORC_MIN_DENSITY ? d0 : ORC_MIN_DENSITY, b = d1 > ORC_MIN_DENSITY ? d1 : ORC_MIN_DENSITY;
    const float ref = a > b ? a : b;
    for (int64_t p = 0; p < count; ++p) {
        orc_tri t;
        tri_weights(&pos[3 * p], data_origin, h, face_n, &t);
        const float value = vel[3 * p + comp];
        const float mass = (a + (b - a) * clamp01(conc[p])) / ref;
        for (int c = 0; c < 8; ++c) {
            int i, j, k;
            tri_corner(&t, c, &i, &j, &k);
            const int64_t f = orc_cidx(face_n, i, j, k);
            const float w = t.w[c] * mass;
            numer[f] = (int32_t)((uint32_t)numer[f] + (uint32_t)f32_to_i32_sat(rintf((value * w) * scale)));
            denom[f] = (int32_t)((uint32_t)denom[f] + (uint32_t)f32_to_i32_sat(rintf(w * scale)));
        }
    }
}
