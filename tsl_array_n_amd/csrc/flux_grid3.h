// Grid geometry and trilinear sampling shared by the particle/grid kernels
// (grid_math3.js:61-126, grid_data3.js:121-176).  Translation units that
// include it are built with -ffp-contract=off so every product and sum
// rounds like the oracle's.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/fluxflow_hip.h"

namespace flux {

struct G3 {
    int n0, n1, n2;
    float h0, h1, h2;
    float o0, o1, o2;
};

inline G3 make_g3(const flux_grid3* g) {
    return G3{g->n[0], g->n[1], g->n[2], g->h[0], g->h[1], g->h[2], g->origin[0], g->origin[1], g->origin[2]};
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// MAC data origins (grid_data3.js:128-130)
__device__ __forceinline__ void mac_origin(const G3& g, int comp, float o[3]) {
    o[0] = g.o0;
    o[1] = g.o1;
    o[2] = g.o2;
    if (comp != 0) o[0] = g.o0 + g.h0 * 0.5f;
    if (comp != 1) o[1] = g.o1 + g.h1 * 0.5f;
    if (comp != 2) o[2] = g.o2 + g.h2 * 0.5f;
}

struct Tri {
    int i0, i1, j0, j1, k0, k1;
    float w[8];  // 000,100,010,110,001,101,011,111
};

// trilinearCoordsAndWeights3 (grid_math3.js:61-95)
__device__ __forceinline__ void tri_weights(const float p[3], const float o[3], const float h[3], const int s[3],
                                            Tri& t) {
    float f[3];
    int i0[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float g = (p[a] - o[a]) / h[a];
        i0[a] = (int)floorf(g);
        f[a] = g - (float)i0[a];
    }
    t.i0 = clampi(i0[0], 0, s[0] - 1);
    t.i1 = clampi(i0[0] + 1, 0, s[0] - 1);
    t.j0 = clampi(i0[1], 0, s[1] - 1);
    t.j1 = clampi(i0[1] + 1, 0, s[1] - 1);
    t.k0 = clampi(i0[2], 0, s[2] - 1);
    t.k1 = clampi(i0[2] + 1, 0, s[2] - 1);
    float gx = 1.0f - f[0], gy = 1.0f - f[1], gz = 1.0f - f[2];
    float fx = f[0], fy = f[1], fz = f[2];
    t.w[0] = (gx * gy) * gz;
    t.w[1] = (fx * gy) * gz;
    t.w[2] = (gx * fy) * gz;
    t.w[3] = (fx * fy) * gz;
    t.w[4] = (gx * gy) * fz;
    t.w[5] = (fx * gy) * fz;
    t.w[6] = (gx * fy) * fz;
    t.w[7] = (fx * fy) * fz;
}

// collocatedValueAtPosition3 (grid_math3.js:99-113)
__device__ __forceinline__ float sample_tri(const float* __restrict__ d, const int s[3], const float o[3],
                                           const float h[3], const float p[3]) {
    Tri t;
    tri_weights(p, o, h, s, t);
    float acc = 0.0f;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        int i = (c & 1) ? t.i1 : t.i0, j = (c & 2) ? t.j1 : t.j0, k = (c & 4) ? t.k1 : t.k0;
        float term = d[i + s[0] * (j + s[1] * k)] * t.w[c];
        acc = (c == 0) ? term : acc + term;
    }
    return acc;
}

struct Mac {
    const float* u;
    const float* v;
    const float* w;
};

__device__ __forceinline__ void sample_mac(const G3& g, const Mac& m, const float p[3], float out[3]) {
    const float h[3] = {g.h0, g.h1, g.h2};
    const float* d[3] = {m.u, m.v, m.w};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        int s[3] = {g.n0 + (a == 0), g.n1 + (a == 1), g.n2 + (a == 2)};
        float o[3];
        mac_origin(g, a, o);
        out[a] = sample_tri(d[a], s, o, h, p);
    }
}

// A static collider's SDF, cell-centred on the velocity grid (the collider's
// sample(), sdf_collider3.js:33-37: collocatedValueAtPosition3); d == nullptr
// is no collider, which samples as 1 (advection_solver3.js:78-82).
struct Sdf3 {
    const float* d;
    int n0, n1, n2;
    float o0, o1, o2;  // the cell-centre origin
    float h0, h1, h2;
};

inline Sdf3 sdf3_on(const G3& g, const float* d) {
    return Sdf3{d, g.n0, g.n1, g.n2, g.o0 + g.h0 * 0.5f, g.o1 + g.h1 * 0.5f, g.o2 + g.h2 * 0.5f, g.h0, g.h1, g.h2};
}

__device__ __forceinline__ float sdf_sample(const Sdf3& c, const float p[3]) {
    if (!c.d) return 1.0f;
    const int s[3] = {c.n0, c.n1, c.n2};
    const float o[3] = {c.o0, c.o1, c.o2}, h[3] = {c.h0, c.h1, c.h2};
    return sample_tri(c.d, s, o, h, p);
}

// trilinearGradientAtPosition3 (grid_math3.js:295-356)
__device__ __forceinline__ void grad_axis(float coord, int n, int& lo, int& hi, float& t) {
    const int idx = (int)floorf(coord);
    int l = idx < n - 2 ? idx : n - 2;
    l = l > 0 ? l : 0;
    lo = l;
    hi = l + 1 < n - 1 ? l + 1 : n - 1;
    const float tt = coord - (float)l;
    t = tt < 0.0f ? 0.0f : (tt > 1.0f ? 1.0f : tt);
}

__device__ __forceinline__ void sdf_gradient(const float* __restrict__ d, const int n[3], const float o[3], const float h[3],
                             const float p[3], float g[3]) {
    int xl, xh, yl, yh, zl, zh;
    float tx, ty, tz;
    grad_axis((p[0] - o[0]) / h[0], n[0], xl, xh, tx);
    grad_axis((p[1] - o[1]) / h[1], n[1], yl, yh, ty);
    grad_axis((p[2] - o[2]) / h[2], n[2], zl, zh, tz);
    auto D = [&](int i, int j, int k) { return d[i + n[0] * (j + n[1] * k)]; };
    const float v000 = D(xl, yl, zl), v100 = D(xh, yl, zl), v010 = D(xl, yh, zl), v110 = D(xh, yh, zl);
    const float v001 = D(xl, yl, zh), v101 = D(xh, yl, zh), v011 = D(xl, yh, zh), v111 = D(xh, yh, zh);
    g[0] = ((((1.0f - ty) * (1.0f - tz)) * (v100 - v000) + (ty * (1.0f - tz)) * (v110 - v010)) +
            ((1.0f - ty) * tz) * (v101 - v001) + (ty * tz) * (v111 - v011)) / h[0];
    g[1] = ((((1.0f - tx) * (1.0f - tz)) * (v010 - v000) + (tx * (1.0f - tz)) * (v110 - v100)) +
            ((1.0f - tx) * tz) * (v011 - v001) + (tx * tz) * (v111 - v101)) / h[1];
    g[2] = ((((1.0f - tx) * (1.0f - ty)) * (v001 - v000) + (tx * (1.0f - ty)) * (v101 - v100)) +
            ((1.0f - tx) * ty) * (v011 - v010) + (tx * ty) * (v111 - v110)) / h[2];
}

// backTrace (advection_solver3.js:20-62): RK2 midpoint sub-steps; with a
// collider a sub-step whose end lies in it (phi1 <= 0) stops at the
// SDF-weighted blend of its two ends, w = |phi1| / (|phi0| + |phi1|), and
// ends the trace.
__device__ __forceinline__ void back_trace_c(const G3& g, const Mac& m, const Sdf3& col, const float start[3],
                                             float dt, float hmin, int max_sub, float direction, float out[3]) {
    float pt[3] = {start[0], start[1], start[2]};
    float rem = dt;
    for (int s = 0; s < max_sub; ++s) {
        if (rem <= 1e-6f) break;
        float v0[3];
        sample_mac(g, m, pt, v0);
        float len = sqrtf((v0[0] * v0[0] + v0[1] * v0[1]) + v0[2] * v0[2]);
        float ns = ceilf((len * rem) / hmin);
        if (ns < 1.0f) ns = 1.0f;
        float sub = rem / ns;
        float half = sub * 0.5f;
        float mid[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) mid[a] = pt[a] - (v0[a] * half) * direction;
        float mv[3];
        sample_mac(g, m, mid, mv);
        float next[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) next[a] = pt[a] - (mv[a] * sub) * direction;
        if (col.d) {
            const float phi0 = sdf_sample(col, pt);
            const float phi1 = sdf_sample(col, next);
            if (phi1 <= 0.0f) {
                const float w = fabsf(phi1) / (fabsf(phi0) + fabsf(phi1));
#pragma unroll
                for (int a = 0; a < 3; ++a) pt[a] = pt[a] * w + next[a] * (1.0f - w);
                break;
            }
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) pt[a] = next[a];
        rem = rem - sub;
    }
    out[0] = pt[0];
    out[1] = pt[1];
    out[2] = pt[2];
}

// backTrace without a collider
__device__ __forceinline__ void back_trace(const G3& g, const Mac& m, const float start[3], float dt, float hmin,
                                           int max_sub, float direction, float out[3]) {
    const Sdf3 none{nullptr, 0, 0, 0, 0.f, 0.f, 0.f, 1.f, 1.f, 1.f};
    back_trace_c(g, m, none, start, dt, hmin, max_sub, direction, out);
}

// monotonicCubic1d3 (grid_math3.js:137-157)
__device__ __forceinline__ float mono_cubic(float f0, float f1, float f2, float f3, float f) {
    float D1 = f2 - f1;
    float rawD1 = (f2 - f0) * 0.5f;
    float rawD2 = (f3 - f1) * 0.5f;
    bool flat = fabsf(D1) < 1e-12f;
    float d1 = (flat || rawD1 * D1 < 0.0f) ? 0.0f : rawD1;
    float d2 = (flat || rawD2 * D1 < 0.0f) ? 0.0f : rawD2;
    float a3 = (d1 + d2) - D1 * 2.0f;
    float a2 = (D1 * 3.0f - d1 * 2.0f) - d2;
    return ((((a3 * f) * f) * f + (a2 * f) * f) + d1 * f) + f1;
}

// collocatedCubicValueAtPosition3 (grid_math3.js:183-207)
__device__ __forceinline__ float sample_cubic(const float* __restrict__ d, const int s[3], const float o[3],
                                              const float h[3], const float p[3]) {
    int xi[4], yi[4], zi[4];
    float f[3];
    int* idx[3] = {xi, yi, zi};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float coord = (p[a] - o[a]) / h[a];
        int i0 = (int)floorf(coord);
        f[a] = coord - (float)i0;
        idx[a][0] = clampi(i0 - 1, 0, s[a] - 1);
        idx[a][1] = clampi(i0, 0, s[a] - 1);
        idx[a][2] = clampi(i0 + 1, 0, s[a] - 1);
        idx[a][3] = clampi(i0 + 2, 0, s[a] - 1);
    }
    float zv[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        float rv[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            int base = s[0] * (yi[b] + s[1] * zi[c]);
            rv[b] = mono_cubic(d[xi[0] + base], d[xi[1] + base], d[xi[2] + base], d[xi[3] + base], f[0]);
        }
        zv[c] = mono_cubic(rv[0], rv[1], rv[2], rv[3], f[1]);
    }
    return mono_cubic(zv[0], zv[1], zv[2], zv[3], f[2]);
}

// advectFaceCentered3, order 1, no collider (advection_solver3.js:130-176)
}  // namespace flux
