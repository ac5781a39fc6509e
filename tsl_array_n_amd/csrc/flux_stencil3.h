// 3-D specialisation of the operator for the hot kernels.
//
// Same operator as stencil_row (laplacianAt / laplacianDiagonalAt,
// multigrid.js:273-339, 399-437), written in its coupling form:
//   (A x)_c = sum over present faces f of  g_f * (xn_f - x_c),
//   g_f = w_f / h_axis^2,  xn_f = 0 when the neighbour is Dirichlet,
//   diag_c = -sum over present faces of g_f;  a Dirichlet row is -x_c (diag -1).
// (Expanding dLo/dUp of the reference gives exactly these terms.)
//
// Launch shape: blocks of 64 x 4 threads over (x, y); each thread marches
// over a chunk of z-planes, so there is no index division and consecutive
// planes share the thread's z-neighbour reads in cache.
#pragma once
#include "flux_stencil.h"

namespace flux {

static constexpr int S3_BX = 64, S3_BY = 4;

struct Row3 {
    float acc;   // sum of g_f * (xn_f - x_c) over present faces
    float diag;  // -sum of g_f
};

// Evaluates one free row; `ld(q)` reads x (as float) at flat index q.
template <typename WT, typename LD>
__device__ __forceinline__ Row3 row3(const LevelDevT<WT>& L, int i, int j, int k, int c, float xc, LD ld) {
    const uint8_t* __restrict__ m = L.mask;
    const int n0 = L.n0, n1 = L.n1, n2 = L.n2;
    const int plane = n0 * n1;
    float acc = 0.0f, gs = 0.0f;
    // x faces
    {
        const int fl = i + (n0 + 1) * (j + n1 * k);
        float glo = L.w0 ? to_f32(L.w0[fl]) : 1.0f;
        float gup = L.w0 ? to_f32(L.w0[fl + 1]) : 1.0f;
        glo *= L.ih2[0];
        gup *= L.ih2[0];
        if (i > 0) {
            float xn = (m && m[c - 1]) ? 0.0f : ld(c - 1);
            acc += glo * (xn - xc);
            gs += glo;
        }
        if (i < n0 - 1) {
            float xn = (m && m[c + 1]) ? 0.0f : ld(c + 1);
            acc += gup * (xn - xc);
            gs += gup;
        }
    }
    // y faces
    {
        const int fl = i + n0 * (j + (n1 + 1) * k);
        float glo = L.w1 ? to_f32(L.w1[fl]) : 1.0f;
        float gup = L.w1 ? to_f32(L.w1[fl + n0]) : 1.0f;
        glo *= L.ih2[1];
        gup *= L.ih2[1];
        if (j > 0) {
            float xn = (m && m[c - n0]) ? 0.0f : ld(c - n0);
            acc += glo * (xn - xc);
            gs += glo;
        }
        if (j < n1 - 1) {
            float xn = (m && m[c + n0]) ? 0.0f : ld(c + n0);
            acc += gup * (xn - xc);
            gs += gup;
        }
    }
    // z faces
    {
        float glo = L.w2 ? to_f32(L.w2[c]) : 1.0f;
        float gup = L.w2 ? to_f32(L.w2[c + plane]) : 1.0f;
        glo *= L.ih2[2];
        gup *= L.ih2[2];
        if (k > 0) {
            float xn = (m && m[c - plane]) ? 0.0f : ld(c - plane);
            acc += glo * (xn - xc);
            gs += glo;
        }
        if (k < n2 - 1) {
            float xn = (m && m[c + plane]) ? 0.0f : ld(c + plane);
            acc += gup * (xn - xc);
            gs += gup;
        }
    }
    return Row3{acc, -gs};
}

// The same row in FP64 arithmetic (the PCG's true residual).
template <typename LD>
__device__ __forceinline__ double row3d(const LevelDev& L, int i, int j, int k, int c, double xc, LD ld) {
    const uint8_t* __restrict__ m = L.mask;
    const int n0 = L.n0, n1 = L.n1, n2 = L.n2;
    const int plane = n0 * n1;
    double acc = 0.0;
    {
        const int fl = i + (n0 + 1) * (j + n1 * k);
        double glo = (L.w0 ? (double)L.w0[fl] : 1.0) * L.ih2d[0];
        double gup = (L.w0 ? (double)L.w0[fl + 1] : 1.0) * L.ih2d[0];
        if (i > 0) acc += glo * (((m && m[c - 1]) ? 0.0 : ld(c - 1)) - xc);
        if (i < n0 - 1) acc += gup * (((m && m[c + 1]) ? 0.0 : ld(c + 1)) - xc);
    }
    {
        const int fl = i + n0 * (j + (n1 + 1) * k);
        double glo = (L.w1 ? (double)L.w1[fl] : 1.0) * L.ih2d[1];
        double gup = (L.w1 ? (double)L.w1[fl + n0] : 1.0) * L.ih2d[1];
        if (j > 0) acc += glo * (((m && m[c - n0]) ? 0.0 : ld(c - n0)) - xc);
        if (j < n1 - 1) acc += gup * (((m && m[c + n0]) ? 0.0 : ld(c + n0)) - xc);
    }
    {
        double glo = (L.w2 ? (double)L.w2[c] : 1.0) * L.ih2d[2];
        double gup = (L.w2 ? (double)L.w2[c + plane] : 1.0) * L.ih2d[2];
        if (k > 0) acc += glo * (((m && m[c - plane]) ? 0.0 : ld(c - plane)) - xc);
        if (k < n2 - 1) acc += gup * (((m && m[c + plane]) ? 0.0 : ld(c + plane)) - xc);
    }
    return acc;
}

}  // namespace flux
