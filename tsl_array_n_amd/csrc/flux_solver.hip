// Operator, multigrid preconditioner and GPU-resident PCG for gfx950.
//
// Reference: packages/fluxflow/src/linalg/multigrid.js (operator, V-cycle)
// and linalg.js (PCG).  All arithmetic of the V-cycle and of the search
// directions is FP32; the solution accumulator and the true residual are
// FP64, and every dot product is reduced in FP64 in a fixed order.
#include <vector>
#include <algorithm>
#include <cmath>
#include <cstring>

#include "flux_internal.h"
#include "flux_solver.h"
#include "flux_stencil.h"
#include "flux_stencil3.h"

namespace flux {

static constexpr int TPB = 256;

LevelDev make_level(int ndim, const int* n, const double* h, const uint8_t* mask, const float* const* w) {
    LevelDev L;
    L.n0 = n[0];
    L.n1 = n[1];
    L.n2 = n[2];
    L.ndim = ndim;
    for (int a = 0; a < 3; ++a) {
        double hh = (a < ndim) ? h[a] : 1.0;
        L.ih2d[a] = 1.0 / (hh * hh);
        L.ih2[a] = (float)L.ih2d[a];
    }
    L.mask = mask;
    L.w0 = w ? w[0] : nullptr;
    L.w1 = w ? w[1] : nullptr;
    L.w2 = (w && ndim == 3) ? w[2] : nullptr;
    return L;
}

// ------------------------------------------------------------------ kernels
__global__ void k_fill_f32(float* p, int n, float v) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}
__global__ void k_fill_f64(double* p, int n, double v) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}
__global__ void k_fill_i32(int* p, int n, int v) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}
__global__ void k_fill_u8(uint8_t* p, int n, uint8_t v) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

// out = A in  (createLaplacianOperator)
__global__ void k_apply(LevelDev L, const float* __restrict__ in, float* __restrict__ out) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= L.n0 * L.n1 * L.n2) return;
    int i, j, k;
    cell_coords(L, c, i, j, k);
    out[c] = stencil_row<float>(L, i, j, k, c, [&](int q) { return in[q]; }, (float*)nullptr);
}

__global__ void k_diag(LevelDev L, float* __restrict__ out) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= L.n0 * L.n1 * L.n2) return;
    int i, j, k;
    cell_coords(L, c, i, j, k);
    float d;
    (void)stencil_row<float>(L, i, j, k, c, [&](int) { return 0.0f; }, &d);
    out[c] = d;
}

// Jacobi preconditioner z = r / diag (createJacobiPreconditioner, multigrid.js:568-596)
__global__ void k_jacobi(LevelDev L, const float* __restrict__ r, float* __restrict__ z, const int* __restrict__ stop) {
    if (stop && *stop) return;
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= L.n0 * L.n1 * L.n2) return;
    int i, j, k;
    cell_coords(L, c, i, j, k);
    float d;
    (void)stencil_row<float>(L, i, j, k, c, [&](int) { return 0.0f; }, &d);
    z[c] = (fabsf(d) > 0.0f) ? r[c] / d : r[c];
}

__global__ void k_copy_f32(const float* __restrict__ a, float* __restrict__ b, int n, const int* __restrict__ stop) {
    if (stop && *stop) return;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) b[t] = a[t];
}

// ------------------------------------------------------------ PCG kernels
// PCG scalar block kept on the GPU (the reference's 9-slot `scalars`,
// linalg.js:1030-1047, widened to FP64).
struct Scal {
    double pAp, rr, rz, rz_old, bb, base, alpha, beta, best, thr, tol, rr0;
    int stop, iter, stop_iter, relative;
    int xsel, psel, flexible, cs_m;
    int restarts, pad2;
    double rzo;  // r_{k+1} . z_k of the flexible beta
    // coarse-space correction (cs_m vectors): Einv = (Phi^T A Phi)^-1,
    // c = Phi^T r of the current residual, a = Einv c the coarse amplitudes
    double cs_einv[16], cs_c[4], cs_a[4];
};

// The coarse space of the additive correction: column 0 the free-cell
// indicator when `free0`, then up to three caller fields, each restricted to
// the free cells.
static constexpr int CS_MAX = 4;
struct CoarseSpace {
    const float* f[3] = {nullptr, nullptr, nullptr};
    int nf = 0;
    int free0 = 0;
    __host__ __device__ int m() const { return free0 + nf; }
};

// phi_q at cell c (0 on Dirichlet cells)
__device__ __forceinline__ float cs_phi(const CoarseSpace& cs, int q, int c, bool dir) {
    if (dir) return 0.0f;
    if (cs.free0) {
        if (q == 0) return 1.0f;
        return cs.f[q - 1][c];
    }
    return cs.f[q][c];
}

// a = Einv c, and the flexible beta's r_{k+1}.(z_k + Phi a_k) = rzo + a_k . c
__device__ __forceinline__ void cs_solve(Scal* S, bool fold_rzo) {
    const int m = S->cs_m;
    if (fold_rzo)
        for (int q = 0; q < m; ++q) S->rzo += S->cs_a[q] * S->cs_c[q];
    for (int q = 0; q < m; ++q) {
        double v = 0;
        for (int t = 0; t < m; ++t) v += S->cs_einv[q * CS_MAX + t] * S->cs_c[t];
        S->cs_a[q] = v;
    }
}

__device__ __forceinline__ bool degenerate(double v) { return !(fabs(v) <= 1.7976931348623157e308) || v == 0.0; }

// Every reduction is two-stage and deterministic: each block of the producing
// kernel stores its block sum, and a one-block finishing kernel adds the
// partials in index order and runs the scalar step.  The kernel boundary
// makes the partials visible across XCDs (no per-block cache write-back).
static constexpr int FIN_TPB = 1024;

// Setup: x64 = x; r = b - A x (FP64 arithmetic); partials of rr.
__global__ void k_pcg_setup_residual(LevelDev L, const float* __restrict__ x32, double* __restrict__ x64,
                                     const float* __restrict__ b, float* __restrict__ r, double* partials) {
    double v = 0;
    int n = L.n0 * L.n1 * L.n2;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        int i, j, k;
        cell_coords(L, c, i, j, k);
        double ax = stencil_row<double>(L, i, j, k, c, [&](int q) { return (double)x32[q]; }, (double*)nullptr);
        double rv = (double)b[c] - ax;
        x64[c] = (double)x32[c];
        r[c] = (float)rv;
        v += rv * rv;
    }
    store_block_partial(v, partials);
}

// The same in the column-march shape of k_pcg_update3 (one cell per thread
// and plane, z-neighbours carried), the PCG's setup at every solve.
__global__ void __launch_bounds__(TPB) k_pcg_setup_residual3(LevelDev L, March M, const float* __restrict__ x32,
                                                             double* __restrict__ x64, const float* __restrict__ b,
                                                             float* __restrict__ r, double* partials) {
    double v = 0;
    int t, k0, k1;
    if (march_begin(M, t, k0, k1)) {
        const int n0 = L.n0, n1 = L.n1, n2 = L.n2;
        const int plane = n0 * n1;
        const int j = t / n0;
        const int i = t - j * n0;
        const uint8_t* __restrict__ nbits = L.nb;
        const double ih2[3] = {L.ih2d[0], L.ih2d[1], L.ih2d[2]};
        double xnext = (double)x32[i + n0 * (j + n1 * k0)];
        double xprev = 0.0;
        for (int k = k0; k < k1; ++k) {
            const int c = i + n0 * (j + n1 * k);
            const double xc = xnext;
            if (k + 1 < n2) xnext = (double)x32[c + plane];
            const unsigned bits = nbits ? nbits[c] : 0u;
            float w[6];
            face_w3(L, i, j, k, c, w);
            double xn[6];
            xn[0] = (double)x32[i > 0 ? c - 1 : c];
            xn[1] = (double)x32[i < n0 - 1 ? c + 1 : c];
            xn[2] = (double)x32[j > 0 ? c - n0 : c];
            xn[3] = (double)x32[j < n1 - 1 ? c + n0 : c];
            xn[4] = k > k0 ? xprev : (double)x32[k > 0 ? c - plane : c];
            xn[5] = k + 1 < n2 ? xnext : xc;
#pragma unroll
            for (int a = 0; a < 6; ++a) xn[a] = ((bits >> a) & 1u) ? 0.0 : xn[a];
            double sum, dsum;
            row_terms<double>(xc, xn, w, ih2, sum, dsum);
            const double ax = (bits & 64u) ? -xc : sum;
            x64[c] = xc;
            const double rv = (double)b[c] - ax;
            r[c] = (float)rv;
            v += rv * rv;
            xprev = xc;
        }
    }
    store_block_partial(v, partials);
}

// Terms of a Galerkin-scaled start along a direction g (free cells only):
// partials[blk] = g.(b - A x), partials[nb + blk] = g.(A g), in FP64.
__global__ void k_galerkin_terms(LevelDev L, const float* __restrict__ x, const float* __restrict__ b,
                                 const float* __restrict__ g, double* partials) {
    double gr = 0, gag = 0;
    const int n = L.n0 * L.n1 * L.n2;
    const uint8_t* __restrict__ m = L.mask;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        int i, j, k;
        cell_coords(L, c, i, j, k);
        const double gc = (m && m[c]) ? 0.0 : (double)g[c];
        const double ax = stencil_row<double>(L, i, j, k, c, [&](int q) { return (double)x[q]; }, (double*)nullptr);
        const double ag = stencil_row<double>(
            L, i, j, k, c, [&](int q) { return (m && m[q]) ? 0.0 : (double)g[q]; }, (double*)nullptr);
        gr += gc * ((double)b[c] - ax);
        gag += gc * ag;
    }
    store_block_partial(gr, partials);
    store_block_partial(gag, partials + gridDim.x);
}

// x += alpha g on the free cells, rounded to FP32 once.
__global__ void k_galerkin_step(const uint8_t* __restrict__ m, const float* __restrict__ g, float* __restrict__ x,
                                double alpha, int n) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n || (m && m[c])) return;
    x[c] = (float)((double)x[c] + alpha * (double)g[c]);
}

__global__ void __launch_bounds__(FIN_TPB) k_fin_setup(Scal* S, const double* partials, unsigned nb) {
    const double tot = sum_partials(partials, nb);
    if (threadIdx.x != 0) return;
    S->rr = tot;
    S->rr0 = tot;
    S->xsel = 0;
    S->psel = 0;
    S->iter = 0;
    S->stop_iter = 0;
    S->stop = 0;
}

__global__ void k_pcg_dot_bb(const float* __restrict__ b, int n, double* partials) {
    double v = 0;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x)
        v += (double)b[c] * (double)b[c];
    store_block_partial(v, partials);
}

__global__ void __launch_bounds__(FIN_TPB) k_fin_bb(Scal* S, const double* partials, unsigned nb) {
    const double tot = sum_partials(partials, nb);
    if (threadIdx.x != 0) return;
    S->bb = tot;
    double bn = sqrt(fabs(tot));
    S->thr = S->relative ? fmax(S->tol * bn, S->tol) : S->tol;
    // host solve(): an initial residual already below the threshold
    // skips the loop and reports convergence (linalg.js:2205)
    if (sqrt(fabs(S->rr)) < S->thr) {
        S->stop = 5;
        S->stop_iter = 0;
    }
}

// partials of rz = r.z
__global__ void k_pcg_dot_rz(const float* __restrict__ r, const float* __restrict__ z, int n, const Scal* S,
                             double* partials) {
    if (S->stop) return;
    double v = 0;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x)
        v += (double)r[c] * (double)z[c];
    store_block_partial(v, partials);
}

// rz; seed (first=1: seedScalarsKernel linalg.js:1238-1244) or beta
// (betaKernel, linalg.js:1198-1226).
__global__ void __launch_bounds__(FIN_TPB) k_fin_rz(Scal* S, const double* partials, unsigned nb, int first) {
    if (S->stop) return;
    double tot = sum_partials(partials, nb);
    if (threadIdx.x != 0) return;
    // r.(z + Phi a) = r.z + a.(Phi^T r)
    for (int q = 0; q < S->cs_m; ++q) tot += S->cs_a[q] * S->cs_c[q];
    S->rz = tot;
    if (first) {
        S->rz_old = tot;
        S->base = fmax(fabs(S->rr0), 1e-12);
        S->beta = 0.0;
        S->best = -1.0;
        return;
    }
    double rzOld = S->rz_old;
    if (degenerate(rzOld)) {
        S->stop = 4;  // degenerate-oldRZ
        S->beta = 0.0;
        return;
    }
    bool flip = (rzOld > 0) != (tot > 0);
    // flexible (Polak-Ribiere) beta = r_{k+1}.(z_{k+1} - z_k) / r_k.z_k when
    // the preconditioner is inexact (mixed-precision storage); otherwise the
    // reference's r_{k+1}.z_{k+1} / r_k.z_k (linalg.js:1198-1226)
    double braw = (S->flexible ? tot - S->rzo : tot) / rzOld;
    bool restart = flip || !(fabs(braw) <= 10.0);  // MAX_BETA_MAGNITUDE
    S->beta = restart ? 0.0 : braw;
    S->restarts += restart ? 1 : 0;
    S->rz_old = tot;
}

// Iteration step 1: p = z + beta p (updateP, fused); partials of p.Ap.  Ap
// itself is not stored: the update recomputes r = b - A x from x.  Column-march shape: one cell per thread and plane.
// FREE, NF: the coarse space's shape (free-cell indicator column, number of
// caller fields), fixed at compile time so that every load of a cell's
// values is issued unconditionally and all of them are in flight together.
template <int FREE, int NF>
__global__ void __launch_bounds__(TPB) k_pcg_apply3(LevelDev L, March M, const float* __restrict__ z,
                                                    float* __restrict__ pa, float* __restrict__ pb, const Scal* S,
                                                    CoarseSpace cs, double* partials) {
    if (S->stop) return;
    const float* __restrict__ pold = S->psel ? pb : pa;
    float* __restrict__ pnew = S->psel ? pa : pb;
    const float beta = (float)S->beta;
    constexpr int csm = FREE + NF;
    float ca[CS_MAX] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int q = 0; q < csm; ++q) ca[q] = (float)S->cs_a[q];
    double v = 0;
    int t, k0, k1;
    if (march_begin(M, t, k0, k1)) {
        const int n0 = L.n0, n1 = L.n1, n2 = L.n2;
        const int plane = n0 * n1;
        const int j = t / n0;
        const int i = t - j * n0;
        const uint8_t* __restrict__ nbits = L.nb;
        const float ih2[3] = {L.ih2[0], L.ih2[1], L.ih2[2]};
        // the preconditioned residual with the coarse correction, z + Phi a
        // (Phi is 0 on Dirichlet cells; a Dirichlet neighbour's value is
        // masked to 0 below, so only centre values need the cell's own bit)
        auto zval = [&](int q, bool dir) {
            float v = z[q];
#pragma unroll
            for (int t = 0; t < csm; ++t) {
                const float phi = (FREE && t == 0) ? 1.0f : cs.f[t - FREE][q];
                v = v + ca[t] * (dir ? 0.0f : phi);
            }
            return v;
        };
        auto pval = [&](int q, bool dir) { return zval(q, dir) + pold[q] * beta; };
        // the z-neighbours are the previous and the next plane's centres
        const int c0 = i + n0 * (j + n1 * k0);
        float pnext = pval(c0, csm > 0 && nbits && ((nbits[c0] >> 6) & 1u));
        float pprev = 0.0f;
        for (int k = k0; k < k1; ++k) {
            const int c = i + n0 * (j + n1 * k);
            const float pc = pnext;
            const unsigned bits = nbits ? nbits[c] : 0u;
            // bit 5: the +z neighbour (the next plane's centre) is Dirichlet
            if (k + 1 < n2) pnext = pval(c + plane, (bits >> 5) & 1u);
            float w[6];
            face_w3(L, i, j, k, c, w);
            float xn[6];
            xn[0] = pval(i > 0 ? c - 1 : c, false);
            xn[1] = pval(i < n0 - 1 ? c + 1 : c, false);
            xn[2] = pval(j > 0 ? c - n0 : c, false);
            xn[3] = pval(j < n1 - 1 ? c + n0 : c, false);
            xn[4] = k > k0 ? pprev : pval(k > 0 ? c - plane : c, false);
            xn[5] = k + 1 < n2 ? pnext : pc;
#pragma unroll
            for (int a = 0; a < 6; ++a) xn[a] = ((bits >> a) & 1u) ? 0.0f : xn[a];
            float sum, dsum;
            row_terms<float>(pc, xn, w, ih2, sum, dsum);
            const float ap = (bits & 64u) ? -pc : sum;
            pnew[c] = pc;
            v += (double)pc * (double)ap;
            pprev = pc;
        }
    }
    store_block_partial(v, partials);
}

// pAp; alpha with the guards of alphaKernel (linalg.js:1147-1196).
__global__ void __launch_bounds__(FIN_TPB) k_fin_pap(Scal* S, const double* partials, unsigned nb) {
    if (S->stop) return;
    const double pAp = sum_partials(partials, nb);
    if (threadIdx.x != 0) return;
    S->pAp = pAp;
    S->psel ^= 1;
    S->iter += 1;
    double alpha = S->rz_old / pAp;
    int code = 0;
    if (degenerate(pAp)) code = 1;                            // degenerate-pAp
    else if (fabs(pAp) > S->base * 1e8) code = 2;             // MAX_PAP_GROWTH_FACTOR
    else if (!(fabs(alpha) <= 1e6)) code = 3;                 // MAX_ALPHA_MAGNITUDE
    S->alpha = code ? 0.0 : alpha;
    if (code) S->stop = code;
}

// Iteration step 2: x += alpha p (FP64 accumulator, ping-pong buffers);
// r = b - A x in FP64 arithmetic (the true residual, recomputed every
// iteration); partials of rr.
// With `zold` (flexible beta) the partials of r_{k+1}.z_k follow at
// partials[gridDim.x ...].
// With a coarse space, the partials of Phi^T r follow (one array of
// gridDim.x per column, after the rr and r.z_k arrays).
template <int FREE, int NF>
__global__ void __launch_bounds__(TPB) k_pcg_update3(LevelDev L, March M, double* __restrict__ xa,
                                                     double* __restrict__ xb, const float* __restrict__ pa,
                                                     const float* __restrict__ pb, const float* __restrict__ b,
                                                     float* __restrict__ r, const float* __restrict__ zold,
                                                     const Scal* S, CoarseSpace cs, double* partials) {
    if (S->stop) return;
    const double* __restrict__ xold = S->xsel ? xb : xa;
    double* __restrict__ xnew = S->xsel ? xa : xb;
    const float* __restrict__ p = S->psel ? pb : pa;  // the p written by k_pcg_apply3
    const double alpha = S->alpha;
    double v = 0, vz = 0;
    double vc[CS_MAX] = {0, 0, 0, 0};
    constexpr int csm = FREE + NF;
    int t, k0, k1;
    if (march_begin(M, t, k0, k1)) {
        const int n0 = L.n0, n1 = L.n1, n2 = L.n2;
        const int plane = n0 * n1;
        const int j = t / n0;
        const int i = t - j * n0;
        const uint8_t* __restrict__ nbits = L.nb;
        const double ih2[3] = {L.ih2d[0], L.ih2d[1], L.ih2d[2]};
        auto xval = [&](int q) { return xold[q] + alpha * (double)p[q]; };
        // the z-neighbours are the previous and the next plane's centres
        double xnext = xval(i + n0 * (j + n1 * k0));
        double xprev = 0.0;
        for (int k = k0; k < k1; ++k) {
            const int c = i + n0 * (j + n1 * k);
            const double xc = xnext;
            if (k + 1 < n2) xnext = xval(c + plane);
            const unsigned bits = nbits ? nbits[c] : 0u;
            float w[6];
            face_w3(L, i, j, k, c, w);
            double xn[6];
            xn[0] = xval(i > 0 ? c - 1 : c);
            xn[1] = xval(i < n0 - 1 ? c + 1 : c);
            xn[2] = xval(j > 0 ? c - n0 : c);
            xn[3] = xval(j < n1 - 1 ? c + n0 : c);
            xn[4] = k > k0 ? xprev : xval(k > 0 ? c - plane : c);
            xn[5] = k + 1 < n2 ? xnext : xc;
#pragma unroll
            for (int a = 0; a < 6; ++a) xn[a] = ((bits >> a) & 1u) ? 0.0 : xn[a];
            double sum, dsum;
            row_terms<double>(xc, xn, w, ih2, sum, dsum);
            const double ax = (bits & 64u) ? -xc : sum;
            xnew[c] = xc;
            const double rv = (double)b[c] - ax;
            const float rf = (float)rv;
            r[c] = rf;
            v += rv * rv;
            if (zold) vz += (double)rf * (double)zold[c];
            const bool dir = (bits & 64u) != 0;
#pragma unroll
            for (int q = 0; q < csm; ++q) {
                const float phi = (FREE && q == 0) ? 1.0f : cs.f[q - FREE][c];
                vc[q] += (double)rf * (double)(dir ? 0.0f : phi);
            }
            xprev = xc;
        }
    }
    store_block_partial(v, partials);
    if (zold) store_block_partial(vz, partials + gridDim.x);
#pragma unroll
    for (int q = 0; q < csm; ++q) store_block_partial(vc[q], partials + (size_t)(2 + q) * gridDim.x);
}

// rr; convergence and residual-growth tests.
__global__ void __launch_bounds__(FIN_TPB) k_fin_rr(Scal* S, const double* partials, unsigned nb) {
    if (S->stop) return;
    const double rr = sum_partials(partials, nb);
    const double rzo = S->flexible ? sum_partials(partials + nb, nb) : 0.0;
    double c[CS_MAX];
    for (int q = 0; q < S->cs_m; ++q) c[q] = sum_partials(partials + (size_t)(2 + q) * nb, nb);
    if (threadIdx.x != 0) return;
    S->rzo = rzo;
    for (int q = 0; q < S->cs_m; ++q) S->cs_c[q] = c[q];
    cs_solve(S, S->flexible != 0);
    S->rr = rr;
    S->xsel ^= 1;
    double res = sqrt(fabs(rr));
    if (res <= S->thr) {  // isNonFiniteOrAbove(residual, threshold).not()
        S->stop = 5;
        S->stop_iter = S->iter;
        return;
    }
    if (!(res == res)) return;
    if (S->best < 0 || fabs(rr) < S->best) S->best = fabs(rr);
    else if (S->best > 0 && fabs(rr) > S->best * 100.0) S->stop = 6;  // MAX_RESIDUAL_GROWTH
}

// Coarse-space setup: partials of E_qt = phi_q . A phi_t (FP64, q, t < m)
// at partials[(q * m + t) * gridDim.x + blk].
__global__ void __launch_bounds__(TPB) k_cs_gram(LevelDev L, CoarseSpace cs, double* __restrict__ partials) {
    const int m = cs.m();
    double e[CS_MAX * CS_MAX];
    for (int q = 0; q < m * m; ++q) e[q] = 0.0;
    const int n = L.n0 * L.n1 * L.n2;
    const uint8_t* __restrict__ msk = L.mask;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        if (msk && msk[c]) continue;
        int i, j, k;
        cell_coords(L, c, i, j, k);
        for (int t = 0; t < m; ++t) {
            const double at = stencil_row<double>(
                L, i, j, k, c, [&](int qq) { return (double)cs_phi(cs, t, qq, msk && msk[qq]); }, (double*)nullptr);
            for (int q = 0; q < m; ++q) e[q * m + t] += (double)cs_phi(cs, q, c, false) * at;
        }
    }
    for (int q = 0; q < m * m; ++q) store_block_partial(e[q], partials + (size_t)q * gridDim.x);
}

// The same Gram matrix in the column-march shape, the coarse space's shape
// fixed at compile time: column t's values at the cell and its neighbours
// are loaded together (phi_t = 1 on free cells for the indicator column).
template <int FREE, int NF>
__global__ void __launch_bounds__(TPB) k_cs_gram3(LevelDev L, March M, CoarseSpace cs, double* __restrict__ partials) {
    constexpr int m = FREE + NF;
    double e[CS_MAX * CS_MAX];
#pragma unroll
    for (int q = 0; q < m * m; ++q) e[q] = 0.0;
    int t, k0, k1;
    if (march_begin(M, t, k0, k1)) {
        const int n0 = L.n0, n1 = L.n1, n2 = L.n2;
        const int plane = n0 * n1;
        const int j = t / n0;
        const int i = t - j * n0;
        const uint8_t* __restrict__ nbits = L.nb;
        const double ih2[3] = {L.ih2d[0], L.ih2d[1], L.ih2d[2]};
        auto phi = [&](int q, int cell) { return (FREE && q == 0) ? 1.0 : (double)cs.f[q - FREE][cell]; };
        for (int k = k0; k < k1; ++k) {
            const int c = i + n0 * (j + n1 * k);
            const unsigned bits = nbits ? nbits[c] : 0u;
            if (bits & 64u) continue;
            float w[6];
            face_w3(L, i, j, k, c, w);
            const int nq[6] = {i > 0 ? c - 1 : c, i < n0 - 1 ? c + 1 : c, j > 0 ? c - n0 : c,
                               j < n1 - 1 ? c + n0 : c, k > 0 ? c - plane : c, k + 1 < n2 ? c + plane : c};
#pragma unroll
            for (int tt = 0; tt < m; ++tt) {
                double xn[6];
#pragma unroll
                for (int a = 0; a < 6; ++a) xn[a] = ((bits >> a) & 1u) ? 0.0 : phi(tt, nq[a]);
                double at, dsum;
                row_terms<double>(phi(tt, c), xn, w, ih2, at, dsum);
#pragma unroll
                for (int q = 0; q < m; ++q) e[q * m + tt] += phi(q, c) * at;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < m * m; ++q) store_block_partial(e[q], partials + (size_t)q * gridDim.x);
}

// c0 = Phi^T r0 partials (the setup residual), then a0 in k_fin_cs
__global__ void k_cs_dots(const uint8_t* __restrict__ msk, CoarseSpace cs, const float* __restrict__ r, int n,
                          double* __restrict__ partials) {
    const int m = cs.m();
    double vc[CS_MAX] = {0, 0, 0, 0};
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        const bool dir = msk && msk[c];
        for (int q = 0; q < m; ++q) vc[q] += (double)r[c] * (double)cs_phi(cs, q, c, dir);
    }
    for (int q = 0; q < m; ++q) store_block_partial(vc[q], partials + (size_t)q * gridDim.x);
}

__global__ void __launch_bounds__(FIN_TPB) k_fin_cs(Scal* S, const double* partials, unsigned nb) {
    double c[CS_MAX];
    for (int q = 0; q < S->cs_m; ++q) c[q] = sum_partials(partials + (size_t)q * nb, nb);
    if (threadIdx.x != 0) return;
    for (int q = 0; q < S->cs_m; ++q) S->cs_c[q] = c[q];
    cs_solve(S, false);
}

__global__ void k_pcg_copy_out(const double* __restrict__ xa, const double* __restrict__ xb, float* __restrict__ x32,
                               int n, const Scal* S) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) x32[c] = (float)(S->xsel ? xb[c] : xa[c]);
}

// ------------------------------------------------------------ host objects

static int launch_fill(void* p, int dtype, int64_t n, double v) {
    if (n <= 0) return FLUX_OK;
    unsigned g = blocks_for(n, TPB);
    switch (dtype) {
        case FLUX_F32: k_fill_f32<<<g, TPB, 0, g_stream>>>((float*)p, (int)n, (float)v); break;
        case FLUX_F64: k_fill_f64<<<g, TPB, 0, g_stream>>>((double*)p, (int)n, v); break;
        case FLUX_I32: k_fill_i32<<<g, TPB, 0, g_stream>>>((int*)p, (int)n, (int)v); break;
        case FLUX_U8: k_fill_u8<<<g, TPB, 0, g_stream>>>((uint8_t*)p, (int)n, (uint8_t)v); break;
        case FLUX_VEC3: k_fill_f32<<<blocks_for(3 * n, TPB), TPB, 0, g_stream>>>((float*)p, (int)(3 * n), (float)v); break;
        default: return set_error(FLUX_ERR_ARG, "fill: unsupported element type %d", dtype);
    }
    FLUX_LAUNCH_CHECK();
    return FLUX_OK;
}

int fill_device(void* p, int dtype, int64_t n, double v) { return launch_fill(p, dtype, n, v); }

int check_shape(int ndim, const int* shape, const double* spacing, int spacing_len, const char* who) {
    if (ndim < 1 || ndim > 3) return set_error(FLUX_ERR_ARG, "%s: only 1D/2D/3D shapes are supported, got %dD.", who, ndim);
    if (ndim == 1) return set_error(FLUX_ERR_ARG, "%s: only 2D/3D shapes are supported by this build, got 1D.", who);
    if (spacing_len != ndim)
        return set_error(FLUX_ERR_ARG, "%s: gridSpacing length (%d) must match shape length (%d).", who, spacing_len, ndim);
    int64_t cells = 1;
    for (int a = 0; a < ndim; ++a) {
        if (shape[a] <= 0) return set_error(FLUX_ERR_ARG, "%s: shape[%d] must be positive", who, a);
        if (!(spacing[a] > 0)) return set_error(FLUX_ERR_ARG, "%s: gridSpacing[%d] must be positive", who, a);
        cells *= shape[a];
    }
    if (cells >= (1ll << 31)) return set_error(FLUX_ERR_ARG, "%s: %lld cells exceed the 2^31 index range", who, (long long)cells);
    return FLUX_OK;
}

}  // namespace flux

using namespace flux;

struct flux_operator {
    int ndim = 3;
    int n[3] = {1, 1, 1};
    double h[3] = {1, 1, 1};
    const flux_field* mask = nullptr;
    const flux_field* w[3] = {nullptr, nullptr, nullptr};
    LevelDev L;
    uint8_t* nb = nullptr;  // neighbour-mask bits of the column-march kernels (refreshed per solve)
};

namespace flux {
// The operator's neighbour-mask bits from its (caller-owned, possibly
// changed) mask.
static void operator_refresh_bits(flux_operator* op) {
    if (op->nb) k_nbits3<float><<<blocks_for(cells_of(op->n), TPB), TPB, 0, g_stream>>>(op->L, op->nb);
}
}  // namespace flux

extern "C" {

int flux_field_fill(flux_field* f, double value) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(f, "fill: field is NULL");
    FLUX_TRY(launch_fill(f->d, f->dtype, f->count, value));
    FLUX_HIP(hipStreamSynchronize(g_stream));
    return FLUX_OK;
}

int flux_operator_create(int ndim, const int* shape, const double* spacing, const flux_field* mask,
                         const flux_field* const* w, flux_operator** out) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(out && shape && spacing, "createLaplacianOperator: NULL argument");
    FLUX_TRY(check_shape(ndim, shape, spacing, ndim, "createLaplacianOperator"));
    flux_operator* op = new flux_operator();
    op->ndim = ndim;
    for (int a = 0; a < 3; ++a) {
        op->n[a] = a < ndim ? shape[a] : 1;
        op->h[a] = a < ndim ? spacing[a] : 1.0;
    }
    int64_t cells = cells_of(op->n);
    if (mask) {
        int rc = check_field(mask, FLUX_U8, cells, "createLaplacianOperator: dirichletMask");
        if (rc) { delete op; return rc; }
    }
    const float* wp[3] = {nullptr, nullptr, nullptr};
    if (w) {
        for (int a = 0; a < ndim; ++a) {
            int rc = check_field(w[a], FLUX_F32, faces_of(op->n, a), "createLaplacianOperator: faceWeights");
            if (rc) { delete op; return rc; }
            op->w[a] = w[a];
            wp[a] = (const float*)w[a]->d;
        }
    }
    op->mask = mask;
    op->L = make_level(ndim, op->n, op->h, mask ? (const uint8_t*)mask->d : nullptr, w ? wp : nullptr);
    if (mask) {
        op->nb = (uint8_t*)dmalloc(cells_of(op->n));
        if (!op->nb) {
            delete op;
            return set_error(FLUX_ERR_NOMEM, "createLaplacianOperator: allocation failed");
        }
        op->L.nb = op->nb;
    }
    *out = op;
    return FLUX_OK;
}

int flux_operator_apply(flux_operator* op, const flux_field* in, flux_field* out) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(op, "applyLaplacian: operator is NULL");
    int64_t cells = cells_of(op->n);
    FLUX_TRY(check_field(in, FLUX_F32, cells, "applyLaplacian: input"));
    FLUX_TRY(check_field(out, FLUX_F32, cells, "applyLaplacian: output"));
    kprof_begin("apply_op");
    k_apply<<<blocks_for(cells, TPB), TPB, 0, g_stream>>>(op->L, (const float*)in->d, (float*)out->d);
    kprof_end("apply_op");
    FLUX_LAUNCH_CHECK();
    FLUX_HIP(hipStreamSynchronize(g_stream));
    return FLUX_OK;
}

int flux_operator_diagonal(flux_operator* op, flux_field* out) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(op, "diagonal: operator is NULL");
    int64_t cells = cells_of(op->n);
    FLUX_TRY(check_field(out, FLUX_F32, cells, "diagonal: output"));
    k_diag<<<blocks_for(cells, TPB), TPB, 0, g_stream>>>(op->L, (float*)out->d);
    FLUX_LAUNCH_CHECK();
    FLUX_HIP(hipStreamSynchronize(g_stream));
    return FLUX_OK;
}

int flux_operator_destroy(flux_operator* op) {
    if (!op) return FLUX_OK;
    if (g_init) (void)hipStreamSynchronize(g_stream);
    dfree(op->nb);
    delete op;
    return FLUX_OK;
}

}  // extern "C"

// -------------------------------------------------------------------- PCG
struct flux_pcg {
    flux_operator* op = nullptr;
    flux_multigrid* mg = nullptr;
    int precond = FLUX_PREC_MULTIGRID;
    flux_field* b = nullptr;
    flux_field* x = nullptr;
    int64_t n = 0;
    float *r = nullptr, *z = nullptr, *pa = nullptr, *pb = nullptr;
    double *xa = nullptr, *xb = nullptr;
    Scal* S = nullptr;
    Scal* Sh = nullptr;  // pinned host mirror
    double* partials = nullptr;
    unsigned n_partials = 0;
    March M;  // column-march grid of the iteration kernels
    int last_iterations = 4;
    int flexible = 0;  // Polak-Ribiere beta
    CoarseSpace cs;    // additive coarse-space correction of the preconditioner
};

namespace flux {

// The grid-stride reductions (setup, b.b, r.z) run on a capped grid: the
// block count is the number of partials the finishing kernel adds.
static constexpr unsigned RED_MAX_BLOCKS = 2048;
static unsigned red_blocks(int64_t n) { return std::min<unsigned>(blocks_for(n, TPB), RED_MAX_BLOCKS); }

// z = M r.  With `dot`, the multigrid V-cycle may also produce the r.z
// partials (dot->blocks > 0).
static int pcg_precondition(flux_pcg* s, DotOut* dot) {
    const int* stop = &s->S->stop;
    unsigned g = blocks_for(s->n, TPB);
    if (dot) dot->blocks = 0;
    if (s->precond == FLUX_PREC_MULTIGRID) return mg_apply_raw(s->mg, s->r, s->z, stop, &s->S->rr, dot);
    if (s->precond == FLUX_PREC_JACOBI) k_jacobi<<<g, TPB, 0, g_stream>>>(s->op->L, s->r, s->z, stop);
    else k_copy_f32<<<g, TPB, 0, g_stream>>>(s->r, s->z, (int)s->n, stop);
    return FLUX_OK;
}

static int pcg_iteration(flux_pcg* s) {
    unsigned g = red_blocks(s->n);
    kprof_begin("apply");
#define FLUX_CS_SHAPE(KERNEL, ...)                                                                      \
    switch (s->cs.free0 * 4 + s->cs.nf) {                                                               \
        case 0: KERNEL<0, 0><<<s->M.nb, TPB, 0, g_stream>>>(__VA_ARGS__); break;                           \
        case 1: KERNEL<0, 1><<<s->M.nb, TPB, 0, g_stream>>>(__VA_ARGS__); break;                           \
        case 2: KERNEL<0, 2><<<s->M.nb, TPB, 0, g_stream>>>(__VA_ARGS__); break;                           \
        case 3: KERNEL<0, 3><<<s->M.nb, TPB, 0, g_stream>>>(__VA_ARGS__); break;                           \
        case 4: KERNEL<1, 0><<<s->M.nb, TPB, 0, g_stream>>>(__VA_ARGS__); break;                           \
        case 5: KERNEL<1, 1><<<s->M.nb, TPB, 0, g_stream>>>(__VA_ARGS__); break;                           \
        case 6: KERNEL<1, 2><<<s->M.nb, TPB, 0, g_stream>>>(__VA_ARGS__); break;                           \
        default: KERNEL<1, 3><<<s->M.nb, TPB, 0, g_stream>>>(__VA_ARGS__); break;                          \
    }
    FLUX_CS_SHAPE(k_pcg_apply3, s->op->L, flip_march(s->M), s->z, s->pa, s->pb, s->S, s->cs, s->partials);
    kprof_end("apply");
    k_fin_pap<<<1, FIN_TPB, 0, g_stream>>>(s->S, s->partials, s->M.nb);
    kprof_begin("update");
    FLUX_CS_SHAPE(k_pcg_update3, s->op->L, flip_march(s->M), s->xa, s->xb, s->pa, s->pb, (const float*)s->b->d,
                  s->r, s->flexible ? s->z : nullptr, s->S, s->cs, s->partials);
#undef FLUX_CS_SHAPE
    kprof_end("update");
    k_fin_rr<<<1, FIN_TPB, 0, g_stream>>>(s->S, s->partials, s->M.nb);
    DotOut dot;
    dot.partials = s->partials;
    dot.capacity = s->n_partials;
    FLUX_TRY(pcg_precondition(s, &dot));
    if (dot.blocks == 0) {
        kprof_begin("dot");
        k_pcg_dot_rz<<<g, TPB, 0, g_stream>>>(s->r, s->z, (int)s->n, s->S, s->partials);
        kprof_end("dot");
        dot.blocks = g;
    }
    k_fin_rz<<<1, FIN_TPB, 0, g_stream>>>(s->S, s->partials, dot.blocks, 0);
    FLUX_LAUNCH_CHECK();
    return FLUX_OK;
}

// (-E)^+ for the symmetric m x m matrix -E (m <= 4) by cyclic Jacobi
// rotations: eigenvalues below 1e-12 of the largest are dropped, so a
// degenerate or repeated column contributes nothing.  Returns -(-E)^+.
static void coarse_inverse(const double* E, int m, double* out) {
    double a[CS_MAX][CS_MAX], v[CS_MAX][CS_MAX];
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
            a[i][j] = -0.5 * (E[i * m + j] + E[j * m + i]);
            v[i][j] = i == j ? 1.0 : 0.0;
        }
    for (int sweep = 0; sweep < 50; ++sweep) {
        double off = 0;
        for (int i = 0; i < m; ++i)
            for (int j = i + 1; j < m; ++j) off += a[i][j] * a[i][j];
        if (off < 1e-30) break;
        for (int pq = 0; pq < m; ++pq)
            for (int q = pq + 1; q < m; ++q) {
                if (a[pq][q] == 0.0) continue;
                const double th = (a[q][q] - a[pq][pq]) / (2.0 * a[pq][q]);
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
                for (int k = 0; k < m; ++k) {
                    const double akp = a[k][pq], akq = a[k][q];
                    a[k][pq] = c * akp - sn * akq;
                    a[k][q] = sn * akp + c * akq;
                }
                for (int k = 0; k < m; ++k) {
                    const double apk = a[pq][k], aqk = a[q][k];
                    a[pq][k] = c * apk - sn * aqk;
                    a[q][k] = sn * apk + c * aqk;
                }
                for (int k = 0; k < m; ++k) {
                    const double vkp = v[k][pq], vkq = v[k][q];
                    v[k][pq] = c * vkp - sn * vkq;
                    v[k][q] = sn * vkp + c * vkq;
                }
            }
    }
    double lmax = 0;
    for (int i = 0; i < m; ++i) lmax = std::max(lmax, a[i][i]);
    for (int i = 0; i < CS_MAX * CS_MAX; ++i) out[i] = 0.0;
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
            double sum = 0;
            for (int k = 0; k < m; ++k)
                if (a[k][k] > 1e-12 * lmax && std::isfinite(a[k][k])) sum += v[i][k] * v[j][k] / a[k][k];
            out[i * CS_MAX + j] = -sum;
        }
}

// The Gram matrix E = Phi^T A Phi of the coarse space, reduced on the host
// in a fixed order, and its inverse into the scalar block.
int coarse_space_setup(flux_pcg* s, Scal* Sh) {
    const int m = s->cs.m();
    Sh->cs_m = m;
    if (!m) return FLUX_OK;
    const unsigned g = s->M.nb;
    switch (s->cs.free0 * 4 + s->cs.nf) {
        case 1: k_cs_gram3<0, 1><<<g, TPB, 0, g_stream>>>(s->op->L, s->M, s->cs, s->partials); break;
        case 2: k_cs_gram3<0, 2><<<g, TPB, 0, g_stream>>>(s->op->L, s->M, s->cs, s->partials); break;
        case 3: k_cs_gram3<0, 3><<<g, TPB, 0, g_stream>>>(s->op->L, s->M, s->cs, s->partials); break;
        case 4: k_cs_gram3<1, 0><<<g, TPB, 0, g_stream>>>(s->op->L, s->M, s->cs, s->partials); break;
        case 5: k_cs_gram3<1, 1><<<g, TPB, 0, g_stream>>>(s->op->L, s->M, s->cs, s->partials); break;
        case 6: k_cs_gram3<1, 2><<<g, TPB, 0, g_stream>>>(s->op->L, s->M, s->cs, s->partials); break;
        default: k_cs_gram3<1, 3><<<g, TPB, 0, g_stream>>>(s->op->L, s->M, s->cs, s->partials); break;
    }
    FLUX_LAUNCH_CHECK();
    std::vector<double> h((size_t)m * m * g);
    FLUX_HIP(hipMemcpyAsync(h.data(), s->partials, sizeof(double) * h.size(), hipMemcpyDeviceToHost, g_stream));
    FLUX_HIP(hipStreamSynchronize(g_stream));
    double E[CS_MAX * CS_MAX];
    for (int q = 0; q < m * m; ++q) {
        double acc = 0;
        for (unsigned b = 0; b < g; ++b) acc += h[(size_t)q * g + b];
        E[q] = acc;
    }
    coarse_inverse(E, m, Sh->cs_einv);
    return FLUX_OK;
}

int pcg_solve_raw(flux_pcg* s, double tol, int maxiter, int relative, flux_pcg_state* st) {
    unsigned g = red_blocks(s->n);
    unsigned gfull = blocks_for(s->n, TPB);
    g_march_toggle = 0;
    operator_refresh_bits(s->op);
    // host-written fields of the scalar block
    memset(s->Sh, 0, sizeof(Scal));
    s->Sh->tol = tol;
    s->Sh->relative = relative ? 1 : 0;
    s->Sh->flexible = s->flexible;
    FLUX_TRY(coarse_space_setup(s, s->Sh));
    FLUX_HIP(hipMemcpyAsync(s->S, s->Sh, sizeof(Scal), hipMemcpyHostToDevice, g_stream));
    FLUX_HIP(hipStreamSynchronize(g_stream));
    // setup (runGpuResidentSetup, linalg.js:1613-1637)
    k_pcg_setup_residual3<<<s->M.nb, TPB, 0, g_stream>>>(s->op->L, s->M, (const float*)s->x->d, s->xa,
                                                          (const float*)s->b->d, s->r, s->partials);
    k_fin_setup<<<1, FIN_TPB, 0, g_stream>>>(s->S, s->partials, s->M.nb);
    k_pcg_dot_bb<<<g, TPB, 0, g_stream>>>((const float*)s->b->d, (int)s->n, s->partials);
    k_fin_bb<<<1, FIN_TPB, 0, g_stream>>>(s->S, s->partials, g);
    if (s->cs.m()) {
        k_cs_dots<<<g, TPB, 0, g_stream>>>(s->op->L.mask, s->cs, s->r, (int)s->n, s->partials);
        k_fin_cs<<<1, FIN_TPB, 0, g_stream>>>(s->S, s->partials, g);
    }
    FLUX_TRY(pcg_precondition(s, nullptr));
    k_pcg_dot_rz<<<g, TPB, 0, g_stream>>>(s->r, s->z, (int)s->n, s->S, s->partials);
    k_fin_rz<<<1, FIN_TPB, 0, g_stream>>>(s->S, s->partials, g, 1);
    k_fill_f32<<<gfull, TPB, 0, g_stream>>>(s->pa, (int)s->n, 0.0f);
    k_fill_f32<<<gfull, TPB, 0, g_stream>>>(s->pb, (int)s->n, 0.0f);
    FLUX_LAUNCH_CHECK();
    // chunked self-stopping loop (linalg.js:1854-1905): one scalar read per chunk
    int ran = 0;
    FLUX_HIP(hipMemcpyAsync(s->Sh, s->S, sizeof(Scal), hipMemcpyDeviceToHost, g_stream));
    FLUX_HIP(hipStreamSynchronize(g_stream));
    while (s->Sh->stop == 0 && ran < maxiter) {
        int target = std::max(4, s->last_iterations + 2);
        int chunk = std::min(std::max(target - ran, 4), maxiter - ran);
        for (int q = 0; q < chunk; ++q) FLUX_TRY(pcg_iteration(s));
        ran += chunk;
        FLUX_HIP(hipMemcpyAsync(s->Sh, s->S, sizeof(Scal), hipMemcpyDeviceToHost, g_stream));
        FLUX_HIP(hipStreamSynchronize(g_stream));
    }
    k_pcg_copy_out<<<gfull, TPB, 0, g_stream>>>(s->xa, s->xb, (float*)s->x->d, (int)s->n, s->S);
    FLUX_LAUNCH_CHECK();
    FLUX_HIP(hipStreamSynchronize(g_stream));
    const Scal& h = *s->Sh;
    st->stopped_by = h.stop;
    st->converged = (h.stop == 5);
    st->iterations = st->converged ? h.stop_iter : h.iter;
    st->residual = sqrt(fabs(h.rr));
    st->threshold = h.thr;
    st->b_norm = sqrt(fabs(h.bb));
    st->restarts = h.restarts;
    s->last_iterations = st->iterations;
    return FLUX_OK;
}

}  // namespace flux

extern "C" {

int flux_pcg_create(flux_operator* op, int precond, flux_multigrid* mg, flux_field* b, flux_field* x, flux_pcg** out) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(op && b && x && out, "conjugateGradient: NULL argument");
    FLUX_REQUIRE(b->dtype == x->dtype, "conjugateGradient: element type mismatch, b.type(%d) != x.type(%d).", b->dtype,
                 x->dtype);
    FLUX_REQUIRE(b->count == x->count, "conjugateGradient: shape mismatch, b has %lld elements, x has %lld.",
                 (long long)b->count, (long long)x->count);
    FLUX_REQUIRE(b->dtype == FLUX_F32, "conjugateGradient: GPU dot product only supports type \"float\", got %d.", b->dtype);
    int64_t cells = cells_of(op->n);
    FLUX_REQUIRE(b->count == cells, "conjugateGradient: shape mismatch with the operator (%lld vs %lld)",
                 (long long)b->count, (long long)cells);
    FLUX_REQUIRE(precond >= 0 && precond <= 2, "conjugateGradient: unknown preconditioner %d", precond);
    FLUX_REQUIRE(precond != FLUX_PREC_MULTIGRID || mg, "conjugateGradient: multigrid preconditioner requires mg");
    if (mg) FLUX_REQUIRE(mg_cells(mg) == cells, "conjugateGradient: multigrid shape mismatch");
    flux_pcg* s = new flux_pcg();
    s->op = op;
    s->mg = mg;
    s->precond = precond;
    s->b = b;
    s->x = x;
    s->n = cells;
    s->r = (float*)dmalloc(cells * 4);
    s->z = (float*)dmalloc(cells * 4);
    s->pa = (float*)dmalloc(cells * 4);
    s->pb = (float*)dmalloc(cells * 4);
    s->xa = (double*)dmalloc(cells * 8);
    s->xb = (double*)dmalloc(cells * 8);
    s->S = (Scal*)dmalloc(sizeof(Scal));
    s->M = march_shape(op->n[0], op->n[1], op->n[2], TPB, g_march_blocks);
    // partials of every reduction: the capped grid-stride grids, the column-march
    // grids of the iteration kernels and of the V-cycle's fused r.z (a march
    // grid over cell pairs; larger grids fall back to the separate dot kernel)
    // (2 + CS_MAX) arrays of the update's partials; CS_MAX^2 of the coarse
    // space's Gram matrix
    s->n_partials = (unsigned)std::max<int64_t>(
        (int64_t)CS_MAX * CS_MAX * std::max<int64_t>(RED_MAX_BLOCKS, (int64_t)s->M.nb), 4096);
    // the reference's beta unless asked otherwise (FLUX_FLEXIBLE_PCG=1 or
    // flux_pcg_set_flexible): measured on the 64^3 dam break with FP16
    // finest-level storage, the flexible beta took more iterations (63) than
    // the reference's (51)
    s->flexible = 0;
    if (const char* e = getenv("FLUX_FLEXIBLE_PCG")) s->flexible = e[0] == '1' ? 1 : 0;
    s->partials = (double*)dmalloc(sizeof(double) * s->n_partials);
    void* hp = nullptr;
    if (hipHostMalloc(&hp, sizeof(Scal), hipHostMallocDefault) != hipSuccess) hp = nullptr;
    s->Sh = (Scal*)hp;
    if (!s->r || !s->z || !s->pa || !s->pb || !s->xa || !s->xb || !s->S || !s->partials || !s->Sh) {
        flux_pcg_destroy(s);
        return set_error(FLUX_ERR_NOMEM, "conjugateGradient: allocation failed");
    }
    // the scalar block starts at zero
    FLUX_HIP(hipMemsetAsync(s->S, 0, sizeof(Scal), g_stream));
    FLUX_HIP(hipStreamSynchronize(g_stream));
    *out = s;
    return FLUX_OK;
}

int flux_pcg_scaled_start(flux_pcg* s, const flux_field* g, double* alpha_out) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(s && g, "scaledStart: NULL argument");
    FLUX_TRY(check_field(g, FLUX_F32, s->n, "scaledStart: direction"));
    const unsigned nb = red_blocks(s->n);
    k_galerkin_terms<<<nb, TPB, 0, g_stream>>>(s->op->L, (const float*)s->x->d, (const float*)s->b->d,
                                               (const float*)g->d, s->partials);
    FLUX_LAUNCH_CHECK();
    std::vector<double> h(2 * (size_t)nb);
    FLUX_HIP(hipMemcpyAsync(h.data(), s->partials, sizeof(double) * h.size(), hipMemcpyDeviceToHost, g_stream));
    FLUX_HIP(hipStreamSynchronize(g_stream));
    double gr = 0, gag = 0;
    for (unsigned q = 0; q < nb; ++q) {
        gr += h[q];
        gag += h[nb + q];
    }
    // A is negative definite on the free cells: gag < 0 for any g != 0
    double alpha = (gag < 0 && std::isfinite(gr) && std::isfinite(gag)) ? gr / gag : 0.0;
    if (!std::isfinite(alpha)) alpha = 0.0;
    if (alpha != 0.0) {
        k_galerkin_step<<<blocks_for(s->n, TPB), TPB, 0, g_stream>>>(s->op->L.mask, (const float*)g->d,
                                                                     (float*)s->x->d, alpha, (int)s->n);
        FLUX_LAUNCH_CHECK();
        FLUX_HIP(hipStreamSynchronize(g_stream));
    }
    if (alpha_out) *alpha_out = alpha;
    return FLUX_OK;
}

int flux_pcg_set_coarse_space(flux_pcg* s, int free_indicator, int nfields, const flux_field* const* fields) {
    FLUX_REQUIRE(s, "conjugateGradient: NULL solver");
    FLUX_REQUIRE(nfields >= 0 && nfields <= 3 && (nfields == 0 || fields),
                 "coarseSpace: 0 to 3 fields, got %d", nfields);
    CoarseSpace cs;
    cs.free0 = free_indicator ? 1 : 0;
    cs.nf = nfields;
    for (int q = 0; q < nfields; ++q) {
        FLUX_TRY(check_field(fields[q], FLUX_F32, s->n, "coarseSpace: vector"));
        cs.f[q] = (const float*)fields[q]->d;
    }
    s->cs = cs;
    return FLUX_OK;
}

int flux_pcg_set_flexible(flux_pcg* s, int flexible) {
    FLUX_REQUIRE(s, "conjugateGradient: NULL solver");
    s->flexible = flexible ? 1 : 0;
    return FLUX_OK;
}

int flux_pcg_solve(flux_pcg* s, double tol, int maxiter, int relative, flux_pcg_state* st) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(s && st, "solve: NULL argument");
    FLUX_REQUIRE(maxiter >= 0, "solve: maxiter must be >= 0");
    return pcg_solve_raw(s, tol, maxiter, relative, st);
}

int flux_pcg_destroy(flux_pcg* s) {
    if (!s) return FLUX_OK;
    if (g_init) (void)hipStreamSynchronize(g_stream);
    dfree(s->r);
    dfree(s->z);
    dfree(s->pa);
    dfree(s->pb);
    dfree(s->xa);
    dfree(s->xb);
    dfree(s->S);
    dfree(s->partials);
    if (s->Sh) (void)hipHostFree(s->Sh);
    delete s;
    return FLUX_OK;
}

}  // extern "C"
