// Operator, multigrid preconditioner and GPU-resident PCG for gfx950.
//
// Reference: packages/fluxflow/src/linalg/multigrid.js (operator, V-cycle)
// and linalg.js (PCG).  All arithmetic of the V-cycle and of the search
// directions is FP32; the solution accumulator and the true residual are
// FP64, and every dot product is reduced in FP64 in a fixed order.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "flux_internal.h"
#include "flux_solver.h"
#include "flux_stencil.h"

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

// One red-black Gauss-Seidel colour (buildRelaxKernel, multigrid.js:626-640):
// x_c += ((b_c - (A x)_c) / diag_c) * omega on cells with (i+j+k)%2 == color.
// Thread t covers the t-th cell of that colour.  With `from_zero` the field
// is taken as zero on entry and every cell is written (the colour-0 cells
// with omega*b/diag, the others with 0), which folds zeroX into the sweep.
// FINE marks launches on the finest level and FROM_ZERO the first sweep
// from a zero field; both are template parameters so that profilers report
// the finest-level launches as kernels of their own.
template <int FINE, int FROM_ZERO>
__global__ void k_relax(LevelDev L, float* __restrict__ x, const float* __restrict__ b, float omega, int color,
                        const int* __restrict__ stop) {
    constexpr bool from_zero = FROM_ZERO != 0;
    if (stop && *stop) return;
    const int half = (L.n0 + 1) >> 1;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int rows = L.n1 * L.n2;
    if (t >= half * rows) return;
    int row = t / half;
    int m = t - row * half;
    int k = row / L.n1;
    int j = row - k * L.n1;
    int i = 2 * m + ((j + k + color) & 1);
    if (from_zero) {
        // the other-colour partner of this pair is cleared here too
        int ip = 2 * m + ((j + k + color + 1) & 1);
        if (ip < L.n0) x[ip + L.n0 * row] = 0.0f;
    }
    if (i >= L.n0) return;
    int c = i + L.n0 * row;
    float d;
    float ax;
    if (from_zero) {
        (void)stencil_row<float>(L, i, j, k, c, [&](int) { return 0.0f; }, &d);
        ax = 0.0f;
    } else {
        ax = stencil_row<float>(L, i, j, k, c, [&](int q) { return x[q]; }, &d);
    }
    float cur = from_zero ? 0.0f : x[c];
    x[c] = cur + ((b[c] - ax) / d) * omega;
}

// r = b - A x  (buildResidualKernel, multigrid.js:741-749)
template <int FINE>
__global__ void k_residual(LevelDev L, const float* __restrict__ x, const float* __restrict__ b,
                           float* __restrict__ r, const int* __restrict__ stop) {
    if (stop && *stop) return;
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= L.n0 * L.n1 * L.n2) return;
    int i, j, k;
    cell_coords(L, c, i, j, k);
    r[c] = b[c] - stencil_row<float>(L, i, j, k, c, [&](int q) { return x[q]; }, (float*)nullptr);
}

// Full-weighting restriction (buildRestrictKernel, multigrid.js:866-914):
// per axis taps 1/8,3/8,3/8,1/8 at 2c-1 (clamped to 2c), 2c, 2c+1, 2c+2
// (clamped to 2c+1); masked fine cells contribute 0; masked coarse cells get 0.
// Also clears the coarse x when `xc` is non-null (foldClearIntoRestrict).
template <int FINE>
__global__ void k_restrict(int ndim, int f0, int f1, int f2, const float* __restrict__ rf,
                           const uint8_t* __restrict__ mf, int c0, int c1, int c2, float* __restrict__ bc,
                           const uint8_t* __restrict__ mc, float* __restrict__ xc, const int* __restrict__ stop) {
    if (stop && *stop) return;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= c0 * c1 * c2) return;
    int row = t / c0;
    int I = t - row * c0;
    int K = row / c1;
    int J = row - K * c1;
    if (xc) xc[t] = 0.0f;
    if (mc && mc[t]) {
        bc[t] = 0.0f;
        return;
    }
    const float tw[4] = {0.125f, 0.375f, 0.375f, 0.125f};
    int ti[4], tj[4], tk[4];
    ti[0] = I > 0 ? 2 * I - 1 : 2 * I;
    ti[1] = 2 * I;
    ti[2] = 2 * I + 1;
    ti[3] = (I + 1 < c0) ? 2 * I + 2 : 2 * I + 1;
    tj[0] = J > 0 ? 2 * J - 1 : 2 * J;
    tj[1] = 2 * J;
    tj[2] = 2 * J + 1;
    tj[3] = (J + 1 < c1) ? 2 * J + 2 : 2 * J + 1;
    int nk = 1;
    tk[0] = 0;
    if (ndim == 3) {
        nk = 4;
        tk[0] = K > 0 ? 2 * K - 1 : 2 * K;
        tk[1] = 2 * K;
        tk[2] = 2 * K + 1;
        tk[3] = (K + 1 < c2) ? 2 * K + 2 : 2 * K + 1;
    }
    float sum = 0.0f;
    for (int a = 0; a < 4; ++a)
        for (int bq = 0; bq < 4; ++bq)
            for (int cq = 0; cq < nk; ++cq) {
                float wgt = tw[a] * tw[bq] * (nk == 4 ? tw[cq] : 1.0f);
                int f = ti[a] + f0 * (tj[bq] + f1 * tk[cq]);
                float v = (mf && mf[f]) ? 0.0f : rf[f];
                sum += v * wgt;
            }
    bc[t] = sum;
}

// Bilinear / trilinear prolongation added to the fine x except at masked fine
// cells (buildCorrectKernel, multigrid.js:990-1040, taps 921-934).
template <int FINE>
__global__ void k_correct(int ndim, int c0, int c1, int c2, const float* __restrict__ xc, int f0, int f1, int f2,
                          float* __restrict__ xf, const uint8_t* __restrict__ mf, const int* __restrict__ stop) {
    if (stop && *stop) return;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= f0 * f1 * f2) return;
    if (mf && mf[t]) return;
    int row = t / f0;
    int i = t - row * f0;
    int k = row / f1;
    int j = row - k * f1;
    int F[3] = {i, j, k};
    int nc[3] = {c0, c1, c2};
    int ti[3][2];
    float tw[3][2];
    for (int a = 0; a < 3; ++a) {
        if (a < ndim) {
            int ci = F[a] >> 1;
            bool even = (F[a] & 1) == 0;
            int lower = ci > 0 ? ci - 1 : ci;
            int upper = ci < nc[a] - 1 ? ci + 1 : ci;
            ti[a][0] = even ? lower : ci;
            tw[a][0] = even ? 0.25f : 0.75f;
            ti[a][1] = even ? ci : upper;
            tw[a][1] = even ? 0.75f : 0.25f;
        } else {
            ti[a][0] = ti[a][1] = 0;
            tw[a][0] = 1.0f;
            tw[a][1] = 0.0f;
        }
    }
    int nk = (ndim == 3) ? 2 : 1;
    float sum = 0.0f;
    for (int a = 0; a < 2; ++a)
        for (int bq = 0; bq < 2; ++bq)
            for (int cq = 0; cq < nk; ++cq) {
                float wgt = tw[0][a] * tw[1][bq] * (nk == 2 ? tw[2][cq] : 1.0f);
                sum += xc[ti[0][a] + c0 * (ti[1][bq] + c1 * ti[2][cq])] * wgt;
            }
    xf[t] += sum;
}

// Coarsest level in one workgroup: `iters` x (colour 0, colour 1) sweeps from
// x = 0 (buildCoarseSweepKernel, multigrid.js:679-737).  Block-wide barriers
// order the colours; global writes are visible within the block after them.
__global__ void __launch_bounds__(1024) k_coarse_solve(LevelDev L, float* __restrict__ x, const float* __restrict__ b,
                                                       float omega, int iters, const int* __restrict__ stop) {
    if (stop && *stop) return;
    const int n = L.n0 * L.n1 * L.n2;
    for (int c = threadIdx.x; c < n; c += blockDim.x) x[c] = 0.0f;
    __syncthreads();
    for (int it = 0; it < iters; ++it) {
        for (int color = 0; color < 2; ++color) {
            for (int c = threadIdx.x; c < n; c += blockDim.x) {
                int i, j, k;
                cell_coords(L, c, i, j, k);
                if (((i + j + k) & 1) != color) continue;
                float d;
                float ax = stencil_row<float>(L, i, j, k, c, [&](int q) { return x[q]; }, &d);
                x[c] = x[c] + ((b[c] - ax) / d) * omega;
            }
            __syncthreads();
        }
    }
}

// Coarse face weight = mean of the 2^(d-1) fine faces at index 2I along the
// face axis (buildFaceWeightCoarsenKernel, multigrid.js:1138-1184).
__global__ void k_coarsen_w(int ndim, int axis, int f0, int f1, int f2, const float* __restrict__ wf, int c0, int c1,
                            int c2, float* __restrict__ wc) {
    // coarse face array: (c0 + (axis==0)) x (c1 + (axis==1)) x (c2 + (axis==2))
    int s0 = c0 + (axis == 0), s1 = c1 + (axis == 1), s2 = c2 + (axis == 2);
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= s0 * s1 * s2) return;
    int row = t / s0;
    int I = t - row * s0;
    int K = row / s1;
    int J = row - K * s1;
    int fs0 = f0 + (axis == 0), fs1 = f1 + (axis == 1);
    int C[3] = {I, J, K};
    int cnt[3];
    for (int a = 0; a < 3; ++a) cnt[a] = (a < ndim && a != axis) ? 2 : 1;
    float sum = 0.0f;
    for (int d2 = 0; d2 < cnt[2]; ++d2)
        for (int d1 = 0; d1 < cnt[1]; ++d1)
            for (int d0 = 0; d0 < cnt[0]; ++d0) {
                int fi = 2 * C[0] + d0, fj = 2 * C[1] + d1, fk = (ndim == 3) ? 2 * C[2] + d2 : 0;
                sum += wf[fi + fs0 * (fj + fs1 * fk)];
            }
    wc[t] = sum / (float)(1 << (ndim - 1));
}

// Coarse mask (buildMaskCoarsenKernel, multigrid.js:1066-1122): ALL 2^d fine
// children of `src` masked, OR the six coarse face weights sum below 1e-6.
__global__ void k_coarsen_mask(int ndim, int f0, int f1, const uint8_t* __restrict__ src, int c0, int c1, int c2,
                               const float* __restrict__ w0, const float* __restrict__ w1,
                               const float* __restrict__ w2, uint8_t* __restrict__ mc) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= c0 * c1 * c2) return;
    int row = t / c0;
    int I = t - row * c0;
    int K = row / c1;
    int J = row - K * c1;
    bool have = false, masked = false;
    if (src) {
        masked = true;
        int nk = (ndim == 3) ? 2 : 1;
        for (int d2 = 0; d2 < nk; ++d2)
            for (int d1 = 0; d1 < 2; ++d1)
                for (int d0 = 0; d0 < 2; ++d0) {
                    int fk = (ndim == 3) ? 2 * K + d2 : 0;
                    masked = masked && src[(2 * I + d0) + f0 * ((2 * J + d1) + f1 * fk)];
                }
        have = true;
    }
    if (w0) {
        float total = w0[I + (c0 + 1) * (J + c1 * K)] + w0[I + 1 + (c0 + 1) * (J + c1 * K)];
        total += w1[I + c0 * (J + (c1 + 1) * K)] + w1[I + c0 * (J + 1 + (c1 + 1) * K)];
        if (ndim == 3) total += w2[t] + w2[t + c0 * c1];
        bool enclosed = total < 1e-6f;
        masked = have ? (masked || enclosed) : enclosed;
        have = true;
    }
    mc[t] = (uint8_t)(have && masked);
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
    int xsel, psel, pad0, pad1;
    unsigned ticket[4];
};

__device__ __forceinline__ bool degenerate(double v) { return !(fabs(v) <= 1.7976931348623157e308) || v == 0.0; }

// Setup: x64 = x; r = b - A x (FP64 arithmetic); rr.
__global__ void k_pcg_setup_residual(LevelDev L, const float* __restrict__ x32, double* __restrict__ x64,
                                     const float* __restrict__ b, float* __restrict__ r, Scal* S, double* partials) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0;
    int n = L.n0 * L.n1 * L.n2;
    if (c < n) {
        int i, j, k;
        cell_coords(L, c, i, j, k);
        double ax = stencil_row<double>(L, i, j, k, c, [&](int q) { return (double)x32[q]; }, (double*)nullptr);
        double rv = (double)b[c] - ax;
        x64[c] = (double)x32[c];
        r[c] = (float)rv;
        v = rv * rv;
    }
    grid_sum_last(v, partials, &S->ticket[0], [&](double tot) {
        S->rr = tot;
        S->rr0 = tot;
        S->xsel = 0;
        S->psel = 0;
        S->iter = 0;
        S->stop_iter = 0;
        S->stop = 0;
    });
}

__global__ void k_pcg_dot_bb(const float* __restrict__ b, int n, Scal* S, double* partials) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0;
    if (c < n) v = (double)b[c] * (double)b[c];
    grid_sum_last(v, partials, &S->ticket[1], [&](double tot) {
        S->bb = tot;
        double nb = sqrt(fabs(tot));
        S->thr = S->relative ? fmax(S->tol * nb, S->tol) : S->tol;
        // host solve(): an initial residual already below the threshold
        // skips the loop and reports convergence (linalg.js:2205)
        if (sqrt(fabs(S->rr)) < S->thr) {
            S->stop = 5;
            S->stop_iter = 0;
        }
    });
}

// rz = r.z; seed (first=1: seedScalarsKernel linalg.js:1238-1244) or beta
// (betaKernel, linalg.js:1198-1226).
__global__ void k_pcg_dot_rz(const float* __restrict__ r, const float* __restrict__ z, int n, Scal* S,
                             double* partials, int first) {
    if (S->stop) return;
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0;
    if (c < n) v = (double)r[c] * (double)z[c];
    grid_sum_last(v, partials, &S->ticket[2], [&](double tot) {
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
        double braw = tot / rzOld;
        bool restart = flip || !(fabs(braw) <= 10.0);  // MAX_BETA_MAGNITUDE
        S->beta = restart ? 0.0 : braw;
        S->rz_old = tot;
    });
}

// Iteration step 1: p = z + beta p (updateP, fused); Ap = A p; pAp; alpha
// with the guards of alphaKernel (linalg.js:1147-1196).
__global__ void k_pcg_apply(LevelDev L, const float* __restrict__ z, float* __restrict__ pa, float* __restrict__ pb,
                            float* __restrict__ Ap, Scal* S, double* partials) {
    if (S->stop) return;
    const float* __restrict__ pold = S->psel ? pb : pa;
    float* __restrict__ pnew = S->psel ? pa : pb;
    const float beta = (float)S->beta;
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    int n = L.n0 * L.n1 * L.n2;
    double v = 0;
    if (c < n) {
        int i, j, k;
        cell_coords(L, c, i, j, k);
        auto ld = [&](int q) { return z[q] + pold[q] * beta; };
        float ap = stencil_row<float>(L, i, j, k, c, ld, (float*)nullptr);
        float pc = ld(c);
        pnew[c] = pc;
        Ap[c] = ap;
        v = (double)pc * (double)ap;
    }
    grid_sum_last(v, partials, &S->ticket[3], [&](double pAp) {
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
    });
}

// Iteration step 2: x += alpha p (FP64 accumulator, ping-pong buffers);
// r = b - A x in FP64 arithmetic (the true residual, recomputed every
// iteration); rr; convergence and residual-growth tests.
__global__ void k_pcg_update(LevelDev L, double* __restrict__ xa, double* __restrict__ xb, const float* __restrict__ pa,
                             const float* __restrict__ pb, const float* __restrict__ b, float* __restrict__ r, Scal* S,
                             double* partials) {
    if (S->stop) return;
    const double* __restrict__ xold = S->xsel ? xb : xa;
    double* __restrict__ xnew = S->xsel ? xa : xb;
    const float* __restrict__ p = S->psel ? pb : pa;  // the p written by k_pcg_apply
    const double alpha = S->alpha;
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    int n = L.n0 * L.n1 * L.n2;
    double v = 0;
    if (c < n) {
        int i, j, k;
        cell_coords(L, c, i, j, k);
        auto ld = [&](int q) { return xold[q] + alpha * (double)p[q]; };
        double ax = stencil_row<double>(L, i, j, k, c, ld, (double*)nullptr);
        xnew[c] = ld(c);
        double rv = (double)b[c] - ax;
        r[c] = (float)rv;
        v = rv * rv;
    }
    grid_sum_last(v, partials, &S->ticket[0], [&](double rr) {
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
    });
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
};

// -------------------------------------------------------------- multigrid
struct MGLevel {
    int n[3];
    double h[3];
    float* x = nullptr;   // level 0: the caller's output
    float* b = nullptr;   // level 0: the caller's input
    float* buf = nullptr; // residual
    uint8_t* mask = nullptr;  // level 0: the caller's mask (not owned)
    float* w[3] = {nullptr, nullptr, nullptr};  // level 0: the caller's (not owned)
    bool own = false;
    LevelDev dev() const {
        LevelDev L = make_level(ndim_, n, h, mask, (w[0] ? w : nullptr));
        return L;
    }
    int ndim_ = 3;
};

struct flux_multigrid {
    int ndim = 3;
    flux_mg_options o;
    std::vector<MGLevel> lv;
    const flux_field* mask = nullptr;
    const flux_field* coarse_mask = nullptr;
    const flux_field* w[3] = {nullptr, nullptr, nullptr};
    bool has_w = false, has_masks = false;
};

static constexpr int COARSE_SINGLE_BLOCK_MAX = 4096;

namespace flux {

int mg_refresh(flux_multigrid* mg) {
    for (size_t l = 1; l < mg->lv.size(); ++l) {
        MGLevel& f = mg->lv[l - 1];
        MGLevel& c = mg->lv[l];
        if (mg->has_w) {
            for (int a = 0; a < mg->ndim; ++a) {
                int64_t nf = faces_of(c.n, a);
                k_coarsen_w<<<blocks_for(nf, TPB), TPB, 0, g_stream>>>(mg->ndim, a, f.n[0], f.n[1], f.n[2], f.w[a], c.n[0],
                                                                         c.n[1], c.n[2], c.w[a]);
            }
        }
        if (!mg->has_masks) continue;
        const uint8_t* src = nullptr;
        if (mg->mask) src = (l == 1) ? (const uint8_t*)(mg->coarse_mask ? mg->coarse_mask->d : mg->mask->d) : f.mask;
        int64_t cc = cells_of(c.n);
        k_coarsen_mask<<<blocks_for(cc, TPB), TPB, 0, g_stream>>>(mg->ndim, f.n[0], f.n[1], src, c.n[0], c.n[1], c.n[2],
                                                                   mg->has_w ? c.w[0] : nullptr,
                                                                   mg->has_w ? c.w[1] : nullptr,
                                                                   mg->has_w ? c.w[2] : nullptr, c.mask);
    }
    FLUX_LAUNCH_CHECK();
    return FLUX_OK;
}

static int relax_sweeps(flux_multigrid* mg, int l, int iters, bool reversed, bool from_zero, const int* stop) {
    MGLevel& v = mg->lv[l];
    LevelDev L = v.dev();
    int64_t work = (int64_t)((v.n[0] + 1) / 2) * v.n[1] * v.n[2];
    unsigned g = blocks_for(work, TPB);
    float om = (float)mg->o.omega;
    for (int it = 0; it < iters; ++it) {
        int c0 = reversed ? 1 : 0, c1 = reversed ? 0 : 1;
        bool fz = from_zero && it == 0;
        if (l == 0) {
            if (fz) {
                k_relax<1, 1><<<g, TPB, 0, g_stream>>>(L, v.x, v.b, om, c0, stop);
            } else {
                kprof_begin("relax");
                k_relax<1, 0><<<g, TPB, 0, g_stream>>>(L, v.x, v.b, om, c0, stop);
                kprof_end("relax");
            }
            kprof_begin("relax");
            k_relax<1, 0><<<g, TPB, 0, g_stream>>>(L, v.x, v.b, om, c1, stop);
            kprof_end("relax");
        } else {
            if (fz) k_relax<0, 1><<<g, TPB, 0, g_stream>>>(L, v.x, v.b, om, c0, stop);
            else k_relax<0, 0><<<g, TPB, 0, g_stream>>>(L, v.x, v.b, om, c0, stop);
            k_relax<0, 0><<<g, TPB, 0, g_stream>>>(L, v.x, v.b, om, c1, stop);
        }
    }
    return FLUX_OK;
}

static int vcycle(flux_multigrid* mg, int l, const int* stop) {
    MGLevel& v = mg->lv[l];
    int nl = (int)mg->lv.size();
    if (l == nl - 1) {
        int64_t cells = cells_of(v.n);
        if (cells <= COARSE_SINGLE_BLOCK_MAX) {
            k_coarse_solve<<<1, 1024, 0, g_stream>>>(v.dev(), v.x, v.b, (float)mg->o.omega, mg->o.n_coarsest, stop);
        } else if (mg->o.n_coarsest > 0) {
            relax_sweeps(mg, l, mg->o.n_coarsest, false, true, stop);
        } else {
            k_fill_f32<<<blocks_for(cells, TPB), TPB, 0, g_stream>>>(v.x, (int)cells, 0.0f);
        }
        return FLUX_OK;
    }
    MGLevel& c = mg->lv[l + 1];
    LevelDev L = v.dev();
    int64_t cells = cells_of(v.n);
    if (mg->o.nu_down > 0) {
        relax_sweeps(mg, l, mg->o.nu_down, false, true, stop);
    } else {
        k_fill_f32<<<blocks_for(cells, TPB), TPB, 0, g_stream>>>(v.x, (int)cells, 0.0f);
    }
    int64_t cc = cells_of(c.n);
    if (l == 0) {
        kprof_begin("residual");
        k_residual<1><<<blocks_for(cells, TPB), TPB, 0, g_stream>>>(L, v.x, v.b, v.buf, stop);
        kprof_end("residual");
        kprof_begin("restrict");
        k_restrict<1><<<blocks_for(cc, TPB), TPB, 0, g_stream>>>(mg->ndim, v.n[0], v.n[1], v.n[2], v.buf, v.mask, c.n[0],
                                                                   c.n[1], c.n[2], c.b, c.mask, nullptr, stop);
        kprof_end("restrict");
    } else {
        k_residual<0><<<blocks_for(cells, TPB), TPB, 0, g_stream>>>(L, v.x, v.b, v.buf, stop);
        k_restrict<0><<<blocks_for(cc, TPB), TPB, 0, g_stream>>>(mg->ndim, v.n[0], v.n[1], v.n[2], v.buf, v.mask, c.n[0],
                                                                   c.n[1], c.n[2], c.b, c.mask, nullptr, stop);
    }
    FLUX_TRY(vcycle(mg, l + 1, stop));
    if (l == 0) {
        kprof_begin("correct");
        k_correct<1><<<blocks_for(cells, TPB), TPB, 0, g_stream>>>(mg->ndim, c.n[0], c.n[1], c.n[2], c.x, v.n[0], v.n[1],
                                                                     v.n[2], v.x, v.mask, stop);
        kprof_end("correct");
    } else {
        k_correct<0><<<blocks_for(cells, TPB), TPB, 0, g_stream>>>(mg->ndim, c.n[0], c.n[1], c.n[2], c.x, v.n[0], v.n[1],
                                                                     v.n[2], v.x, v.mask, stop);
    }
    relax_sweeps(mg, l, l > 0 ? mg->o.nu_up : mg->o.n_final, true, false, stop);
    return FLUX_OK;
}

// z = M r on the stream; `stop` (device) short-circuits every kernel.
int mg_apply_raw(flux_multigrid* mg, const float* r, float* z, const int* stop) {
    mg->lv[0].x = z;
    mg->lv[0].b = (float*)r;
    FLUX_TRY(vcycle(mg, 0, stop));
    FLUX_LAUNCH_CHECK();
    return FLUX_OK;
}

}  // namespace flux

extern "C" {

void flux_mg_options_default(flux_mg_options* o) {
    o->levels = 1;
    o->nu_down = 2;
    o->nu_up = 2;
    o->n_coarsest = 20;
    o->n_final = 2;
    o->omega = 1.0;
    o->storage = FLUX_STORE_F32;
}

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
    delete op;
    return FLUX_OK;
}

int flux_multigrid_create(int ndim, const int* shape, const double* spacing, const flux_mg_options* opts,
                          const flux_field* mask, const flux_field* coarse_mask, const flux_field* const* w,
                          flux_multigrid** out) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(out && shape && spacing, "multigrid: NULL argument");
    FLUX_TRY(check_shape(ndim, shape, spacing, ndim, "multigrid"));
    flux_mg_options o;
    if (opts) o = *opts;
    else flux_mg_options_default(&o);
    FLUX_REQUIRE(o.levels >= 1 && o.levels <= 12, "multigrid: numberOfLevels %d out of range", o.levels);
    FLUX_REQUIRE(o.storage == FLUX_STORE_F32, "multigrid: storage %d is not available in this build", o.storage);
    int div = 1 << (o.levels - 1);
    for (int a = 0; a < ndim; ++a)
        if (shape[a] % div)
            return set_error(FLUX_ERR_ARG,
                             "multigrid: shape [%d, %d, %d] is not divisible by 2^(numberOfLevels-1)=%d (numberOfLevels=%d).",
                             shape[0], shape[1], ndim > 2 ? shape[2] : 1, div, o.levels);
    int n0[3] = {shape[0], shape[1], ndim > 2 ? shape[2] : 1};
    if (mask) FLUX_TRY(check_field(mask, FLUX_U8, cells_of(n0), "multigrid: dirichletMask"));
    if (coarse_mask) FLUX_TRY(check_field(coarse_mask, FLUX_U8, cells_of(n0), "multigrid: coarseDirichletMask"));
    if (w)
        for (int a = 0; a < ndim; ++a) FLUX_TRY(check_field(w[a], FLUX_F32, faces_of(n0, a), "multigrid: faceWeights"));
    flux_multigrid* mg = new flux_multigrid();
    mg->ndim = ndim;
    mg->o = o;
    mg->mask = mask;
    mg->coarse_mask = coarse_mask;
    mg->has_w = (w != nullptr);
    mg->has_masks = (mask != nullptr) || (w != nullptr);
    mg->lv.resize(o.levels);
    bool ok = true;
    for (int l = 0; l < o.levels; ++l) {
        MGLevel& v = mg->lv[l];
        v.ndim_ = ndim;
        for (int a = 0; a < 3; ++a) {
            v.n[a] = (a < ndim) ? (n0[a] >> l) : 1;
            v.h[a] = (a < ndim) ? spacing[a] * (double)(1 << l) : 1.0;
        }
        int64_t cells = cells_of(v.n);
        v.buf = (float*)dmalloc(cells * 4);
        ok = ok && v.buf;
        if (l == 0) {
            v.mask = mask ? (uint8_t*)mask->d : nullptr;
            for (int a = 0; a < ndim; ++a) v.w[a] = w ? (float*)w[a]->d : nullptr;
        } else {
            v.own = true;
            v.x = (float*)dmalloc(cells * 4);
            v.b = (float*)dmalloc(cells * 4);
            ok = ok && v.x && v.b;
            if (mg->has_masks) {
                v.mask = (uint8_t*)dmalloc(cells);
                ok = ok && v.mask;
                if (v.mask) launch_fill(v.mask, FLUX_U8, cells, 0);
            }
            if (mg->has_w)
                for (int a = 0; a < ndim; ++a) {
                    int64_t nf = faces_of(v.n, a);
                    v.w[a] = (float*)dmalloc(nf * 4);
                    ok = ok && v.w[a];
                    if (v.w[a]) launch_fill(v.w[a], FLUX_F32, nf, 1.0);
                }
            if (v.x) launch_fill(v.x, FLUX_F32, cells, 0.0);
            if (v.b) launch_fill(v.b, FLUX_F32, cells, 0.0);
        }
    }
    if (!ok) {
        flux_multigrid_destroy(mg);
        return set_error(FLUX_ERR_NOMEM, "multigrid: device allocation failed");
    }
    int rc = mg_refresh(mg);
    if (rc == FLUX_OK && hipStreamSynchronize(g_stream) != hipSuccess) rc = set_error(FLUX_ERR_HIP, "multigrid: setup failed");
    if (rc) {
        flux_multigrid_destroy(mg);
        return rc;
    }
    *out = mg;
    return FLUX_OK;
}

int flux_multigrid_refresh(flux_multigrid* mg) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(mg, "refreshCoarseLevels: NULL multigrid");
    FLUX_TRY(mg_refresh(mg));
    FLUX_HIP(hipStreamSynchronize(g_stream));
    return FLUX_OK;
}

int flux_multigrid_apply(flux_multigrid* mg, const flux_field* in, flux_field* out) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(mg, "multigrid: NULL");
    int64_t cells = cells_of(mg->lv[0].n);
    FLUX_TRY(check_field(in, FLUX_F32, cells, "multigrid: input"));
    FLUX_TRY(check_field(out, FLUX_F32, cells, "multigrid: output"));
    FLUX_REQUIRE(in->d != out->d, "multigrid: input and output must differ");
    FLUX_TRY(mg_apply_raw(mg, (const float*)in->d, (float*)out->d, nullptr));
    FLUX_HIP(hipStreamSynchronize(g_stream));
    return FLUX_OK;
}

int flux_multigrid_level(flux_multigrid* mg, int level, int* shape_out, uint8_t* mask_out, float* const* w_out) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(mg && level >= 0 && level < (int)mg->lv.size(), "multigrid: level out of range");
    MGLevel& v = mg->lv[level];
    if (shape_out)
        for (int a = 0; a < 3; ++a) shape_out[a] = v.n[a];
    FLUX_HIP(hipStreamSynchronize(g_stream));
    if (mask_out && v.mask) FLUX_HIP(hipMemcpy(mask_out, v.mask, cells_of(v.n), hipMemcpyDeviceToHost));
    if (w_out)
        for (int a = 0; a < mg->ndim; ++a)
            if (w_out[a] && v.w[a]) FLUX_HIP(hipMemcpy(w_out[a], v.w[a], faces_of(v.n, a) * 4, hipMemcpyDeviceToHost));
    return FLUX_OK;
}

int flux_multigrid_destroy(flux_multigrid* mg) {
    if (!mg) return FLUX_OK;
    if (g_init) (void)hipStreamSynchronize(g_stream);
    for (size_t l = 0; l < mg->lv.size(); ++l) {
        MGLevel& v = mg->lv[l];
        dfree(v.buf);
        if (v.own) {
            dfree(v.x);
            dfree(v.b);
            dfree(v.mask);
            for (int a = 0; a < 3; ++a) dfree(v.w[a]);
        }
    }
    delete mg;
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
    float *r = nullptr, *z = nullptr, *pa = nullptr, *pb = nullptr, *Ap = nullptr;
    double *xa = nullptr, *xb = nullptr;
    Scal* S = nullptr;
    Scal* Sh = nullptr;  // pinned host mirror
    double* partials = nullptr;
    int last_iterations = 4;
};

namespace flux {

static int pcg_precondition(flux_pcg* s) {
    const int* stop = &s->S->stop;
    unsigned g = blocks_for(s->n, TPB);
    if (s->precond == FLUX_PREC_MULTIGRID) return mg_apply_raw(s->mg, s->r, s->z, stop);
    if (s->precond == FLUX_PREC_JACOBI) k_jacobi<<<g, TPB, 0, g_stream>>>(s->op->L, s->r, s->z, stop);
    else k_copy_f32<<<g, TPB, 0, g_stream>>>(s->r, s->z, (int)s->n, stop);
    return FLUX_OK;
}

static int pcg_iteration(flux_pcg* s) {
    unsigned g = blocks_for(s->n, TPB);
    kprof_begin("apply");
    k_pcg_apply<<<g, TPB, 0, g_stream>>>(s->op->L, s->z, s->pa, s->pb, s->Ap, s->S, s->partials);
    kprof_end("apply");
    kprof_begin("update");
    k_pcg_update<<<g, TPB, 0, g_stream>>>(s->op->L, s->xa, s->xb, s->pa, s->pb, (const float*)s->b->d, s->r, s->S,
                                           s->partials);
    kprof_end("update");
    FLUX_TRY(pcg_precondition(s));
    k_pcg_dot_rz<<<g, TPB, 0, g_stream>>>(s->r, s->z, (int)s->n, s->S, s->partials, 0);
    FLUX_LAUNCH_CHECK();
    return FLUX_OK;
}

int pcg_solve_raw(flux_pcg* s, double tol, int maxiter, int relative, flux_pcg_state* st) {
    unsigned g = blocks_for(s->n, TPB);
    // host-written fields of the scalar block
    memset(s->Sh, 0, sizeof(Scal));
    s->Sh->tol = tol;
    s->Sh->relative = relative ? 1 : 0;
    FLUX_HIP(hipMemcpyAsync(s->S, s->Sh, sizeof(Scal), hipMemcpyHostToDevice, g_stream));
    FLUX_HIP(hipStreamSynchronize(g_stream));
    // setup (runGpuResidentSetup, linalg.js:1613-1637)
    k_pcg_setup_residual<<<g, TPB, 0, g_stream>>>(s->op->L, (const float*)s->x->d, s->xa, (const float*)s->b->d, s->r,
                                                   s->S, s->partials);
    k_pcg_dot_bb<<<g, TPB, 0, g_stream>>>((const float*)s->b->d, (int)s->n, s->S, s->partials);
    FLUX_TRY(pcg_precondition(s));
    k_pcg_dot_rz<<<g, TPB, 0, g_stream>>>(s->r, s->z, (int)s->n, s->S, s->partials, 1);
    k_fill_f32<<<g, TPB, 0, g_stream>>>(s->pa, (int)s->n, 0.0f);
    k_fill_f32<<<g, TPB, 0, g_stream>>>(s->pb, (int)s->n, 0.0f);
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
    k_pcg_copy_out<<<g, TPB, 0, g_stream>>>(s->xa, s->xb, (float*)s->x->d, (int)s->n, s->S);
    FLUX_LAUNCH_CHECK();
    FLUX_HIP(hipStreamSynchronize(g_stream));
    const Scal& h = *s->Sh;
    st->stopped_by = h.stop;
    st->converged = (h.stop == 5);
    st->iterations = st->converged ? h.stop_iter : h.iter;
    st->residual = sqrt(fabs(h.rr));
    st->threshold = h.thr;
    st->b_norm = sqrt(fabs(h.bb));
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
    if (mg) FLUX_REQUIRE(cells_of(mg->lv[0].n) == cells, "conjugateGradient: multigrid shape mismatch");
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
    s->Ap = (float*)dmalloc(cells * 4);
    s->xa = (double*)dmalloc(cells * 8);
    s->xb = (double*)dmalloc(cells * 8);
    s->S = (Scal*)dmalloc(sizeof(Scal));
    s->partials = (double*)dmalloc(sizeof(double) * (blocks_for(cells, TPB) + 64));
    void* hp = nullptr;
    if (hipHostMalloc(&hp, sizeof(Scal), hipHostMallocDefault) != hipSuccess) hp = nullptr;
    s->Sh = (Scal*)hp;
    if (!s->r || !s->z || !s->pa || !s->pb || !s->Ap || !s->xa || !s->xb || !s->S || !s->partials || !s->Sh) {
        flux_pcg_destroy(s);
        return set_error(FLUX_ERR_NOMEM, "conjugateGradient: allocation failed");
    }
    // the ticket counters start at zero
    FLUX_HIP(hipMemsetAsync(s->S, 0, sizeof(Scal), g_stream));
    FLUX_HIP(hipStreamSynchronize(g_stream));
    *out = s;
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
    dfree(s->Ap);
    dfree(s->xa);
    dfree(s->xb);
    dfree(s->S);
    dfree(s->partials);
    if (s->Sh) (void)hipHostFree(s->Sh);
    delete s;
    return FLUX_OK;
}

}  // extern "C"
