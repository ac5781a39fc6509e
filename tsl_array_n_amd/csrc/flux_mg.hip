// Multigrid preconditioner for gfx950: createMultigridPreconditioner
// (packages/fluxflow/src/linalg/multigrid.js:1213-1735).
//
// Every kernel computes in FP32; storage is FP32 (the reference's form) or,
// for the mixed-precision form, FP16/BF16 for all coarse-level fields and the
// face weights.  The finest level's x and b are the PCG's z and r and stay
// FP32.  With FP16 storage the restricted residual is scaled to unit RMS
// (s = sqrt(N / r.r), undone in the prolongation) so late, small residuals
// never fall into FP16's subnormal range; the V-cycle is linear, so the
// scaling does not change what it computes.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "flux_internal.h"
#include "flux_solver.h"
#include "flux_stencil.h"
#include "flux_stencil3.h"
#include "flux_sweep.h"
#include "flux_brick.h"

namespace flux {

static constexpr int MTPB = 256;

template <typename T>
__global__ void k_fill_t(T* __restrict__ p, int n, float v) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = to_store<T>(v);
}

template <typename T>
__global__ void k_copy_t(const T* __restrict__ a, T* __restrict__ b, int n, const int* __restrict__ stop) {
    if (stop && *stop) return;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) b[t] = a[t];
}

template <typename I, typename O>
__global__ void k_convert(const I* __restrict__ a, O* __restrict__ b, int n) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) b[t] = to_store<O>(to_f32(a[t]));
}

// residual scale for FP16 storage: unit RMS of the finest residual
__device__ __forceinline__ float residual_scale(const double* rr, int n) {
    if (!rr) return 1.0f;
    double v = *rr;
    if (!(v > 0.0) || !(v < 1e300)) return 1.0f;
    return (float)sqrt((double)n / v);
}

// One red-black Gauss-Seidel colour (buildRelaxKernel, multigrid.js:626-640):
// x_c += ((b_c - (A x)_c) / diag_c) * omega on cells with (i+j+k)%2 == color.
// Thread t covers the t-th cell of that colour.  FROM_ZERO takes the field as
// zero on entry and writes every cell (the colour's cells with omega*b/diag,
// the others with 0), folding zeroX into the sweep.  FINE marks finest-level
// launches so profilers report them as kernels of their own.
template <int FINE, int FROM_ZERO, typename XT, typename WT>
__global__ void k_relax(LevelDevT<WT> L, XT* __restrict__ x, const XT* __restrict__ b, float omega, int color,
                        const int* __restrict__ stop) {
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
    if (FROM_ZERO) {
        int ip = 2 * m + ((j + k + color + 1) & 1);
        if (ip < L.n0) x[ip + L.n0 * row] = to_store<XT>(0.0f);
    }
    if (i >= L.n0) return;
    int c = i + L.n0 * row;
    float d;
    float ax;
    if (FROM_ZERO) {
        (void)stencil_row<float>(L, i, j, k, c, [&](int) { return 0.0f; }, &d);
        ax = 0.0f;
    } else {
        ax = stencil_row<float>(L, i, j, k, c, [&](int q) { return to_f32(x[q]); }, &d);
    }
    float cur = FROM_ZERO ? 0.0f : to_f32(x[c]);
    x[c] = to_store<XT>(cur + ((to_f32(b[c]) - ax) / d) * omega);
}

// k_relax for levels with an even number of cells along x, in the column-march
// shape: thread t owns the cell pair (2p, 2p+1) of row j and marches over a
// chunk of planes.  In every plane exactly one cell of the pair has the
// colour; the pair is read and written as one aligned vector (the partner's
// value is written back unchanged: no other thread writes it in this colour),
// so every lane updates a cell and the stores fill whole lines.
//
// DOT (level 0's last launch of the V-cycle): each block also stores its
// partial of b.x over both cells of every pair (the PCG's r.z with the final
// z), for a one-block finishing kernel.
template <int FINE, int FROM_ZERO, int DOT, typename XT, typename WT>
__global__ void __launch_bounds__(MTPB) k_relax3(LevelDevT<WT> L, March M, XT* __restrict__ x,
                                                 const XT* __restrict__ b, float omega, int color,
                                                 const int* __restrict__ stop, double* __restrict__ partials) {
    if (stop && *stop) return;
    double dot = 0.0;
    int t, k0, k1;
    // Every thread of a DOT launch reaches the block reduction exactly once
    // (it holds a barrier): a thread past the plane skips the march, it does
    // not return early, since a plane need not fill whole wavefronts.
    const bool active = march_begin(M, t, k0, k1);
    if (!DOT && !active) return;
    if (active) {
        const int half = M.run;
        const int j = t / half;
        const int p = t - j * half;
        const int n0 = L.n0, n1 = L.n1, n2 = L.n2;
        const int plane = n0 * n1;
        const uint8_t* __restrict__ nbits = L.nb;
        const float ih2[3] = {L.ih2[0], L.ih2[1], L.ih2[2]};
        // The pair of the next plane is loaded one plane ahead: in plane k the
        // coloured cell is 2p+s, and in planes k-1 and k+1 the cell 2p+s has the
        // other colour, so its (unchanged) value is the partner this thread read
        // there -- the z-neighbours come from registers.
        Pair2<XT> xnext;
        Pair2<uint8_t> nnext;
        const int cp0 = 2 * p + n0 * (j + n1 * k0);
        if (!FROM_ZERO) xnext = *reinterpret_cast<const Pair2<XT>*>(x + cp0);
        if (nbits) nnext = *reinterpret_cast<const Pair2<uint8_t>*>(nbits + cp0);
        float xprev = 0.0f;
        for (int k = k0; k < k1; ++k) {
            const int s = (j + k + color) & 1;  // which cell of the pair has the colour
            const int i = 2 * p + s;
            const int cp = 2 * p + n0 * (j + n1 * k);
            const int c = cp + s;
            const Pair2<XT> xp = xnext;
            const Pair2<uint8_t> np = nnext;
            if (k + 1 < n2) {
                if (!FROM_ZERO) xnext = *reinterpret_cast<const Pair2<XT>*>(x + cp + plane);
                if (nbits && k + 1 < k1) nnext = *reinterpret_cast<const Pair2<uint8_t>*>(nbits + cp + plane);
            }
            const unsigned bits = nbits ? (s ? np.hi : np.lo) : 0u;
            float w[6];
            face_w3(L, i, j, k, c, w);
            const bool dir = (bits & 64u) != 0;
            const Pair2<XT> bp = *reinterpret_cast<const Pair2<XT>*>(b + cp);
            const float bc = to_f32(s ? bp.hi : bp.lo);
            if (FROM_ZERO) {
                // x = 0 on entry: A x = 0 on a free row and -0 on a Dirichlet row
                float sum, dsum;
                const float zn[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                row_terms<float>(0.0f, zn, w, ih2, sum, dsum);
                const float d = dir ? -1.0f : dsum;
                const float ax = dir ? -0.0f : sum;
                const float v = 0.0f + ((bc - ax) / d) * omega;
                Pair2<XT> o;
                o.lo = to_store<XT>(s ? 0.0f : v);
                o.hi = to_store<XT>(s ? v : 0.0f);
                *reinterpret_cast<Pair2<XT>*>(x + cp) = o;
            } else {
                const float xc = to_f32(s ? xp.hi : xp.lo);
                const float xpart = to_f32(s ? xp.lo : xp.hi);
                // the pair partner is the upper x neighbour of the even cell and the lower of the odd one
                const int qo = s ? (i < n0 - 1 ? c + 1 : c) : (i > 0 ? c - 1 : c);
                const float xo = to_f32(x[qo]);
                float xn[6];
                xn[0] = s ? xpart : xo;
                xn[1] = s ? xo : xpart;
                xn[2] = to_f32(x[j > 0 ? c - n0 : c]);
                xn[3] = to_f32(x[j < n1 - 1 ? c + n0 : c]);
                xn[4] = k > k0 ? xprev : to_f32(x[k > 0 ? c - plane : c]);
                xn[5] = k + 1 < n2 ? to_f32(s ? xnext.hi : xnext.lo) : xc;
    #pragma unroll
                for (int a = 0; a < 6; ++a) xn[a] = ((bits >> a) & 1u) ? 0.0f : xn[a];
                float sum, dsum;
                row_terms<float>(xc, xn, w, ih2, sum, dsum);
                const float d = dir ? -1.0f : dsum;
                const float ax = dir ? -xc : sum;
                const XT nv = to_store<XT>(xc + ((bc - ax) / d) * omega);
                Pair2<XT> o = xp;
                if (s) o.hi = nv;
                else o.lo = nv;
                *reinterpret_cast<Pair2<XT>*>(x + cp) = o;
                xprev = xpart;
                if (DOT) {
                    dot += (double)to_f32(bp.lo) * (double)to_f32(o.lo);
                    dot += (double)to_f32(bp.hi) * (double)to_f32(o.hi);
                }
            }
        }
    }
    if (DOT) store_block_partial(dot, partials);
}

// One red-black colour of the finest level in the mixed form, where the
// level's x and b are stored in ST (FP16/BF16) in scaled units (b~ = r s,
// x~ = x s / xs; the operator carries xs / h^2), in k_relax3's pair shape:
//   MODE 1 (first launch, colour 0, from x~ = 0): reads r (FP32), stores
//          b~ = r s for the later launches and the residual, and x~;
//   MODE 0: one colour in place on x~ with b~;
//   MODE 2 (the V-cycle's last launch): the colour's update with b from r
//          (FP32, times s); both cells of the pair are written to the FP32
//          output z = x~ xs / s (the PCG's z), with DOT the r.z partials.
// In MODE 2 the pair is read from x~ and written to z, so no x~ value is
// overwritten while another thread may still read it.
template <int MODE, int DOT, typename ST>
__global__ void __launch_bounds__(MTPB) k_relax3m(LevelDevT<ST> L, March M, ST* __restrict__ x,
                                                  const ST* __restrict__ b16, ST* __restrict__ b16out,
                                                  const float* __restrict__ r, float* __restrict__ z,
                                                  const double* __restrict__ rr, float xs, float omega, int color,
                                                  const int* __restrict__ stop, double* __restrict__ partials) {
    if (stop && *stop) return;
    double dot = 0.0;
    int t, k0, k1;
    const bool active = march_begin(M, t, k0, k1);
    if (!DOT && !active) return;
    if (active) {
        const int half = M.run;
        const int j = t / half;
        const int p = t - j * half;
        const int n0 = L.n0, n1 = L.n1, n2 = L.n2;
        const int plane = n0 * n1;
        const uint8_t* __restrict__ nbits = L.nb;
        const float ih2[3] = {L.ih2[0], L.ih2[1], L.ih2[2]};
        const float s = residual_scale(rr, n0 * n1 * n2);
        const float zscale = xs / s;
        Pair2<ST> xnext;
        Pair2<uint8_t> nnext;
        const int cp0 = 2 * p + n0 * (j + n1 * k0);
        if (MODE != 1) xnext = *reinterpret_cast<const Pair2<ST>*>(x + cp0);
        if (nbits) nnext = *reinterpret_cast<const Pair2<uint8_t>*>(nbits + cp0);
        float xprev = 0.0f;
        for (int k = k0; k < k1; ++k) {
            const int sel = (j + k + color) & 1;  // which cell of the pair has the colour
            const int i = 2 * p + sel;
            const int cp = 2 * p + n0 * (j + n1 * k);
            const int c = cp + sel;
            const Pair2<ST> xp = xnext;
            const Pair2<uint8_t> np = nnext;
            if (k + 1 < n2) {
                if (MODE != 1) xnext = *reinterpret_cast<const Pair2<ST>*>(x + cp + plane);
                if (nbits && k + 1 < k1) nnext = *reinterpret_cast<const Pair2<uint8_t>*>(nbits + cp + plane);
            }
            const unsigned bits = nbits ? (sel ? np.hi : np.lo) : 0u;
            float w[6];
            face_w3(L, i, j, k, c, w);
            const bool dir = (bits & 64u) != 0;
            float bc, rlo = 0.0f, rhi = 0.0f;
            if (MODE == 0) {
                const Pair2<ST> bp = *reinterpret_cast<const Pair2<ST>*>(b16 + cp);
                bc = to_f32(sel ? bp.hi : bp.lo);
            } else {
                const Pair2<float> rp = *reinterpret_cast<const Pair2<float>*>(r + cp);
                rlo = rp.lo;
                rhi = rp.hi;
                if (MODE == 1) {
                    Pair2<ST> bo;
                    bo.lo = to_store<ST>(rp.lo * s);
                    bo.hi = to_store<ST>(rp.hi * s);
                    *reinterpret_cast<Pair2<ST>*>(b16out + cp) = bo;
                    bc = to_f32(sel ? bo.hi : bo.lo);
                } else {
                    bc = (sel ? rp.hi : rp.lo) * s;
                }
            }
            if (MODE == 1) {
                float sum, dsum;
                const float zn[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                row_terms<float>(0.0f, zn, w, ih2, sum, dsum);
                const float d = dir ? -1.0f : dsum;
                const float ax = dir ? -0.0f : sum;
                const float v = 0.0f + ((bc - ax) / d) * omega;
                Pair2<ST> o;
                o.lo = to_store<ST>(sel ? 0.0f : v);
                o.hi = to_store<ST>(sel ? v : 0.0f);
                *reinterpret_cast<Pair2<ST>*>(x + cp) = o;
            } else {
                const float xc = to_f32(sel ? xp.hi : xp.lo);
                const float xpart = to_f32(sel ? xp.lo : xp.hi);
                const int qo = sel ? (i < n0 - 1 ? c + 1 : c) : (i > 0 ? c - 1 : c);
                const float xo = to_f32(x[qo]);
                float xn[6];
                xn[0] = sel ? xpart : xo;
                xn[1] = sel ? xo : xpart;
                xn[2] = to_f32(x[j > 0 ? c - n0 : c]);
                xn[3] = to_f32(x[j < n1 - 1 ? c + n0 : c]);
                xn[4] = k > k0 ? xprev : to_f32(x[k > 0 ? c - plane : c]);
                xn[5] = k + 1 < n2 ? to_f32(sel ? xnext.hi : xnext.lo) : xc;
#pragma unroll
                for (int a = 0; a < 6; ++a) xn[a] = ((bits >> a) & 1u) ? 0.0f : xn[a];
                float sum, dsum;
                row_terms<float>(xc, xn, w, ih2, sum, dsum);
                const float d = dir ? -1.0f : dsum;
                const float ax = dir ? -xc : sum;
                const ST nv = to_store<ST>(xc + ((bc - ax) / d) * omega);
                Pair2<ST> o = xp;
                if (sel) o.hi = nv;
                else o.lo = nv;
                if (MODE == 2) {
                    Pair2<float> zo;
                    zo.lo = to_f32(o.lo) * zscale;
                    zo.hi = to_f32(o.hi) * zscale;
                    *reinterpret_cast<Pair2<float>*>(z + cp) = zo;
                    if (DOT) dot += (double)rlo * (double)zo.lo + (double)rhi * (double)zo.hi;
                } else {
                    *reinterpret_cast<Pair2<ST>*>(x + cp) = o;
                }
                xprev = xpart;
            }
        }
    }
    if (DOT) store_block_partial(dot, partials);
}

// One damped-Jacobi sweep (buildJacobiRelaxKernel,
// sandbox/jacobi-smoother/multigrid_jacobi.js:632-642): every cell reads the
// old iterate xr and writes xr_c + ((b - A xr)_c / diag_c) * omega to xw, so
// the sweep is out of place (ping-pong).  Generic shape: one cell per thread.
template <int FINE, int FROM_ZERO, typename XT, typename WT>
__global__ void k_jacobi(LevelDevT<WT> L, const XT* __restrict__ xr, XT* __restrict__ xw, const XT* __restrict__ b,
                         float omega, const int* __restrict__ stop) {
    if (stop && *stop) return;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= L.n0 * L.n1 * L.n2) return;
    int i, j, k;
    cell_coords(L, c, i, j, k);
    float d;
    float ax;
    if (FROM_ZERO) {
        (void)stencil_row<float>(L, i, j, k, c, [&](int) { return 0.0f; }, &d);
        ax = 0.0f;
    } else {
        ax = stencil_row<float>(L, i, j, k, c, [&](int q) { return to_f32(xr[q]); }, &d);
    }
    const float cur = FROM_ZERO ? 0.0f : to_f32(xr[c]);
    xw[c] = to_store<XT>(cur + ((to_f32(b[c]) - ax) / d) * omega);
}

// The damped-Jacobi sweep for an even row length in the column-march shape:
// thread t owns the cell pair (2p, 2p+1) of row j and marches over a chunk of
// planes; both cells are updated from the old iterate and stored as one
// aligned vector.  The z-neighbour pairs come from registers (the previous
// plane's centre pair and the next plane's, loaded one plane ahead), and so
// does the lower z-face weight pair (the previous plane's upper one).
// FROM_ZERO takes xr as zero (xr is not read); DOT stores each block's
// partial of b.xw (the PCG's r.z with the final z).
template <int FINE, int FROM_ZERO, int DOT, typename XT, typename WT>
__global__ void __launch_bounds__(MTPB) k_jacobi3(LevelDevT<WT> L, March M, const XT* __restrict__ xr,
                                                  XT* __restrict__ xw, const XT* __restrict__ b, float omega,
                                                  const int* __restrict__ stop, double* __restrict__ partials) {
    if (stop && *stop) return;
    double dot = 0.0;
    int t, k0, k1;
    const bool active = march_begin(M, t, k0, k1);
    if (!DOT && !active) return;
    if (active) {
        const int half = M.run;
        const int j = t / half;
        const int p = t - j * half;
        const int n0 = L.n0, n1 = L.n1, n2 = L.n2;
        const int plane = n0 * n1;
        const int i0 = 2 * p;
        const uint8_t* __restrict__ nbits = L.nb;
        const float ih2[3] = {L.ih2[0], L.ih2[1], L.ih2[2]};
        const bool has_w = L.w0 != nullptr;
        Pair2<XT> xc2, xnext, xprev;
        const int cp0 = i0 + n0 * (j + n1 * k0);
        if (!FROM_ZERO) {
            xnext = *reinterpret_cast<const Pair2<XT>*>(xr + cp0);
            if (k0 > 0) xprev = *reinterpret_cast<const Pair2<XT>*>(xr + cp0 - plane);
        }
        Pair2<WT> wz_lo;
        if (has_w) wz_lo = *reinterpret_cast<const Pair2<WT>*>(L.w2 + cp0);
        for (int k = k0; k < k1; ++k) {
            const int cp = i0 + n0 * (j + n1 * k);
            if (!FROM_ZERO) {
                xc2 = xnext;
                if (k + 1 < n2) xnext = *reinterpret_cast<const Pair2<XT>*>(xr + cp + plane);
            }
            const Pair2<uint8_t> np = nbits ? *reinterpret_cast<const Pair2<uint8_t>*>(nbits + cp) : Pair2<uint8_t>{0, 0};
            const Pair2<XT> bp = *reinterpret_cast<const Pair2<XT>*>(b + cp);
            // face weights of both cells: x faces i0, i0+1, i0+2; y faces of rows j and j+1; z faces k and k+1
            float wx0 = 1.0f, wx1 = 1.0f, wx2 = 1.0f, wyl[2] = {1.0f, 1.0f}, wyu[2] = {1.0f, 1.0f};
            float wzl[2] = {1.0f, 1.0f}, wzu[2] = {1.0f, 1.0f};
            Pair2<WT> wz_up;
            if (has_w) {
                const int fx = i0 + (n0 + 1) * (j + n1 * k);
                wx0 = to_f32(L.w0[fx]);
                wx1 = to_f32(L.w0[fx + 1]);
                wx2 = to_f32(L.w0[fx + 2]);
                const int fy = i0 + n0 * (j + (n1 + 1) * k);
                const Pair2<WT> yl = *reinterpret_cast<const Pair2<WT>*>(L.w1 + fy);
                const Pair2<WT> yu = *reinterpret_cast<const Pair2<WT>*>(L.w1 + fy + n0);
                wyl[0] = to_f32(yl.lo);
                wyl[1] = to_f32(yl.hi);
                wyu[0] = to_f32(yu.lo);
                wyu[1] = to_f32(yu.hi);
                wz_up = *reinterpret_cast<const Pair2<WT>*>(L.w2 + cp + plane);
                wzl[0] = to_f32(wz_lo.lo);
                wzl[1] = to_f32(wz_lo.hi);
                wzu[0] = to_f32(wz_up.lo);
                wzu[1] = to_f32(wz_up.hi);
            }
            float w[2][6];
            w[0][0] = i0 > 0 ? wx0 : 0.0f;
            w[0][1] = wx1;
            w[1][0] = wx1;
            w[1][1] = i0 + 1 < n0 - 1 ? wx2 : 0.0f;
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                w[s][2] = j > 0 ? wyl[s] : 0.0f;
                w[s][3] = j < n1 - 1 ? wyu[s] : 0.0f;
                w[s][4] = k > 0 ? wzl[s] : 0.0f;
                w[s][5] = k < n2 - 1 ? wzu[s] : 0.0f;
            }
            float out[2];
            if (FROM_ZERO) {
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    const unsigned bits = s ? np.hi : np.lo;
                    const float zn[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                    float sum, dsum;
                    row_terms<float>(0.0f, zn, w[s], ih2, sum, dsum);
                    const bool dir = (bits & 64u) != 0;
                    const float d = dir ? -1.0f : dsum;
                    const float ax = dir ? -0.0f : sum;
                    out[s] = 0.0f + ((to_f32(s ? bp.hi : bp.lo) - ax) / d) * omega;
                }
            } else {
                const float xl = to_f32(xr[i0 > 0 ? cp - 1 : cp]);
                const float xu = to_f32(xr[i0 + 1 < n0 - 1 ? cp + 2 : cp + 1]);
                const Pair2<XT> ys = *reinterpret_cast<const Pair2<XT>*>(xr + (j > 0 ? cp - n0 : cp));
                const Pair2<XT> yn = *reinterpret_cast<const Pair2<XT>*>(xr + (j < n1 - 1 ? cp + n0 : cp));
                const float c0 = to_f32(xc2.lo), c1 = to_f32(xc2.hi);
                float xn[2][6];
                xn[0][0] = xl;
                xn[0][1] = c1;
                xn[1][0] = c0;
                xn[1][1] = xu;
                xn[0][2] = to_f32(ys.lo);
                xn[1][2] = to_f32(ys.hi);
                xn[0][3] = to_f32(yn.lo);
                xn[1][3] = to_f32(yn.hi);
                xn[0][4] = k > 0 ? to_f32(xprev.lo) : c0;
                xn[1][4] = k > 0 ? to_f32(xprev.hi) : c1;
                xn[0][5] = k + 1 < n2 ? to_f32(xnext.lo) : c0;
                xn[1][5] = k + 1 < n2 ? to_f32(xnext.hi) : c1;
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    const unsigned bits = s ? np.hi : np.lo;
#pragma unroll
                    for (int a = 0; a < 6; ++a) xn[s][a] = ((bits >> a) & 1u) ? 0.0f : xn[s][a];
                    const float xc = s ? c1 : c0;
                    float sum, dsum;
                    row_terms<float>(xc, xn[s], w[s], ih2, sum, dsum);
                    const bool dir = (bits & 64u) != 0;
                    const float d = dir ? -1.0f : dsum;
                    const float ax = dir ? -xc : sum;
                    out[s] = xc + ((to_f32(s ? bp.hi : bp.lo) - ax) / d) * omega;
                }
                xprev = xc2;
            }
            Pair2<XT> o;
            o.lo = to_store<XT>(out[0]);
            o.hi = to_store<XT>(out[1]);
            *reinterpret_cast<Pair2<XT>*>(xw + cp) = o;
            if (has_w) wz_lo = wz_up;
            if (DOT) {
                dot += (double)to_f32(bp.lo) * (double)to_f32(o.lo);
                dot += (double)to_f32(bp.hi) * (double)to_f32(o.hi);
            }
        }
    }
    if (DOT) store_block_partial(dot, partials);
}

// The damped-Jacobi sweep of the finest level in the mixed form, where the
// level's x and b are stored in ST (FP16/BF16) in scaled units: b~ = r s
// (s the residual scale, unit RMS for FP16, 1 for BF16) and x~ = x s / xs
// (xs the level's storage scale, the operator carrying xs / h^2).  The
// PCG's vectors stay FP32:
//   MODE 1 (first sweep, from x~ = 0): reads r (FP32), stores b~ = r s (ST)
//          for the later sweeps and the residual, and x~' (ST);
//   MODE 0 (middle sweeps): x~ -> x~' with b~;
//   MODE 2 (last sweep): reads b from r (FP32, times s), writes the FP32
//          output z = x~' xs / s and, with DOT, the r.z partials.
// Same column-march pair shape and per-cell arithmetic as k_jacobi3.
template <int MODE, int DOT, typename ST>
__global__ void __launch_bounds__(MTPB) k_jacobi3m(LevelDevT<ST> L, March M, const ST* __restrict__ xr,
                                                   ST* __restrict__ xw, const ST* __restrict__ b16,
                                                   ST* __restrict__ b16out, const float* __restrict__ r,
                                                   float* __restrict__ z, const double* __restrict__ rr, float xs,
                                                   float omega, const int* __restrict__ stop,
                                                   double* __restrict__ partials) {
    if (stop && *stop) return;
    double dot = 0.0;
    int t, k0, k1;
    const bool active = march_begin(M, t, k0, k1);
    if (!DOT && !active) return;
    if (active) {
        const int half = M.run;
        const int j = t / half;
        const int p = t - j * half;
        const int n0 = L.n0, n1 = L.n1, n2 = L.n2;
        const int plane = n0 * n1;
        const int i0 = 2 * p;
        const uint8_t* __restrict__ nbits = L.nb;
        const float ih2[3] = {L.ih2[0], L.ih2[1], L.ih2[2]};
        const bool has_w = L.w0 != nullptr;
        const float s = residual_scale(rr, n0 * n1 * n2);
        const float zscale = xs / s;
        Pair2<ST> xc2, xnext, xprev;
        const int cp0 = i0 + n0 * (j + n1 * k0);
        if (MODE != 1) {
            xnext = *reinterpret_cast<const Pair2<ST>*>(xr + cp0);
            if (k0 > 0) xprev = *reinterpret_cast<const Pair2<ST>*>(xr + cp0 - plane);
        }
        Pair2<ST> wz_lo;
        if (has_w) wz_lo = *reinterpret_cast<const Pair2<ST>*>(L.w2 + cp0);
        for (int k = k0; k < k1; ++k) {
            const int cp = i0 + n0 * (j + n1 * k);
            if (MODE != 1) {
                xc2 = xnext;
                if (k + 1 < n2) xnext = *reinterpret_cast<const Pair2<ST>*>(xr + cp + plane);
            }
            const Pair2<uint8_t> np = nbits ? *reinterpret_cast<const Pair2<uint8_t>*>(nbits + cp) : Pair2<uint8_t>{0, 0};
            float bv[2], rv[2] = {0.0f, 0.0f};
            if (MODE == 0) {
                const Pair2<ST> bp = *reinterpret_cast<const Pair2<ST>*>(b16 + cp);
                bv[0] = to_f32(bp.lo);
                bv[1] = to_f32(bp.hi);
            } else {
                const Pair2<float> rp = *reinterpret_cast<const Pair2<float>*>(r + cp);
                rv[0] = rp.lo;
                rv[1] = rp.hi;
                bv[0] = rp.lo * s;
                bv[1] = rp.hi * s;
                if (MODE == 1) {
                    Pair2<ST> bo;
                    bo.lo = to_store<ST>(bv[0]);
                    bo.hi = to_store<ST>(bv[1]);
                    *reinterpret_cast<Pair2<ST>*>(b16out + cp) = bo;
                    // the later sweeps see b~ rounded to storage; so does this one
                    bv[0] = to_f32(bo.lo);
                    bv[1] = to_f32(bo.hi);
                }
            }
            float wx0 = 1.0f, wx1 = 1.0f, wx2 = 1.0f, wyl[2] = {1.0f, 1.0f}, wyu[2] = {1.0f, 1.0f};
            float wzl[2] = {1.0f, 1.0f}, wzu[2] = {1.0f, 1.0f};
            Pair2<ST> wz_up;
            if (has_w) {
                const int fx = i0 + (n0 + 1) * (j + n1 * k);
                wx0 = to_f32(L.w0[fx]);
                wx1 = to_f32(L.w0[fx + 1]);
                wx2 = to_f32(L.w0[fx + 2]);
                const int fy = i0 + n0 * (j + (n1 + 1) * k);
                const Pair2<ST> yl = *reinterpret_cast<const Pair2<ST>*>(L.w1 + fy);
                const Pair2<ST> yu = *reinterpret_cast<const Pair2<ST>*>(L.w1 + fy + n0);
                wyl[0] = to_f32(yl.lo);
                wyl[1] = to_f32(yl.hi);
                wyu[0] = to_f32(yu.lo);
                wyu[1] = to_f32(yu.hi);
                wz_up = *reinterpret_cast<const Pair2<ST>*>(L.w2 + cp + plane);
                wzl[0] = to_f32(wz_lo.lo);
                wzl[1] = to_f32(wz_lo.hi);
                wzu[0] = to_f32(wz_up.lo);
                wzu[1] = to_f32(wz_up.hi);
            }
            float w[2][6];
            w[0][0] = i0 > 0 ? wx0 : 0.0f;
            w[0][1] = wx1;
            w[1][0] = wx1;
            w[1][1] = i0 + 1 < n0 - 1 ? wx2 : 0.0f;
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                w[q][2] = j > 0 ? wyl[q] : 0.0f;
                w[q][3] = j < n1 - 1 ? wyu[q] : 0.0f;
                w[q][4] = k > 0 ? wzl[q] : 0.0f;
                w[q][5] = k < n2 - 1 ? wzu[q] : 0.0f;
            }
            float out[2];
            if (MODE == 1) {
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    const unsigned bits = q ? np.hi : np.lo;
                    const float zn[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                    float sum, dsum;
                    row_terms<float>(0.0f, zn, w[q], ih2, sum, dsum);
                    const bool dir = (bits & 64u) != 0;
                    const float d = dir ? -1.0f : dsum;
                    const float ax = dir ? -0.0f : sum;
                    out[q] = 0.0f + ((bv[q] - ax) / d) * omega;
                }
            } else {
                const float xl = to_f32(xr[i0 > 0 ? cp - 1 : cp]);
                const float xu = to_f32(xr[i0 + 1 < n0 - 1 ? cp + 2 : cp + 1]);
                const Pair2<ST> ys = *reinterpret_cast<const Pair2<ST>*>(xr + (j > 0 ? cp - n0 : cp));
                const Pair2<ST> yn = *reinterpret_cast<const Pair2<ST>*>(xr + (j < n1 - 1 ? cp + n0 : cp));
                const float c0 = to_f32(xc2.lo), c1 = to_f32(xc2.hi);
                float xn[2][6];
                xn[0][0] = xl;
                xn[0][1] = c1;
                xn[1][0] = c0;
                xn[1][1] = xu;
                xn[0][2] = to_f32(ys.lo);
                xn[1][2] = to_f32(ys.hi);
                xn[0][3] = to_f32(yn.lo);
                xn[1][3] = to_f32(yn.hi);
                xn[0][4] = k > 0 ? to_f32(xprev.lo) : c0;
                xn[1][4] = k > 0 ? to_f32(xprev.hi) : c1;
                xn[0][5] = k + 1 < n2 ? to_f32(xnext.lo) : c0;
                xn[1][5] = k + 1 < n2 ? to_f32(xnext.hi) : c1;
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    const unsigned bits = q ? np.hi : np.lo;
#pragma unroll
                    for (int a = 0; a < 6; ++a) xn[q][a] = ((bits >> a) & 1u) ? 0.0f : xn[q][a];
                    const float xc = q ? c1 : c0;
                    float sum, dsum;
                    row_terms<float>(xc, xn[q], w[q], ih2, sum, dsum);
                    const bool dir = (bits & 64u) != 0;
                    const float d = dir ? -1.0f : dsum;
                    const float ax = dir ? -xc : sum;
                    out[q] = xc + ((bv[q] - ax) / d) * omega;
                }
                xprev = xc2;
            }
            if (MODE == 2) {
                Pair2<float> o;
                o.lo = to_f32(to_store<ST>(out[0])) * zscale;
                o.hi = to_f32(to_store<ST>(out[1])) * zscale;
                *reinterpret_cast<Pair2<float>*>(z + cp) = o;
                if (DOT) {
                    dot += (double)rv[0] * (double)o.lo;
                    dot += (double)rv[1] * (double)o.hi;
                }
            } else {
                Pair2<ST> o;
                o.lo = to_store<ST>(out[0]);
                o.hi = to_store<ST>(out[1]);
                *reinterpret_cast<Pair2<ST>*>(xw + cp) = o;
            }
            if (has_w) wz_lo = wz_up;
        }
    }
    if (DOT) store_block_partial(dot, partials);
}

// r = b - A x (buildResidualKernel, multigrid.js:741-749) in the column-march
// shape (one cell per thread and plane), as the restriction's input.
template <int FINE, typename XT, typename RT, typename WT>
__global__ void __launch_bounds__(MTPB) k_residual3(LevelDevT<WT> L, March M, const XT* __restrict__ x,
                                                    const XT* __restrict__ b, RT* __restrict__ r,
                                                    const int* __restrict__ stop) {
    if (stop && *stop) return;
    int t, k0, k1;
    if (!march_begin(M, t, k0, k1)) return;
    const int n0 = L.n0, n1 = L.n1, n2 = L.n2;
    const int plane = n0 * n1;
    const int j = t / n0;
    const int i = t - j * n0;
    const uint8_t* __restrict__ nbits = L.nb;
    const float ih2[3] = {L.ih2[0], L.ih2[1], L.ih2[2]};
    // the z-neighbours are the previous and the next plane's centres
    float xnext = to_f32(x[i + n0 * (j + n1 * k0)]);
    float xprev = 0.0f;
    for (int k = k0; k < k1; ++k) {
        const int c = i + n0 * (j + n1 * k);
        const float xc = xnext;
        if (k + 1 < n2) xnext = to_f32(x[c + plane]);
        const unsigned bits = nbits ? nbits[c] : 0u;
        float w[6];
        face_w3(L, i, j, k, c, w);
        float xn[6];
        xn[0] = to_f32(x[i > 0 ? c - 1 : c]);
        xn[1] = to_f32(x[i < n0 - 1 ? c + 1 : c]);
        xn[2] = to_f32(x[j > 0 ? c - n0 : c]);
        xn[3] = to_f32(x[j < n1 - 1 ? c + n0 : c]);
        xn[4] = k > k0 ? xprev : to_f32(x[k > 0 ? c - plane : c]);
        xn[5] = k + 1 < n2 ? xnext : xc;
#pragma unroll
        for (int a = 0; a < 6; ++a) xn[a] = ((bits >> a) & 1u) ? 0.0f : xn[a];
        float sum, dsum;
        row_terms<float>(xc, xn, w, ih2, sum, dsum);
        // a Dirichlet cell's residual is stored as 0: its only reader, the
        // restriction, counts masked fine cells as 0
        const bool dir = (bits & 64u) != 0;
        r[c] = to_store<RT>(dir ? 0.0f : to_f32(b[c]) - sum);
        xprev = xc;
    }
}

// The residual r = b - A x of a 3-D level (k_residual3's per-cell arithmetic)
// folded at once into the z-taps of the full-weighting restriction:
//   rz[K][j][i] = sum_c tz[c] r(i, j, fk_c(K)),  fk = 2K-1, 2K, 2K+1, 2K+2
// (clamped to 2K and 2K+1 at the ends, tz = 1/8, 3/8, 3/8, 1/8, summed in
// tap order), so r itself is never stored and the restriction that follows
// (k_restrict_xy) reads half as many values and sums only the x-y taps.  A
// thread marches its column over an even z-chunk [k0, k1) of fine planes
// and also evaluates the residual at the planes k0 - 1 and k1 its first
// and last coarse planes reach.  Dirichlet cells contribute 0, as their
// stored residual does in the unfolded form.
template <typename XT, typename WT>
__global__ void __launch_bounds__(MTPB) k_residual3z(LevelDevT<WT> L, March M, const XT* __restrict__ x,
                                                     const XT* __restrict__ b, float* __restrict__ rz,
                                                     const int* __restrict__ stop) {
    if (stop && *stop) return;
    int t, k0, k1;
    if (!march_begin(M, t, k0, k1)) return;
    const int n0 = L.n0, n1 = L.n1, n2 = L.n2;
    const int plane = n0 * n1;
    const int j = t / n0;
    const int i = t - j * n0;
    const uint8_t* __restrict__ nbits = L.nb;
    const float ih2[3] = {L.ih2[0], L.ih2[1], L.ih2[2]};
    const float tz[4] = {0.125f, 0.375f, 0.375f, 0.125f};
    const int K0 = k0 >> 1, K1 = k1 >> 1;  // coarse planes of this chunk
    const int kb = k0 > 0 ? k0 - 1 : 0;
    const int ke = k1 < n2 ? k1 : n2 - 1;  // inclusive
    float xnext = to_f32(x[i + n0 * (j + n1 * kb)]);
    float xprev = 0.0f;
    // z-sums in progress: K = m - 1 (awaiting its last tap), K = m, K = m + 1
    float acc_prev = 0.0f, acc_cur = 0.0f, acc_next = 0.0f;
    for (int k = kb; k <= ke; ++k) {
        const int c = i + n0 * (j + n1 * k);
        const float xc = xnext;
        if (k + 1 < n2) xnext = to_f32(x[c + plane]);
        const unsigned bits = nbits ? nbits[c] : 0u;
        float w[6];
        face_w3(L, i, j, k, c, w);
        float xn[6];
        xn[0] = to_f32(x[i > 0 ? c - 1 : c]);
        xn[1] = to_f32(x[i < n0 - 1 ? c + 1 : c]);
        xn[2] = to_f32(x[j > 0 ? c - n0 : c]);
        xn[3] = to_f32(x[j < n1 - 1 ? c + n0 : c]);
        xn[4] = k > kb ? xprev : to_f32(x[k > 0 ? c - plane : c]);
        xn[5] = k + 1 < n2 ? xnext : xc;
#pragma unroll
        for (int a = 0; a < 6; ++a) xn[a] = ((bits >> a) & 1u) ? 0.0f : xn[a];
        float sum, dsum;
        row_terms<float>(xc, xn, w, ih2, sum, dsum);
        const bool dir = (bits & 64u) != 0;
        const float r = dir ? 0.0f : to_f32(b[c]) - sum;
        const int m = k >> 1;
        if ((k & 1) == 0) {
            // k = 2m: the last tap of K = m - 1, the second of K = m (for
            // K = 0 also its first: 2K - 1 clamps to 2K)
            if (m - 1 >= K0 && m - 1 < K1) {
                acc_prev = acc_prev + r * tz[3];
                rz[i + n0 * (j + n1 * (m - 1))] = acc_prev;
            }
            if (m >= K0 && m < K1) {
                if (m == 0) acc_cur = r * tz[0];
                acc_cur = acc_cur + r * tz[1];
            }
        } else {
            // k = 2m + 1: the third tap of K = m (for the last K also its
            // fourth: 2K + 2 clamps to 2K + 1), the first of K = m + 1
            if (m >= K0 && m < K1) {
                acc_cur = acc_cur + r * tz[2];
                if (k == n2 - 1) {
                    acc_cur = acc_cur + r * tz[3];
                    rz[i + n0 * (j + n1 * m)] = acc_cur;
                }
            }
            if (m + 1 >= K0 && m + 1 < K1) acc_next = r * tz[0];
            acc_prev = acc_cur;
            acc_cur = acc_next;
        }
        xprev = xc;
    }
}

// The x-y taps of the full-weighting restriction over the z-folded residual
// of k_residual3z (same taps and clamps as k_restrict3 per axis); masked
// coarse cells get 0; the residual scale of the finest level when rr is set.
template <typename OT>
__global__ void __launch_bounds__(MTPB) k_restrict_xy(March M, int f0, int f1, int f2, const float* __restrict__ rz,
                                                      int c0, int c1, OT* __restrict__ bc,
                                                      const uint8_t* __restrict__ mc, const double* __restrict__ rr,
                                                      const int* __restrict__ stop) {
    if (stop && *stop) return;
    int t, K0, K1;
    if (!march_begin(M, t, K0, K1)) return;
    const int J = t / c0;
    const int I = t - J * c0;
    const float tw[4] = {0.125f, 0.375f, 0.375f, 0.125f};
    int ti[4], tj[4];
    ti[0] = I > 0 ? 2 * I - 1 : 2 * I;
    ti[1] = 2 * I;
    ti[2] = 2 * I + 1;
    ti[3] = (I + 1 < c0) ? 2 * I + 2 : 2 * I + 1;
    tj[0] = J > 0 ? 2 * J - 1 : 2 * J;
    tj[1] = 2 * J;
    tj[2] = 2 * J + 1;
    tj[3] = (J + 1 < c1) ? 2 * J + 2 : 2 * J + 1;
    const float scale = residual_scale(rr, f0 * f1 * f2);
    for (int K = K0; K < K1; ++K) {
        const float* __restrict__ p = rz + (int64_t)f0 * f1 * K;
        float sum = 0.0f;
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int bq = 0; bq < 4; ++bq) sum += p[ti[a] + f0 * tj[bq]] * (tw[a] * tw[bq]);
        const int c = I + c0 * (J + c1 * K);
        bc[c] = to_store<OT>((mc && mc[c]) ? 0.0f : sum * scale);
    }
}

// Full-weighting restriction (restrictionTapsForAxis 766-780 + buildRestrictKernel
// 866-914): per axis taps 1/8,3/8,3/8,1/8 at 2c-1 (clamped to 2c), 2c, 2c+1,
// 2c+2 (clamped to 2c+1); masked fine cells contribute 0; masked coarse cells
// get 0.  The result is multiplied by the residual scale when rr is given.
template <int FINE, typename IT, typename OT>
__global__ void k_restrict(int ndim, int f0, int f1, int f2, const IT* __restrict__ rf, const uint8_t* __restrict__ mf,
                           int c0, int c1, int c2, OT* __restrict__ bc, const uint8_t* __restrict__ mc,
                           const double* __restrict__ rr, const int* __restrict__ stop) {
    if (stop && *stop) return;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= c0 * c1 * c2) return;
    int row = t / c0;
    int I = t - row * c0;
    int K = row / c1;
    int J = row - K * c1;
    if (mc && mc[t]) {
        bc[t] = to_store<OT>(0.0f);
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
                float v = (mf && mf[f]) ? 0.0f : to_f32(rf[f]);
                sum += v * wgt;
            }
    bc[t] = to_store<OT>(sum * residual_scale(rr, f0 * f1 * f2));
}

// Bilinear / trilinear prolongation added to the fine x except at masked fine
// cells (correctionTapsForAxis 921-934, buildCorrectKernel 990-1040); the
// residual scale is undone when rr is given.
template <int FINE, typename CT, typename FT>
__global__ void k_correct(int ndim, int c0, int c1, int c2, const CT* __restrict__ xc, int f0, int f1, int f2,
                          FT* __restrict__ xf, const uint8_t* __restrict__ mf, const double* __restrict__ rr,
                          float mult, const int* __restrict__ stop) {
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
                sum += to_f32(xc[ti[0][a] + c0 * (ti[1][bq] + c1 * ti[2][cq])]) * wgt;
            }
    if (rr) sum = sum / residual_scale(rr, f0 * f1 * f2);
    if (mult != 1.0f) sum = sum * mult;
    xf[t] = to_store<FT>(to_f32(xf[t]) + sum);
}

// 3-D full-weighting restriction in the column-march shape: thread (I,J)
// marches over coarse planes K.  Each fine plane f is first reduced over its
// 4 x 4 in-plane taps to P(f); the coarse value is sum_c tz[c] * P(fk[c]).
// Consecutive coarse planes share two fine planes, whose P values are carried
// in registers.  Same taps, clamps and masking as k_restrict.
template <int FINE, typename IT, typename OT>
__global__ void __launch_bounds__(MTPB) k_restrict3(March M, int f0, int f1, int f2, const IT* __restrict__ rf,
                                                    const uint8_t* __restrict__ mf, int c0, int c1, int c2,
                                                    OT* __restrict__ bc, const uint8_t* __restrict__ mc,
                                                    const double* __restrict__ rr, const int* __restrict__ stop) {
    if (stop && *stop) return;
    int t, K0, K1;
    if (!march_begin(M, t, K0, K1)) return;
    const int J = t / c0;
    const int I = t - J * c0;
    const float tw[4] = {0.125f, 0.375f, 0.375f, 0.125f};
    int ti[4], tj[4];
    ti[0] = I > 0 ? 2 * I - 1 : 2 * I;
    ti[1] = 2 * I;
    ti[2] = 2 * I + 1;
    ti[3] = (I + 1 < c0) ? 2 * I + 2 : 2 * I + 1;
    tj[0] = J > 0 ? 2 * J - 1 : 2 * J;
    tj[1] = 2 * J;
    tj[2] = 2 * J + 1;
    tj[3] = (J + 1 < c1) ? 2 * J + 2 : 2 * J + 1;
    const int fplane = f0 * f1;
    const float scale = residual_scale(rr, f0 * f1 * f2);
    auto reduce_plane = [&](int fk) {
        float sum = 0.0f;
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int bq = 0; bq < 4; ++bq) {
                const int f = ti[a] + f0 * tj[bq] + fplane * fk;
                const float v = (mf && mf[f]) ? 0.0f : to_f32(rf[f]);
                sum += v * (tw[a] * tw[bq]);
            }
        return sum;
    };
    float Pa = 0.0f, Pb = 0.0f;  // P at fine planes 2K+1 and 2K+2 of the previous K
    for (int K = K0; K < K1; ++K) {
        const int tk0 = K > 0 ? 2 * K - 1 : 2 * K;
        const int tk3 = (K + 1 < c2) ? 2 * K + 2 : 2 * K + 1;
        float P0, P1;
        if (K == K0) {
            P0 = reduce_plane(tk0);
            P1 = (tk0 == 2 * K) ? P0 : reduce_plane(2 * K);
        } else {
            P0 = Pa;  // 2(K-1)+1 = 2K-1
            P1 = Pb;  // 2(K-1)+2 = 2K
        }
        const float P2 = reduce_plane(2 * K + 1);
        const float P3 = (tk3 == 2 * K + 1) ? P2 : reduce_plane(tk3);
        Pa = P2;
        Pb = P3;
        const int c = I + c0 * (J + c1 * K);
        const float v = P0 * tw[0] + P1 * tw[1] + P2 * tw[2] + P3 * tw[3];
        bc[c] = to_store<OT>((mc && mc[c]) ? 0.0f : v * scale);
    }
}

// 3-D trilinear prolongation added to the fine x in the column-march shape
// (same taps and masking as k_correct).
template <int FINE, typename CT, typename FT>
__global__ void __launch_bounds__(MTPB) k_correct3(March M, int c0, int c1, int c2, const CT* __restrict__ xc,
                                                   int f0, int f1, int f2, FT* __restrict__ xf,
                                                   const uint8_t* __restrict__ mf, const double* __restrict__ rr,
                                                   float mult, const int* __restrict__ stop) {
    if (stop && *stop) return;
    int t, k0, k1;
    if (!march_begin(M, t, k0, k1)) return;
    const int j = t / f0;
    const int i = t - j * f0;
    auto taps = [](int F, int nc, int* ti, float* tw) {
        const int ci = F >> 1;
        const bool even = (F & 1) == 0;
        const int lower = ci > 0 ? ci - 1 : ci;
        const int upper = ci < nc - 1 ? ci + 1 : ci;
        ti[0] = even ? lower : ci;
        tw[0] = even ? 0.25f : 0.75f;
        ti[1] = even ? ci : upper;
        tw[1] = even ? 0.75f : 0.25f;
    };
    int xi[2], yj[2];
    float wx[2], wy[2];
    taps(i, c0, xi, wx);
    taps(j, c1, yj, wy);
    const float scale = rr ? residual_scale(rr, f0 * f1 * f2) : 1.0f;
    const int cplane = c0 * c1;
    for (int k = k0; k < k1; ++k) {
        const int c = i + f0 * (j + f1 * k);
        // every load is issued before the mask is known (one memory round
        // trip per plane); a masked cell is not written
        const bool masked = mf && mf[c];
        const float old = to_f32(xf[c]);
        int zk[2];
        float wz[2];
        taps(k, c2, zk, wz);
        float sum = 0.0f;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int bq = 0; bq < 2; ++bq)
#pragma unroll
                for (int cq = 0; cq < 2; ++cq) {
                    const float wgt = wx[a] * wy[bq] * wz[cq];
                    sum += to_f32(xc[xi[a] + c0 * yj[bq] + cplane * zk[cq]]) * wgt;
                }
        if (rr) sum = sum / scale;
        if (mult != 1.0f) sum = sum * mult;
        if (!masked) xf[c] = to_store<FT>(old + sum);
    }
}

// k_correct3 for an even fine row length: each thread owns the cell pair
// (2p, 2p+1), whose x-taps are (p-1, p) and (p, p+1) (clamped), so the pair
// reads 3 coarse columns instead of 4 and the fine x / mask as one vector.
// Every cell's sum runs over the same taps in the same order as k_correct3.
template <int FINE, typename CT, typename FT>
__global__ void __launch_bounds__(MTPB) k_correct3p(March M, int c0, int c1, int c2, const CT* __restrict__ xc,
                                                    int f0, int f1, int f2, FT* __restrict__ xf,
                                                    const uint8_t* __restrict__ mf, const double* __restrict__ rr,
                                                    float mult, const int* __restrict__ stop) {
    if (stop && *stop) return;
    int t, k0, k1;
    if (!march_begin(M, t, k0, k1)) return;
    const int h0 = f0 >> 1;
    const int j = t / h0;
    const int p = t - j * h0;
    auto taps = [](int F, int nc, int* ti, float* tw) {
        const int ci = F >> 1;
        const bool even = (F & 1) == 0;
        const int lower = ci > 0 ? ci - 1 : ci;
        const int upper = ci < nc - 1 ? ci + 1 : ci;
        ti[0] = even ? lower : ci;
        tw[0] = even ? 0.25f : 0.75f;
        ti[1] = even ? ci : upper;
        tw[1] = even ? 0.75f : 0.25f;
    };
    // x-columns: lower = (p-1 clamped), p, upper = (p+1 clamped)
    const int xl = p > 0 ? p - 1 : p, xu = p < c0 - 1 ? p + 1 : p;
    int yj[2];
    float wy[2];
    taps(j, c1, yj, wy);
    const float scale = rr ? residual_scale(rr, f0 * f1 * f2) : 1.0f;
    const int cplane = c0 * c1;
    for (int k = k0; k < k1; ++k) {
        const int c = 2 * p + f0 * (j + f1 * k);
        Pair2<uint8_t> mp{0, 0};
        if (mf) mp = *reinterpret_cast<const Pair2<uint8_t>*>(mf + c);
        const Pair2<FT> old = *reinterpret_cast<const Pair2<FT>*>(xf + c);
        int zk[2];
        float wz[2];
        taps(k, c2, zk, wz);
        float v[3][2][2];  // [x column][y tap][z tap]
#pragma unroll
        for (int bq = 0; bq < 2; ++bq)
#pragma unroll
            for (int cq = 0; cq < 2; ++cq) {
                const int base = c0 * yj[bq] + cplane * zk[cq];
                v[0][bq][cq] = to_f32(xc[xl + base]);
                v[1][bq][cq] = to_f32(xc[p + base]);
                v[2][bq][cq] = to_f32(xc[xu + base]);
            }
        // even cell: taps (lower 0.25, p 0.75); odd cell: (p 0.75, upper 0.25)
        const float wxe[2] = {0.25f, 0.75f}, wxo[2] = {0.75f, 0.25f};
        float se = 0.0f, so = 0.0f;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int bq = 0; bq < 2; ++bq)
#pragma unroll
                for (int cq = 0; cq < 2; ++cq) {
                    se += v[a][bq][cq] * (wxe[a] * wy[bq] * wz[cq]);
                    so += v[a + 1][bq][cq] * (wxo[a] * wy[bq] * wz[cq]);
                }
        if (rr) {
            se = se / scale;
            so = so / scale;
        }
        if (mult != 1.0f) {
            se = se * mult;
            so = so * mult;
        }
        Pair2<FT> out;
        out.lo = mp.lo ? old.lo : to_store<FT>(to_f32(old.lo) + se);
        out.hi = mp.hi ? old.hi : to_store<FT>(to_f32(old.hi) + so);
        *reinterpret_cast<Pair2<FT>*>(xf + c) = out;
    }
}

// Coarsest level in one workgroup: `iters` x (colour 0, colour 1) sweeps from
// x = 0 (buildCoarseSweepKernel, multigrid.js:679-737).  x lives in LDS for
// the whole solve; each thread keeps its cells' b, face weights, neighbour
// indices and Dirichlet bits in registers.  Every update is rounded through
// the storage type, as a store to the level's x would be.
static constexpr int COARSE_LDS_CELLS = 4096;
static constexpr int COARSE_TPB = 1024;
// With `jacobi` the sweeps are damped-Jacobi single sweeps (`iters` of them,
// even), each computed from the previous iterate before any value changes.
template <typename XT, typename WT>
__device__ void coarse_solve_block(const LevelDevT<WT>& L, XT* __restrict__ x, const XT* __restrict__ b, float omega,
                                   int iters, int jacobi) {
    constexpr int PER = COARSE_LDS_CELLS / COARSE_TPB;
    __shared__ float sx[COARSE_LDS_CELLS];
    const int n = L.n0 * L.n1 * L.n2;
    const uint8_t* __restrict__ m = L.mask;
    const float ih2[3] = {L.ih2[0], L.ih2[1], L.ih2[2]};
    float bc[PER], w[PER][6];
    int q[PER][6];
    unsigned bits[PER];  // bit a: neighbour a is Dirichlet; bit 6: the cell is; bit 7: colour
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int c = threadIdx.x + u * COARSE_TPB;
        bits[u] = 0;
        bc[u] = 0.0f;
        if (c < n) {
            int i, j, k;
            cell_coords(L, c, i, j, k);
            face_w3(L, i, j, k, c, w[u]);
            nbr_index3(L.n0, L.n1, L.n2, i, j, k, c, q[u]);
            if (m) {
                for (int a = 0; a < 6; ++a) bits[u] |= m[q[u][a]] ? (1u << a) : 0u;
                bits[u] |= m[c] ? 64u : 0u;
            }
            bits[u] |= ((i + j + k) & 1) ? 128u : 0u;
            bc[u] = to_f32(b[c]);
        }
    }
    for (int c = threadIdx.x; c < n; c += COARSE_TPB) sx[c] = 0.0f;
    __syncthreads();
    if (jacobi) {
        for (int it = 0; it < iters; ++it) {
            float nv[PER];
#pragma unroll
            for (int u = 0; u < PER; ++u) {
                const int c = threadIdx.x + u * COARSE_TPB;
                nv[u] = 0.0f;
                if (c >= n) continue;
                const float xc = sx[c];
                float xn[6];
#pragma unroll
                for (int a = 0; a < 6; ++a) xn[a] = ((bits[u] >> a) & 1u) ? 0.0f : sx[q[u][a]];
                float sum, dsum;
                row_terms<float>(xc, xn, w[u], ih2, sum, dsum);
                const bool dir = (bits[u] >> 6) & 1u;
                const float d = dir ? -1.0f : dsum;
                const float ax = dir ? -xc : sum;
                nv[u] = to_f32(to_store<XT>(xc + ((bc[u] - ax) / d) * omega));
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < PER; ++u) {
                const int c = threadIdx.x + u * COARSE_TPB;
                if (c < n) sx[c] = nv[u];
            }
            __syncthreads();
        }
        iters = 0;
    }
    for (int it = 0; it < iters; ++it) {
        for (int color = 0; color < 2; ++color) {
#pragma unroll
            for (int u = 0; u < PER; ++u) {
                const int c = threadIdx.x + u * COARSE_TPB;
                if (c >= n || (int)((bits[u] >> 7) & 1u) != color) continue;
                const float xc = sx[c];
                float xn[6];
#pragma unroll
                for (int a = 0; a < 6; ++a) xn[a] = ((bits[u] >> a) & 1u) ? 0.0f : sx[q[u][a]];
                float sum, dsum;
                row_terms<float>(xc, xn, w[u], ih2, sum, dsum);
                const bool dir = (bits[u] >> 6) & 1u;
                const float d = dir ? -1.0f : dsum;
                const float ax = dir ? -xc : sum;
                sx[c] = to_f32(to_store<XT>(xc + ((bc[u] - ax) / d) * omega));
            }
            __syncthreads();
        }
    }
    for (int c = threadIdx.x; c < n; c += COARSE_TPB) x[c] = to_store<XT>(sx[c]);
}


template <typename XT, typename WT>
__global__ void __launch_bounds__(COARSE_TPB) k_coarse_solve(LevelDevT<WT> L, XT* __restrict__ x,
                                                             const XT* __restrict__ b, float omega, int iters,
                                                             int jacobi, const int* __restrict__ stop) {
    if (stop && *stop) return;
    coarse_solve_block<XT, WT>(L, x, b, omega, iters, jacobi);
}

// The V-cycle's tail in one workgroup: a 3-D level `lt` of at most
// COARSE_LDS_CELLS cells and the coarsest level below it.  Exactly the
// launches it replaces -- nu_down colour pairs from x = 0, the residual,
// the full-weighting restriction, the coarsest solve, the trilinear
// correction and nu_up reversed colour pairs -- with the same per-cell
// arithmetic and the same rounding through the storage type, but with every
// vector in LDS and one barrier between phases.  x of level lt is written out
// for the next finer level's correction.
template <typename ST>
__global__ void __launch_bounds__(COARSE_TPB) k_vcycle_tail(LevelDevT<ST> L0, LevelDevT<ST> L1,
                                                            const ST* __restrict__ b0, ST* __restrict__ x0,
                                                            float omega, int nu_down, int nu_up, int n_coarsest,
                                                            float mult, int jacobi, int post_forward,
                                                            const int* __restrict__ stop) {
    constexpr int PER = COARSE_LDS_CELLS / COARSE_TPB;
    __shared__ float sx0[COARSE_LDS_CELLS];
    __shared__ float sr0[COARSE_LDS_CELLS];
    __shared__ float sx1[COARSE_LDS_CELLS];
    if (stop && *stop) return;
    const int n0 = L0.n0 * L0.n1 * L0.n2;
    const int n1 = L1.n0 * L1.n1 * L1.n2;
    const float ih0[3] = {L0.ih2[0], L0.ih2[1], L0.ih2[2]};
    const float ih1[3] = {L1.ih2[0], L1.ih2[1], L1.ih2[2]};
    // this thread's level-lt cells: weights, neighbour slots, mask/colour bits, b
    float bc[PER], w[PER][6];
    int q[PER][6];
    unsigned bits[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int c = threadIdx.x + u * COARSE_TPB;
        bits[u] = 0;
        bc[u] = 0.0f;
        if (c < n0) {
            int i, j, k;
            cell_coords(L0, c, i, j, k);
            face_w3(L0, i, j, k, c, w[u]);
            nbr_index3(L0.n0, L0.n1, L0.n2, i, j, k, c, q[u]);
            if (L0.mask) {
                for (int a = 0; a < 6; ++a) bits[u] |= L0.mask[q[u][a]] ? (1u << a) : 0u;
                bits[u] |= L0.mask[c] ? 64u : 0u;
            }
            bits[u] |= ((i + j + k) & 1) ? 128u : 0u;
            bc[u] = to_f32(b0[c]);
        }
    }
    for (int c = threadIdx.x; c < n0; c += COARSE_TPB) sx0[c] = 0.0f;
    __syncthreads();
    auto sweep0 = [&](int color) {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int c = threadIdx.x + u * COARSE_TPB;
            if (c >= n0 || (int)((bits[u] >> 7) & 1u) != color) continue;
            const float xc = sx0[c];
            float xn[6];
#pragma unroll
            for (int a = 0; a < 6; ++a) xn[a] = ((bits[u] >> a) & 1u) ? 0.0f : sx0[q[u][a]];
            float sum, dsum;
            row_terms<float>(xc, xn, w[u], ih0, sum, dsum);
            const bool dir = (bits[u] >> 6) & 1u;
            const float d = dir ? -1.0f : dsum;
            const float ax = dir ? -xc : sum;
            sx0[c] = to_f32(to_store<ST>(xc + ((bc[u] - ax) / d) * omega));
        }
        __syncthreads();
    };
    // damped-Jacobi single sweep of level lt (k_jacobi3's arithmetic)
    auto jsweep0 = [&]() {
        float nv[PER];
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int c = threadIdx.x + u * COARSE_TPB;
            nv[u] = 0.0f;
            if (c >= n0) continue;
            const float xc = sx0[c];
            float xn[6];
#pragma unroll
            for (int a = 0; a < 6; ++a) xn[a] = ((bits[u] >> a) & 1u) ? 0.0f : sx0[q[u][a]];
            float sum, dsum;
            row_terms<float>(xc, xn, w[u], ih0, sum, dsum);
            const bool dir = (bits[u] >> 6) & 1u;
            const float d = dir ? -1.0f : dsum;
            const float ax = dir ? -xc : sum;
            nv[u] = to_f32(to_store<ST>(xc + ((bc[u] - ax) / d) * omega));
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int c = threadIdx.x + u * COARSE_TPB;
            if (c < n0) sx0[c] = nv[u];
        }
        __syncthreads();
    };
    for (int it = 0; it < nu_down; ++it) {
        if (jacobi) {
            jsweep0();
            continue;
        }
        sweep0(0);
        sweep0(1);
    }
    // residual (k_residual3): a Dirichlet cell's residual is 0
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int c = threadIdx.x + u * COARSE_TPB;
        if (c >= n0) continue;
        const float xc = sx0[c];
        float xn[6];
#pragma unroll
        for (int a = 0; a < 6; ++a) xn[a] = ((bits[u] >> a) & 1u) ? 0.0f : sx0[q[u][a]];
        float sum, dsum;
        row_terms<float>(xc, xn, w[u], ih0, sum, dsum);
        const bool dir = (bits[u] >> 6) & 1u;
        sr0[c] = to_f32(to_store<ST>(dir ? 0.0f : bc[u] - sum));
    }
    __syncthreads();
    // restriction (k_restrict3) into the coarsest level's b; its cells' data
    float bc1 = 0.0f, w1[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    int q1[6] = {0, 0, 0, 0, 0, 0};
    unsigned bits1 = 0;
    const int f0 = L0.n0, f1 = L0.n1, f2 = L0.n2;
    const int c0 = L1.n0, c1 = L1.n1, c2 = L1.n2;
    for (int t = threadIdx.x; t < n1; t += COARSE_TPB) {
        int I, J, K;
        cell_coords(L1, t, I, J, K);
        const float tw[4] = {0.125f, 0.375f, 0.375f, 0.125f};
        int ti[4], tj[4];
        ti[0] = I > 0 ? 2 * I - 1 : 2 * I;
        ti[1] = 2 * I;
        ti[2] = 2 * I + 1;
        ti[3] = (I + 1 < c0) ? 2 * I + 2 : 2 * I + 1;
        tj[0] = J > 0 ? 2 * J - 1 : 2 * J;
        tj[1] = 2 * J;
        tj[2] = 2 * J + 1;
        tj[3] = (J + 1 < c1) ? 2 * J + 2 : 2 * J + 1;
        auto reduce_plane = [&](int fk) {
            float sum = 0.0f;
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int bq = 0; bq < 4; ++bq) sum += sr0[ti[a] + f0 * tj[bq] + f0 * f1 * fk] * (tw[a] * tw[bq]);
            return sum;
        };
        const int tk0 = K > 0 ? 2 * K - 1 : 2 * K;
        const int tk3 = (K + 1 < c2) ? 2 * K + 2 : 2 * K + 1;
        const float P0 = reduce_plane(tk0);
        const float P1 = (tk0 == 2 * K) ? P0 : reduce_plane(2 * K);
        const float P2 = reduce_plane(2 * K + 1);
        const float P3 = (tk3 == 2 * K + 1) ? P2 : reduce_plane(tk3);
        const float v = P0 * tw[0] + P1 * tw[1] + P2 * tw[2] + P3 * tw[3];
        bc1 = to_f32(to_store<ST>((L1.mask && L1.mask[t]) ? 0.0f : v * 1.0f));
        face_w3(L1, I, J, K, t, w1);
        nbr_index3(c0, c1, c2, I, J, K, t, q1);
        if (L1.mask) {
            for (int a = 0; a < 6; ++a) bits1 |= L1.mask[q1[a]] ? (1u << a) : 0u;
            bits1 |= L1.mask[t] ? 64u : 0u;
        }
        bits1 |= ((I + J + K) & 1) ? 128u : 0u;
        sx1[t] = 0.0f;
    }
    (void)f2;
    __syncthreads();
    // the coarsest solve (coarse_solve_block), one cell per thread
    const bool own1 = (int)threadIdx.x < n1;
    if (jacobi) {
        for (int it = 0; it < n_coarsest; ++it) {
            float nv = 0.0f;
            if (own1) {
                const int c = threadIdx.x;
                const float xc = sx1[c];
                float xn[6];
#pragma unroll
                for (int a = 0; a < 6; ++a) xn[a] = ((bits1 >> a) & 1u) ? 0.0f : sx1[q1[a]];
                float sum, dsum;
                row_terms<float>(xc, xn, w1, ih1, sum, dsum);
                const bool dir = (bits1 >> 6) & 1u;
                const float d = dir ? -1.0f : dsum;
                const float ax = dir ? -xc : sum;
                nv = to_f32(to_store<ST>(xc + ((bc1 - ax) / d) * omega));
            }
            __syncthreads();
            if (own1) sx1[threadIdx.x] = nv;
            __syncthreads();
        }
    }
    for (int it = 0; it < (jacobi ? 0 : n_coarsest); ++it) {
        for (int color = 0; color < 2; ++color) {
            if (own1 && (int)((bits1 >> 7) & 1u) == color) {
                const int c = threadIdx.x;
                const float xc = sx1[c];
                float xn[6];
#pragma unroll
                for (int a = 0; a < 6; ++a) xn[a] = ((bits1 >> a) & 1u) ? 0.0f : sx1[q1[a]];
                float sum, dsum;
                row_terms<float>(xc, xn, w1, ih1, sum, dsum);
                const bool dir = (bits1 >> 6) & 1u;
                const float d = dir ? -1.0f : dsum;
                const float ax = dir ? -xc : sum;
                sx1[c] = to_f32(to_store<ST>(xc + ((bc1 - ax) / d) * omega));
            }
            __syncthreads();
        }
    }
    // trilinear correction (k_correct3) of level lt
    auto taps = [](int F, int nc, int* ti, float* tw) {
        const int ci = F >> 1;
        const bool even = (F & 1) == 0;
        const int lower = ci > 0 ? ci - 1 : ci;
        const int upper = ci < nc - 1 ? ci + 1 : ci;
        ti[0] = even ? lower : ci;
        tw[0] = even ? 0.25f : 0.75f;
        ti[1] = even ? ci : upper;
        tw[1] = even ? 0.75f : 0.25f;
    };
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int c = threadIdx.x + u * COARSE_TPB;
        if (c >= n0 || ((bits[u] >> 6) & 1u)) continue;
        int i, j, k;
        cell_coords(L0, c, i, j, k);
        int xi[2], yj[2], zk[2];
        float wx[2], wy[2], wz[2];
        taps(i, c0, xi, wx);
        taps(j, c1, yj, wy);
        taps(k, c2, zk, wz);
        float sum = 0.0f;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int bq = 0; bq < 2; ++bq)
#pragma unroll
                for (int cq = 0; cq < 2; ++cq) {
                    const float wgt = wx[a] * wy[bq] * wz[cq];
                    sum += sx1[xi[a] + c0 * yj[bq] + c0 * c1 * zk[cq]] * wgt;
                }
        if (mult != 1.0f) sum = sum * mult;
        sx0[c] = to_f32(to_store<ST>(sx0[c] + sum));
    }
    __syncthreads();
    for (int it = 0; it < nu_up; ++it) {
        if (jacobi) {
            jsweep0();
            continue;
        }
        sweep0(post_forward ? 0 : 1);
        sweep0(post_forward ? 1 : 0);
    }
    for (int c = threadIdx.x; c < n0; c += COARSE_TPB) x0[c] = to_store<ST>(sx0[c]);
}

// Coarse face weight = mean of the 2^(d-1) fine faces at index 2I along the
// face axis (buildFaceWeightCoarsenKernel, multigrid.js:1138-1184).
template <typename IT, typename OT>
__global__ void k_coarsen_w(int ndim, int axis, int f0, int f1, int f2, const IT* __restrict__ wf, int c0, int c1,
                            int c2, OT* __restrict__ wc) {
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
                sum += to_f32(wf[fi + fs0 * (fj + fs1 * fk)]);
            }
    wc[t] = to_store<OT>(sum / (float)(1 << (ndim - 1)));
}

// Coarse mask (buildMaskCoarsenKernel, multigrid.js:1066-1122): ALL 2^d fine
// children of `src` masked, OR the six coarse face weights sum below 1e-6.
template <typename WT>
__global__ void k_coarsen_mask(int ndim, int f0, int f1, const uint8_t* __restrict__ src, int c0, int c1, int c2,
                               const WT* __restrict__ w0, const WT* __restrict__ w1, const WT* __restrict__ w2,
                               uint8_t* __restrict__ mc) {
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
        float total = to_f32(w0[I + (c0 + 1) * (J + c1 * K)]) + to_f32(w0[I + 1 + (c0 + 1) * (J + c1 * K)]);
        total += to_f32(w1[I + c0 * (J + (c1 + 1) * K)]) + to_f32(w1[I + c0 * (J + 1 + (c1 + 1) * K)]);
        if (ndim == 3) total += to_f32(w2[t]) + to_f32(w2[t + c0 * c1]);
        bool enclosed = total < 1e-6f;
        masked = have ? (masked || enclosed) : enclosed;
        have = true;
    }
    mc[t] = (uint8_t)(have && masked);
}

}  // namespace flux

using namespace flux;

// ------------------------------------------------------------ host objects
struct MGLevel {
    int n[3];
    double h[3];
    void* x = nullptr;    // level 0: the caller's output (FP32)
    void* b = nullptr;    // level 0: the caller's input (FP32)
    void* buf = nullptr;  // residual (FP32 on level 0)
    void* alt = nullptr;  // ping-pong partner of x for the fused sweep
    // the finest level's storage-type copies in the mixed Jacobi form
    // (x~, b~ and the ping-pong / residual buffer), scaled units
    void *x16 = nullptr, *b16 = nullptr, *buf16 = nullptr;
    uint8_t* mask = nullptr;              // level 0: the caller's (not owned)
    uint8_t* nb = nullptr;                // neighbour-mask bits of the column-march kernels
    void* w[3] = {nullptr, nullptr, nullptr};  // level 0: caller's FP32 or own converted copy
    bool own = false;
    bool own_w = false;
    // the level stores x / xscale and applies xscale * A (1 except on FP16
    // coarse levels); Gauss-Seidel and the residual are unchanged by it
    double xscale = 1.0;
};

struct flux_multigrid {
    int ndim = 3;
    flux_mg_options o;
    std::vector<MGLevel> lv;
    const flux_field* mask = nullptr;
    const flux_field* coarse_mask = nullptr;
    const flux_field* w[3] = {nullptr, nullptr, nullptr};
    bool has_w = false, has_masks = false;
    // fused red-black sweeps (FLUX_FUSED_SWEEP): 1 the LDS-tile form on every
    // level wide enough, 2 the register form on the finest level
    int fused = 0;
    bool tail = true;    // the last two levels in one workgroup (FLUX_VCYCLE_TAIL=0 turns it off)
    bool pairs = true;   // paired prolongation-correction on even rows (FLUX_CORRECT_PAIRS=0 turns it off)
    // mixed form: the finest level stores x and b in the storage type too
    // (opt-in, FLUX_FINE_STORAGE=1).  Off by default: the rounding of z at
    // every smoothing update makes the V-cycle a slightly nonlinear operator
    // and PCG needs more iterations (64^3 dam break 50 -> 51 with FP16 and
    // red-black, 256^3 62 -> 82), more than the halved fine-level traffic
    // saves.
    bool fine16 = false;
    // post-smoothing in the pre-smoothing colour order, a non-symmetric
    // V-cycle (FLUX_MG_POST_FORWARD=1): a host-side switch that forces the
    // PCG guards in tests
    bool post_forward = false;
};

static constexpr int COARSE_SINGLE_BLOCK_MAX = COARSE_LDS_CELLS;
static constexpr int FUSED_MIN_N = 64;  // narrower levels keep the two-launch form

namespace {

size_t store_size(int st) { return st == FLUX_STORE_F32 ? 4 : 2; }

template <typename WT>
LevelDevT<WT> level_dev(const flux_multigrid* mg, const MGLevel& v) {
    LevelDevT<WT> L;
    L.n0 = v.n[0];
    L.n1 = v.n[1];
    L.n2 = v.n[2];
    L.ndim = mg->ndim;
    for (int a = 0; a < 3; ++a) {
        double hh = (a < mg->ndim) ? v.h[a] : 1.0;
        L.ih2d[a] = v.xscale / (hh * hh);
        L.ih2[a] = (float)L.ih2d[a];
    }
    L.mask = v.mask;
    L.w0 = (const WT*)v.w[0];
    L.w1 = (const WT*)v.w[1];
    L.w2 = mg->ndim == 3 ? (const WT*)v.w[2] : nullptr;
    L.nb = v.mask ? v.nb : nullptr;
    return L;
}

template <typename T>
void fill_t(void* p, int64_t n, float v) {
    if (n > 0) k_fill_t<T><<<blocks_for(n, MTPB), MTPB, 0, g_stream>>>((T*)p, (int)n, v);
}

// refreshCoarseLevels (multigrid.js:1415-1437): level-l weights from level
// l-1, then the level-l mask from level l-1's source (level 1 reads the
// vent-only mask), and in the mixed form the finest weights' storage copy.
template <typename ST>
int refresh_t(flux_multigrid* mg) {
    if (mg->has_w && mg->lv[0].own_w)
        for (int a = 0; a < mg->ndim; ++a) {
            int64_t nf = faces_of(mg->lv[0].n, a);
            k_convert<float, ST><<<blocks_for(nf, MTPB), MTPB, 0, g_stream>>>((const float*)mg->w[a]->d,
                                                                               (ST*)mg->lv[0].w[a], (int)nf);
        }
    for (size_t l = 1; l < mg->lv.size(); ++l) {
        MGLevel& f = mg->lv[l - 1];
        MGLevel& c = mg->lv[l];
        if (mg->has_w) {
            for (int a = 0; a < mg->ndim; ++a) {
                int64_t nf = faces_of(c.n, a);
                k_coarsen_w<ST, ST><<<blocks_for(nf, MTPB), MTPB, 0, g_stream>>>(
                    mg->ndim, a, f.n[0], f.n[1], f.n[2], (const ST*)f.w[a], c.n[0], c.n[1], c.n[2], (ST*)c.w[a]);
            }
        }
        if (!mg->has_masks) continue;
        const uint8_t* src = nullptr;
        if (mg->mask) src = (l == 1) ? (const uint8_t*)(mg->coarse_mask ? mg->coarse_mask->d : mg->mask->d) : f.mask;
        int64_t cc = cells_of(c.n);
        k_coarsen_mask<ST><<<blocks_for(cc, MTPB), MTPB, 0, g_stream>>>(
            mg->ndim, f.n[0], f.n[1], src, c.n[0], c.n[1], c.n[2], mg->has_w ? (const ST*)c.w[0] : nullptr,
            mg->has_w ? (const ST*)c.w[1] : nullptr, mg->has_w ? (const ST*)c.w[2] : nullptr, c.mask);
    }
    for (MGLevel& v : mg->lv)
        if (v.nb && v.mask)
            k_nbits3<ST><<<blocks_for(cells_of(v.n), MTPB), MTPB, 0, g_stream>>>(level_dev<ST>(mg, v), v.nb);
    FLUX_LAUNCH_CHECK();
    return FLUX_OK;
}

bool use_fused(const flux_multigrid* mg, const MGLevel& v) {
    if (mg->ndim != 3 || !v.alt || (v.n[0] & 1) != 0) return false;
    if (mg->fused == 2) return &v == &mg->lv[0] && v.n[0] <= 2 * SweepR::PX;
    return mg->fused == 1 && v.n[0] >= FUSED_MIN_N && v.n[1] >= Sweep3::TY && v.n[2] >= 2;
}

// planes per z-chunk of the register form (FLUX_SWEEP_KZ, default 8)
static int sweepr_chunk() {
    static const int v = [] {
        const char* e = getenv("FLUX_SWEEP_KZ");
        const long q = e ? strtol(e, nullptr, 10) : 8;
        return (int)(q >= 2 && q <= 256 ? q : 8);
    }();
    return v;
}

// `iters` fused sweeps (colour A then B each), ping-ponging between v.x and
// v.alt so the result lands in v.x; with `dot`, the last sweep also stores
// the r.z partials of the final field.
template <typename XT, typename WT>
void fused_sweeps(flux_multigrid* mg, int l, int iters, bool reversed, bool from_zero, const int* stop, DotOut* dot) {
    MGLevel& v = mg->lv[l];
    LevelDevT<WT> L = level_dev<WT>(mg, v);
    const int gx = (v.n[0] + Sweep3::TX - 1) / Sweep3::TX, gy = (v.n[1] + Sweep3::TY - 1) / Sweep3::TY;
    const int gxy = gx * gy;
    int gz = std::max(1, std::min<int>(v.n[2], (int)((g_march_blocks / 2 + gxy - 1) / gxy)));
    const int kz = (v.n[2] + gz - 1) / gz;
    gz = (v.n[2] + kz - 1) / kz;
    const unsigned nb = (unsigned)(gxy * gz);
    const float om = (float)mg->o.omega;
    const int colorA = reversed ? 1 : 0;
    if constexpr (std::is_same<XT, float>::value) {
        if (mg->fused == 2) {
            // the register form, finest level: whole rows of cell pairs
            const int kz = sweepr_chunk();
            const int gyr = (v.n[1] + SweepR::TY - 1) / SweepR::TY, gzr = (v.n[2] + kz - 1) / kz;
            const unsigned nbr = (unsigned)(gyr * gzr);
            const float* srcr = (const float*)v.x;
            if (!from_zero && (iters & 1)) {
                k_copy_t<float><<<blocks_for(cells_of(v.n), MTPB), MTPB, 0, g_stream>>>(
                    (const float*)v.x, (float*)v.alt, (int)cells_of(v.n), stop);
                srcr = (const float*)v.alt;
            }
            for (int it = 0; it < iters; ++it) {
                float* dst = (((iters - 1 - it) & 1) == 0) ? (float*)v.x : (float*)v.alt;
                const bool zero = from_zero && it == 0;
                const bool with_dot = dot && it == iters - 1 && nbr <= dot->capacity;
                kprof_begin("sweep");
                if (zero)
                    k_sweep3r<true, false, WT><<<nbr, SweepR::NT, 0, g_stream>>>(L, nullptr, (const float*)v.b, dst, om,
                                                                               colorA, kz, stop, nullptr);
                else if (with_dot)
                    k_sweep3r<false, true, WT><<<nbr, SweepR::NT, 0, g_stream>>>(L, srcr, (const float*)v.b, dst, om,
                                                                               colorA, kz, stop, dot->partials);
                else
                    k_sweep3r<false, false, WT><<<nbr, SweepR::NT, 0, g_stream>>>(L, srcr, (const float*)v.b, dst, om,
                                                                                colorA, kz, stop, nullptr);
                kprof_end("sweep");
                if (with_dot) dot->blocks = nbr;
                srcr = dst;
            }
            return;
        }
    }
    const XT* src = (const XT*)v.x;
    if (!from_zero && (iters & 1)) {
        k_copy_t<XT><<<blocks_for(cells_of(v.n), MTPB), MTPB, 0, g_stream>>>((const XT*)v.x, (XT*)v.alt,
                                                                              (int)cells_of(v.n), stop);
        src = (const XT*)v.alt;
    }
    for (int it = 0; it < iters; ++it) {
        XT* dst = (((iters - 1 - it) & 1) == 0) ? (XT*)v.x : (XT*)v.alt;
        const bool zero = from_zero && it == 0;
        const bool with_dot = dot && it == iters - 1 && nb <= dot->capacity;
        const bool fine = (l == 0);
        if (fine) kprof_begin("sweep");
#define FLUX_SWEEP(FINE_, ZERO_, DOT_)                                                                       \
    k_sweep3<FINE_, ZERO_, DOT_, XT, WT><<<nb, Sweep3::NT, 0, g_stream>>>(L, zero ? nullptr : src,            \
                                                                          (const XT*)v.b, dst, om, colorA, kz, \
                                                                          stop, with_dot ? dot->partials : nullptr)
        if (fine) {
            if (zero) FLUX_SWEEP(1, true, false);
            else if (with_dot) FLUX_SWEEP(1, false, true);
            else FLUX_SWEEP(1, false, false);
        } else {
            if (zero) FLUX_SWEEP(0, true, false);
            else FLUX_SWEEP(0, false, false);
        }
#undef FLUX_SWEEP
        if (with_dot) dot->blocks = nb;
        if (fine) kprof_end("sweep");
        src = dst;
    }
}

float smooth_omega(const flux_multigrid* mg) {
    return (float)(mg->o.smoother == FLUX_SMOOTH_JACOBI ? mg->o.jacobi_omega : mg->o.omega);
}

// `iters` damped-Jacobi sweeps (an even count): x -> buf -> x ...; the
// residual buffer is free while the level smooths.  With `dot`, the last
// sweep also stores the r.z partials of the final field.
template <int FINE, typename XT, typename WT>
void jacobi_sweeps(flux_multigrid* mg, int l, int iters, bool from_zero, const int* stop, DotOut* dot) {
    MGLevel& v = mg->lv[l];
    LevelDevT<WT> L = level_dev<WT>(mg, v);
    const float om = smooth_omega(mg);
    const bool pairs = mg->ndim == 3 && (v.n[0] & 1) == 0;
    const March M = pairs ? march_shape(v.n[0] / 2, v.n[1], v.n[2], MTPB, g_march_blocks) : March{};
    const unsigned g = blocks_for(cells_of(v.n), MTPB);
    for (int it = 0; it < iters; ++it) {
        const XT* src = (it & 1) ? (const XT*)v.buf : (const XT*)v.x;
        XT* dst = (it & 1) ? (XT*)v.x : (XT*)v.buf;
        const bool zero = from_zero && it == 0;
        const bool with_dot = dot && pairs && it == iters - 1 && M.nb <= dot->capacity;
        const char* tag = zero ? "jacobi_zero" : "jacobi";
        if (FINE) kprof_begin(tag);
        if (pairs) {
            const March Mf = FINE ? flip_march(M) : M;
            if (zero) k_jacobi3<FINE, 1, 0, XT, WT><<<M.nb, MTPB, 0, g_stream>>>(L, Mf, src, dst, (const XT*)v.b, om,
                                                                               stop, nullptr);
            else if (with_dot) k_jacobi3<FINE, 0, 1, XT, WT><<<M.nb, MTPB, 0, g_stream>>>(L, Mf, src, dst,
                                                                                         (const XT*)v.b, om, stop,
                                                                                         dot->partials);
            else k_jacobi3<FINE, 0, 0, XT, WT><<<M.nb, MTPB, 0, g_stream>>>(L, Mf, src, dst, (const XT*)v.b, om,
                                                                          stop, nullptr);
        } else {
            if (zero) k_jacobi<FINE, 1, XT, WT><<<g, MTPB, 0, g_stream>>>(L, src, dst, (const XT*)v.b, om, stop);
            else k_jacobi<FINE, 0, XT, WT><<<g, MTPB, 0, g_stream>>>(L, src, dst, (const XT*)v.b, om, stop);
        }
        if (FINE) kprof_end(tag);
        if (with_dot) dot->blocks = M.nb;
    }
}

template <int FINE, typename XT, typename WT>
void relax_sweeps(flux_multigrid* mg, int l, int iters, bool reversed, bool from_zero, const int* stop,
                  DotOut* dot = nullptr) {
    MGLevel& v = mg->lv[l];
    if (iters <= 0) return;
    if (mg->o.smoother == FLUX_SMOOTH_JACOBI) {
        jacobi_sweeps<FINE, XT, WT>(mg, l, iters, from_zero, stop, dot);
        return;
    }
    if (use_fused(mg, v)) {
        fused_sweeps<XT, WT>(mg, l, iters, reversed, from_zero, stop, dot);
        return;
    }
    LevelDevT<WT> L = level_dev<WT>(mg, v);
    float om = (float)mg->o.omega;
    XT* x = (XT*)v.x;
    const XT* b = (const XT*)v.b;
    if ((v.n[0] & 1) == 0) {
        const March M = march_shape(v.n[0] / 2, v.n[1], v.n[2], MTPB, g_march_blocks);
        for (int it = 0; it < iters; ++it) {
            int c0 = reversed ? 1 : 0, c1 = reversed ? 0 : 1;
            if (from_zero && it == 0) {
                k_relax3<FINE, 1, 0, XT, WT><<<M.nb, MTPB, 0, g_stream>>>(L, FINE ? flip_march(M) : M, x, b, om, c0,
                                                                          stop, nullptr);
            } else {
                if (FINE) kprof_begin("relax");
                k_relax3<FINE, 0, 0, XT, WT><<<M.nb, MTPB, 0, g_stream>>>(L, FINE ? flip_march(M) : M, x, b, om, c0,
                                                                          stop, nullptr);
                if (FINE) kprof_end("relax");
            }
            if (FINE) kprof_begin("relax");
            if (dot && it == iters - 1 && M.nb <= dot->capacity) {
                k_relax3<FINE, 0, 1, XT, WT><<<M.nb, MTPB, 0, g_stream>>>(L, FINE ? flip_march(M) : M, x, b, om, c1,
                                                                          stop, dot->partials);
                dot->blocks = M.nb;
            } else {
                k_relax3<FINE, 0, 0, XT, WT><<<M.nb, MTPB, 0, g_stream>>>(L, FINE ? flip_march(M) : M, x, b, om, c1,
                                                                          stop, nullptr);
            }
            if (FINE) kprof_end("relax");
        }
        return;
    }
    int64_t work = (int64_t)((v.n[0] + 1) / 2) * v.n[1] * v.n[2];
    unsigned g = blocks_for(work, MTPB);
    for (int it = 0; it < iters; ++it) {
        int c0 = reversed ? 1 : 0, c1 = reversed ? 0 : 1;
        if (from_zero && it == 0) {
            k_relax<FINE, 1, XT, WT><<<g, MTPB, 0, g_stream>>>(L, x, b, om, c0, stop);
        } else {
            if (FINE) kprof_begin("relax");
            k_relax<FINE, 0, XT, WT><<<g, MTPB, 0, g_stream>>>(L, x, b, om, c0, stop);
            if (FINE) kprof_end("relax");
        }
        if (FINE) kprof_begin("relax");
        k_relax<FINE, 0, XT, WT><<<g, MTPB, 0, g_stream>>>(L, x, b, om, c1, stop);
        if (FINE) kprof_end("relax");
    }
}

template <int FINE, typename XT, typename WT>
void level_smooth(flux_multigrid* mg, int l, int iters, bool reversed, bool from_zero, const int* stop,
                  DotOut* dot = nullptr) {
    relax_sweeps<FINE, XT, WT>(mg, l, iters, reversed, from_zero, stop, dot);
}

// Restriction of level v's residual `r` into level c's b.
template <int FINE, typename IT, typename OT>
void restrict_level(flux_multigrid* mg, const MGLevel& v, const MGLevel& c, const IT* r, const double* scale,
                    const int* stop) {
    if (mg->ndim == 3) {
        // fewer, longer columns: consecutive coarse planes share two fine planes.
        // The residual (k_residual3) is already 0 on masked fine cells.
        static const unsigned rdiv = [] {
            const char* e = getenv("FLUX_RESTRICT_DIV");
            const long v = e ? strtol(e, nullptr, 10) : 4;
            return (unsigned)(v >= 1 && v <= 64 ? v : 4);
        }();
        const March M = march_shape(c.n[0], c.n[1], c.n[2], MTPB, std::max(64u, g_march_blocks / rdiv));
        k_restrict3<FINE, IT, OT><<<M.nb, MTPB, 0, g_stream>>>(FINE ? flip_march(M) : M, v.n[0], v.n[1], v.n[2], r,
                                                                nullptr, c.n[0], c.n[1],
                                                                c.n[2], (OT*)c.b, c.mask, scale, stop);
    } else {
        k_restrict<FINE, IT, OT><<<blocks_for(cells_of(c.n), MTPB), MTPB, 0, g_stream>>>(
            mg->ndim, v.n[0], v.n[1], v.n[2], r, v.mask, c.n[0], c.n[1], c.n[2], (OT*)c.b, c.mask, scale, stop);
    }
}

// Prolongation of level c's x added into level v's x.
template <int FINE, typename CT, typename FT>
void correct_level(flux_multigrid* mg, const MGLevel& v, const MGLevel& c, const double* scale, const int* stop) {
    const float mult = (float)(c.xscale / v.xscale);  // stored coarse x to the fine level's storage scale
    if (mg->ndim == 3 && mg->pairs && (v.n[0] & 1) == 0) {
        const March M = march_shape(v.n[0] / 2, v.n[1], v.n[2], MTPB, g_march_blocks);
        k_correct3p<FINE, CT, FT><<<M.nb, MTPB, 0, g_stream>>>(FINE ? flip_march(M) : M, c.n[0], c.n[1], c.n[2],
                                                                (const CT*)c.x, v.n[0], v.n[1], v.n[2], (FT*)v.x,
                                                                v.mask, scale, mult, stop);
    } else if (mg->ndim == 3) {
        const March M = march_shape(v.n[0], v.n[1], v.n[2], MTPB, g_march_blocks);
        k_correct3<FINE, CT, FT><<<M.nb, MTPB, 0, g_stream>>>(FINE ? flip_march(M) : M, c.n[0], c.n[1], c.n[2],
                                                               (const CT*)c.x, v.n[0],
                                                               v.n[1], v.n[2], (FT*)v.x, v.mask, scale, mult, stop);
    } else {
        k_correct<FINE, CT, FT><<<blocks_for(cells_of(v.n), MTPB), MTPB, 0, g_stream>>>(
            mg->ndim, c.n[0], c.n[1], c.n[2], (const CT*)c.x, v.n[0], v.n[1], v.n[2], (FT*)v.x, v.mask, scale, mult,
            stop);
    }
}

template <typename ST>
int vcycle_t(flux_multigrid* mg, int l, const int* stop, const double* rr, DotOut* dot);

// The levels the brick smoother covers: every 3-D level but the coarsest
// with more than 4096 cells (the smaller ones run in one workgroup, in the
// V-cycle tail or the coarsest solve, and smooth globally).
bool brick_level(const flux_multigrid* mg, int l) {
    const MGLevel& v = mg->lv[l];
    return mg->o.smoother == FLUX_SMOOTH_BRICK && mg->ndim == 3 && l + 1 < (int)mg->lv.size() &&
           (v.n[0] & 1) == 0 && cells_of(v.n) > (int64_t)COARSE_LDS_CELLS;
}

// One brick pass of `sweeps` colour pairs from xin (nullptr: zero) into xout.
template <int FINE, typename XT, typename WT>
void brick_pass(flux_multigrid* mg, const MGLevel& v, bool shifted, const XT* xin, XT* xout, int sweeps,
                bool reversed, const int* stop, DotOut* dot) {
    const LevelDevT<WT> L = level_dev<WT>(mg, v);
    const BrickGrid G = brick_grid(v.n[0], v.n[1], v.n[2], shifted);
    const unsigned nb = (unsigned)G.nbx * G.nby * G.nbz;
    const float om = smooth_omega(mg);
    const int rev = reversed ? 1 : 0;
    const XT* b = (const XT*)v.b;
    // "brick" times the passes that read the field, "brick0" the first
    // pre-smoothing pass from zero
    const char* tag = xin ? "brick" : "brick0";
    if (FINE) kprof_begin(tag);
    const bool full = L.w0 && L.w1 && L.w2 && L.nb;
#define FLUX_BRICK(Z_, D_, F_)                                                                                 \
    k_brick3<Z_, D_, F_, XT, WT><<<nb, BRICK_TPB, 0, g_stream>>>(L, G, Z_ ? nullptr : xin, xout, b, om, sweeps, rev, \
                                                                 stop, D_ ? dot->partials : nullptr)
    const bool with_dot = xin && dot && nb <= dot->capacity;
    if (!xin) {
        if (full) FLUX_BRICK(1, 0, 1);
        else FLUX_BRICK(1, 0, 0);
    } else if (with_dot) {
        if (full) FLUX_BRICK(0, 1, 1);
        else FLUX_BRICK(0, 1, 0);
        dot->blocks = nb;
    } else {
        if (full) FLUX_BRICK(0, 0, 1);
        else FLUX_BRICK(0, 0, 0);
    }
#undef FLUX_BRICK
    if (FINE) kprof_end(tag);
}

// A V-cycle level smoothed brick-wise (FLUX_SMOOTH_BRICK).  Pre-smoothing is
// two passes from x = 0: bricks at the origin, then bricks shifted by half a
// brick (so every cell near a brick face of the first pass lies inside a
// brick of the second); post-smoothing runs the adjoint: shifted bricks then
// unshifted, colours reversed.  Each pass reads one buffer and writes
// another: x (first pre pass) -> alt (smoothed) -> residual in buf; after
// the correction alt -> buf -> x.
template <int FINE, typename XT, typename ST>
int vcycle_brick(flux_multigrid* mg, int l, const int* stop, const double* rr, DotOut* dot) {
    MGLevel& v = mg->lv[l];
    MGLevel& c = mg->lv[l + 1];
    XT* x = (XT*)v.x;
    XT* alt = (XT*)v.alt;
    XT* buf = (XT*)v.buf;
    const double* scale = FINE ? rr : nullptr;
    const int nu = mg->o.nu_down;
    if (nu > 0) {
        brick_pass<FINE, XT, ST>(mg, v, false, nullptr, x, nu, false, stop, nullptr);
        brick_pass<FINE, XT, ST>(mg, v, true, x, alt, nu, false, stop, nullptr);
    } else {
        fill_t<XT>(alt, cells_of(v.n), 0.0f);
    }
    // the residual folded in z into the restriction (FLUX_ZFOLD=0: stored
    // residual, then k_restrict3; read at every call so one process can
    // compare the two); the folded sums fill half of buf's bytes
    const char* zenv = getenv("FLUX_ZFOLD");
    const bool zfold = !(zenv && zenv[0] == '0') && (v.n[2] & 1) == 0 && c.n[2] * 2 == v.n[2];
    if (zfold) {
        if (FINE) kprof_begin("residual");
        March Mz = march_shape(v.n[0], v.n[1], v.n[2], MTPB, g_march_blocks);
        if (Mz.kz & 1) {  // even z-chunks: a chunk holds whole coarse planes
            Mz.kz += 1;
            Mz.nb = Mz.bxy * (unsigned)((v.n[2] + Mz.kz - 1) / Mz.kz);
        }
        float* rz = reinterpret_cast<float*>(buf);
        k_residual3z<XT, ST><<<Mz.nb, MTPB, 0, g_stream>>>(level_dev<ST>(mg, v), FINE ? flip_march(Mz) : Mz, alt,
                                                            (const XT*)v.b, rz, stop);
        if (FINE) kprof_end("residual");
        if (FINE) kprof_begin("restrict");
        const March Mc = march_shape(c.n[0], c.n[1], c.n[2], MTPB, g_march_blocks);
        k_restrict_xy<ST><<<Mc.nb, MTPB, 0, g_stream>>>(FINE ? flip_march(Mc) : Mc, v.n[0], v.n[1], v.n[2], rz,
                                                         c.n[0], c.n[1], (ST*)c.b, c.mask, scale, stop);
        if (FINE) kprof_end("restrict");
    } else {
        if (FINE) kprof_begin("residual");
        const March M = march_shape(v.n[0], v.n[1], v.n[2], MTPB, g_march_blocks);
        k_residual3<FINE, XT, XT, ST><<<M.nb, MTPB, 0, g_stream>>>(level_dev<ST>(mg, v), FINE ? flip_march(M) : M,
                                                                    alt, (const XT*)v.b, buf, stop);
        if (FINE) kprof_end("residual");
        if (FINE) kprof_begin("restrict");
        restrict_level<FINE, XT, ST>(mg, v, c, buf, scale, stop);
        if (FINE) kprof_end("restrict");
    }
    FLUX_TRY(vcycle_t<ST>(mg, l + 1, stop, rr, nullptr));
    // the correction adds into alt (correct_level works on v.x)
    void* xsave = v.x;
    v.x = alt;
    if (FINE) kprof_begin("correct");
    correct_level<FINE, ST, XT>(mg, v, c, scale, stop);
    if (FINE) kprof_end("correct");
    v.x = xsave;
    const int nf = FINE ? mg->o.n_final : mg->o.nu_up;
    if (nf > 0) {
        const bool rev = !mg->post_forward;
        brick_pass<FINE, XT, ST>(mg, v, rev, alt, buf, nf, rev, stop, nullptr);
        brick_pass<FINE, XT, ST>(mg, v, !rev, buf, x, nf, rev, stop, dot);
    } else {
        k_copy_t<XT><<<blocks_for(cells_of(v.n), MTPB), MTPB, 0, g_stream>>>(alt, x, (int)cells_of(v.n), stop);
    }
    return FLUX_OK;
}

// The finest level of the mixed form: r (FP32) is scaled into b~ by the
// first pre-smoothing launch, the level smooths, computes its residual and
// corrects in ST, and the last post-smoothing launch writes z in FP32 (and
// the r.z partials).  Red-black runs its colour pairs in place on x~;
// Jacobi's sweep counts are even, so the ping-pong between x~ and buf~ ends
// in x~ before the residual and before the last sweep.
template <typename ST>
int vcycle_fine16(flux_multigrid* mg, const int* stop, const double* rr, DotOut* dot) {
    MGLevel& v = mg->lv[0];
    MGLevel& c = mg->lv[1];
    LevelDevT<ST> L = level_dev<ST>(mg, v);
    const float om = smooth_omega(mg);
    const float xs = (float)v.xscale;
    const double* srr = (mg->o.storage == FLUX_STORE_F16) ? rr : nullptr;
    const float* r = (const float*)v.b;
    float* z = (float*)v.x;
    const March M = march_shape(v.n[0] / 2, v.n[1], v.n[2], MTPB, g_march_blocks);
    ST* x16 = (ST*)v.x16;
    ST* b16 = (ST*)v.b16;
    ST* alt = (ST*)v.buf16;
    const int nu = mg->o.nu_down;
    const bool rbgs = mg->o.smoother == FLUX_SMOOTH_RBGS;
    // red-black: colour pairs in place, post-smoothing reversed
    for (int it = 0; rbgs && it < nu; ++it) {
        kprof_begin(it == 0 ? "relax_io" : "relax");
        if (it == 0)
            k_relax3m<1, 0, ST><<<M.nb, MTPB, 0, g_stream>>>(L, flip_march(M), x16, nullptr, b16, r, nullptr, srr, xs,
                                                             om, 0, stop, nullptr);
        else
            k_relax3m<0, 0, ST><<<M.nb, MTPB, 0, g_stream>>>(L, flip_march(M), x16, b16, nullptr, nullptr, nullptr,
                                                             srr, xs, om, 0, stop, nullptr);
        kprof_end(it == 0 ? "relax_io" : "relax");
        kprof_begin("relax");
        k_relax3m<0, 0, ST><<<M.nb, MTPB, 0, g_stream>>>(L, flip_march(M), x16, b16, nullptr, nullptr, nullptr, srr,
                                                         xs, om, 1, stop, nullptr);
        kprof_end("relax");
    }
    for (int it = 0; !rbgs && it < nu; ++it) {
        const ST* src = (it & 1) ? alt : x16;
        ST* dst = (it & 1) ? x16 : alt;
        // "jacobi" times the steady-state sweeps, "jacobi_io" the first and last
        const char* tag = it == 0 ? "jacobi_io" : "jacobi";
        kprof_begin(tag);
        if (it == 0)
            k_jacobi3m<1, 0, ST><<<M.nb, MTPB, 0, g_stream>>>(L, flip_march(M), nullptr, dst, nullptr, b16, r, nullptr,
                                                              srr, xs, om, stop, nullptr);
        else
            k_jacobi3m<0, 0, ST><<<M.nb, MTPB, 0, g_stream>>>(L, flip_march(M), src, dst, b16, nullptr, nullptr,
                                                              nullptr, srr, xs, om, stop, nullptr);
        kprof_end(tag);
    }
    kprof_begin("residual");
    const March Mr = march_shape(v.n[0], v.n[1], v.n[2], MTPB, g_march_blocks);
    k_residual3<1, ST, ST, ST><<<Mr.nb, MTPB, 0, g_stream>>>(L, flip_march(Mr), x16, b16, alt, stop);
    kprof_end("residual");
    kprof_begin("restrict");
    restrict_level<1, ST, ST>(mg, v, c, alt, nullptr, stop);
    kprof_end("restrict");
    FLUX_TRY(vcycle_t<ST>(mg, 1, stop, rr, nullptr));
    // the correction adds into x~ (correct_level works on v.x)
    void* zsave = v.x;
    v.x = x16;
    kprof_begin("correct");
    correct_level<1, ST, ST>(mg, v, c, nullptr, stop);
    kprof_end("correct");
    v.x = zsave;
    const int nf = mg->o.n_final;
    for (int it = 0; rbgs && it < nf; ++it) {
        kprof_begin("relax");
        k_relax3m<0, 0, ST><<<M.nb, MTPB, 0, g_stream>>>(L, flip_march(M), x16, b16, nullptr, nullptr, nullptr, srr,
                                                         xs, om, 1, stop, nullptr);
        kprof_end("relax");
        const bool last = it + 1 == nf;
        kprof_begin(last ? "relax_io" : "relax");
        if (!last)
            k_relax3m<0, 0, ST><<<M.nb, MTPB, 0, g_stream>>>(L, flip_march(M), x16, b16, nullptr, nullptr, nullptr,
                                                             srr, xs, om, 0, stop, nullptr);
        else if (dot && M.nb <= dot->capacity) {
            k_relax3m<2, 1, ST><<<M.nb, MTPB, 0, g_stream>>>(L, flip_march(M), x16, nullptr, nullptr, r, z, srr, xs,
                                                             om, 0, stop, dot->partials);
            dot->blocks = M.nb;
        } else
            k_relax3m<2, 0, ST><<<M.nb, MTPB, 0, g_stream>>>(L, flip_march(M), x16, nullptr, nullptr, r, z, srr, xs,
                                                             om, 0, stop, nullptr);
        kprof_end(last ? "relax_io" : "relax");
    }
    for (int it = 0; !rbgs && it < nf; ++it) {
        const ST* src = (it & 1) ? alt : x16;
        ST* dst = (it & 1) ? x16 : alt;
        const char* tag = it + 1 < nf ? "jacobi" : "jacobi_io";
        kprof_begin(tag);
        if (it + 1 < nf) {
            k_jacobi3m<0, 0, ST><<<M.nb, MTPB, 0, g_stream>>>(L, flip_march(M), src, dst, b16, nullptr, nullptr,
                                                              nullptr, srr, xs, om, stop, nullptr);
        } else if (dot && M.nb <= dot->capacity) {
            k_jacobi3m<2, 1, ST><<<M.nb, MTPB, 0, g_stream>>>(L, flip_march(M), src, nullptr, nullptr, nullptr, r, z,
                                                              srr, xs, om, stop, dot->partials);
            dot->blocks = M.nb;
        } else {
            k_jacobi3m<2, 0, ST><<<M.nb, MTPB, 0, g_stream>>>(L, flip_march(M), src, nullptr, nullptr, nullptr, r, z,
                                                              srr, xs, om, stop, nullptr);
        }
        kprof_end(tag);
    }
    return FLUX_OK;
}

// vCycle (multigrid.js:1609-1642).  Level 0 has FP32 vectors and weights of
// type ST; coarse levels store everything in ST.
template <typename ST>
int vcycle_t(flux_multigrid* mg, int l, const int* stop, const double* rr, DotOut* dot) {
    if (l == 0 && mg->fine16 && !brick_level(mg, 0)) return vcycle_fine16<ST>(mg, stop, rr, dot);
    MGLevel& v = mg->lv[l];
    int nl = (int)mg->lv.size();
    const bool fine = (l == 0);
    const int jac = mg->o.smoother == FLUX_SMOOTH_JACOBI ? 1 : 0;
    if (l == nl - 1) {
        int64_t cells = cells_of(v.n);
        if (fine) {
            if (cells <= COARSE_SINGLE_BLOCK_MAX)
                k_coarse_solve<float, ST><<<1, COARSE_TPB, 0, g_stream>>>(level_dev<ST>(mg, v), (float*)v.x,
                                                                    (const float*)v.b, smooth_omega(mg),
                                                                    mg->o.n_coarsest, jac, stop);
            else if (mg->o.n_coarsest > 0)
                level_smooth<1, float, ST>(mg, l, mg->o.n_coarsest, false, true, stop);
            else
                fill_t<float>(v.x, cells, 0.0f);
        } else {
            if (cells <= COARSE_SINGLE_BLOCK_MAX)
                k_coarse_solve<ST, ST><<<1, COARSE_TPB, 0, g_stream>>>(level_dev<ST>(mg, v), (ST*)v.x, (const ST*)v.b,
                                                                 smooth_omega(mg), mg->o.n_coarsest, jac, stop);
            else if (mg->o.n_coarsest > 0)
                level_smooth<0, ST, ST>(mg, l, mg->o.n_coarsest, false, true, stop);
            else
                fill_t<ST>(v.x, cells, 0.0f);
        }
        return FLUX_OK;
    }
    if (brick_level(mg, l))
        return fine ? vcycle_brick<1, float, ST>(mg, l, stop, rr, dot) : vcycle_brick<0, ST, ST>(mg, l, stop, rr, dot);
    MGLevel& c = mg->lv[l + 1];
    int64_t cells = cells_of(v.n), cc = cells_of(c.n);
    const double* scale = fine ? rr : nullptr;
    if (!fine && l + 2 == nl && mg->ndim == 3 && cells <= COARSE_LDS_CELLS && cc <= COARSE_TPB &&
        !use_fused(mg, v) && mg->tail) {
        k_vcycle_tail<ST><<<1, COARSE_TPB, 0, g_stream>>>(level_dev<ST>(mg, v), level_dev<ST>(mg, c), (const ST*)v.b,
                                                          (ST*)v.x, smooth_omega(mg), mg->o.nu_down,
                                                          mg->o.nu_up, mg->o.n_coarsest,
                                                          (float)(c.xscale / v.xscale), jac, mg->post_forward ? 1 : 0,
                                                          stop);
        return FLUX_OK;
    }
    if (fine) {
        if (mg->o.nu_down > 0) level_smooth<1, float, ST>(mg, l, mg->o.nu_down, false, true, stop);
        else fill_t<float>(v.x, cells, 0.0f);
        kprof_begin("residual");
        const March M = march_shape(v.n[0], v.n[1], v.n[2], MTPB, g_march_blocks);
        k_residual3<1, float, float, ST><<<M.nb, MTPB, 0, g_stream>>>(level_dev<ST>(mg, v), flip_march(M),
                                                                      (const float*)v.x,
                                                                      (const float*)v.b, (float*)v.buf, stop);
        kprof_end("residual");
        kprof_begin("restrict");
        restrict_level<1, float, ST>(mg, v, c, (const float*)v.buf, scale, stop);
        kprof_end("restrict");
    } else {
        if (mg->o.nu_down > 0) level_smooth<0, ST, ST>(mg, l, mg->o.nu_down, false, true, stop);
        else fill_t<ST>(v.x, cells, 0.0f);
        const March M = march_shape(v.n[0], v.n[1], v.n[2], MTPB, g_march_blocks);
        k_residual3<0, ST, ST, ST><<<M.nb, MTPB, 0, g_stream>>>(level_dev<ST>(mg, v), M, (const ST*)v.x,
                                                                (const ST*)v.b, (ST*)v.buf, stop);
        restrict_level<0, ST, ST>(mg, v, c, (const ST*)v.buf, nullptr, stop);
    }
    FLUX_TRY(vcycle_t<ST>(mg, l + 1, stop, rr, nullptr));
    if (fine) {
        kprof_begin("correct");
        correct_level<1, ST, float>(mg, v, c, scale, stop);
        kprof_end("correct");
        level_smooth<1, float, ST>(mg, l, mg->o.n_final, !mg->post_forward, false, stop, dot);
    } else {
        correct_level<0, ST, ST>(mg, v, c, nullptr, stop);
        level_smooth<0, ST, ST>(mg, l, mg->o.nu_up, !mg->post_forward, false, stop);
    }
    return FLUX_OK;
}

}  // namespace

namespace flux {

int mg_refresh(flux_multigrid* mg) {
    switch (mg->o.storage) {
        case FLUX_STORE_F16: return refresh_t<__half>(mg);
        case FLUX_STORE_BF16: return refresh_t<__hip_bfloat16>(mg);
        default: return refresh_t<float>(mg);
    }
}

// z = M r on the stream.  `stop` (device) short-circuits every kernel; `rr`
// (device, may be null) is r.r, used for the FP16 residual scale.
int mg_apply_raw(flux_multigrid* mg, const float* r, float* z, const int* stop, const double* rr, DotOut* dot) {
    mg->lv[0].x = z;
    mg->lv[0].b = (void*)r;
    if (dot) dot->blocks = 0;
    switch (mg->o.storage) {
        case FLUX_STORE_F16: FLUX_TRY(vcycle_t<__half>(mg, 0, stop, rr, dot)); break;
        case FLUX_STORE_BF16: FLUX_TRY(vcycle_t<__hip_bfloat16>(mg, 0, stop, nullptr, dot)); break;
        default: FLUX_TRY(vcycle_t<float>(mg, 0, stop, nullptr, dot)); break;
    }
    FLUX_LAUNCH_CHECK();
    return FLUX_OK;
}

int64_t mg_cells(const flux_multigrid* mg) { return cells_of(mg->lv[0].n); }
int mg_storage(const flux_multigrid* mg) { return mg ? mg->o.storage : FLUX_STORE_F32; }

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
    o->smoother = FLUX_SMOOTH_RBGS;
    o->jacobi_omega = 2.0 / 3.0;
}

int flux_multigrid_destroy(flux_multigrid* mg) {
    if (!mg) return FLUX_OK;
    if (g_init) (void)hipStreamSynchronize(g_stream);
    for (size_t l = 0; l < mg->lv.size(); ++l) {
        MGLevel& v = mg->lv[l];
        dfree(v.buf);
        dfree(v.alt);
        dfree(v.nb);
        dfree(v.x16);
        dfree(v.b16);
        dfree(v.buf16);
        if (v.own) {
            dfree(v.x);
            dfree(v.b);
            dfree(v.mask);
        }
        if (v.own || v.own_w)
            for (int a = 0; a < 3; ++a) dfree(v.w[a]);
    }
    delete mg;
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
    FLUX_REQUIRE(o.storage == FLUX_STORE_F32 || o.storage == FLUX_STORE_F16 || o.storage == FLUX_STORE_BF16,
                 "multigrid: unknown storage %d", o.storage);
    FLUX_REQUIRE(o.nu_down >= 0 && o.nu_up >= 0 && o.n_coarsest >= 0 && o.n_final >= 0,
                 "multigrid: iteration counts must be >= 0");
    FLUX_REQUIRE(o.smoother == FLUX_SMOOTH_RBGS || o.smoother == FLUX_SMOOTH_JACOBI || o.smoother == FLUX_SMOOTH_BRICK,
                 "multigrid: unknown smoother %d", o.smoother);
    FLUX_REQUIRE(o.smoother != FLUX_SMOOTH_JACOBI ||
                     !((o.nu_down | o.nu_up | o.n_coarsest | o.n_final) & 1),
                 "multigrid (jacobi): sweep counts must be even so the ping-pong lands back in x, got %d/%d/%d/%d.",
                 o.nu_down, o.nu_up, o.n_coarsest, o.n_final);
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
    for (int a = 0; a < ndim; ++a) mg->w[a] = w ? w[a] : nullptr;
    mg->has_w = (w != nullptr);
    mg->has_masks = (mask != nullptr) || (w != nullptr);
    // the one-pass sweep moves half the bytes of two launches but runs at
    // ~2.1 TB/s (latency-bound z-march): 197 vs 2 x 110 us per sweep at 256^3
    // in FP32, ~1 % per projection, and slower with FP16 weights; opt-in
    // (FLUX_FUSED_SWEEP=1)
    const char* fenv = getenv("FLUX_FUSED_SWEEP");
    mg->fused = (fenv && (fenv[0] == '1' || fenv[0] == '2')) ? fenv[0] - '0' : 0;
    const char* tenv = getenv("FLUX_VCYCLE_TAIL");
    mg->tail = !(tenv && tenv[0] == '0');
    const char* penv = getenv("FLUX_CORRECT_PAIRS");
    mg->pairs = !(penv && penv[0] == '0');
    const char* pfenv = getenv("FLUX_MG_POST_FORWARD");
    mg->post_forward = pfenv && pfenv[0] == '1';
    const char* f16env = getenv("FLUX_FINE_STORAGE");
    mg->fine16 = o.storage != FLUX_STORE_F32 && ndim == 3 && o.levels > 1 && (n0[0] % 2) == 0 && o.nu_down >= 1 &&
                 o.n_final >= 1 && (f16env && f16env[0] == '1');
    mg->lv.resize(o.levels);
    const size_t es = store_size(o.storage);
    bool ok = true;
    for (int l = 0; l < o.levels; ++l) {
        MGLevel& v = mg->lv[l];
        for (int a = 0; a < 3; ++a) {
            v.n[a] = (a < ndim) ? (n0[a] >> l) : 1;
            v.h[a] = (a < ndim) ? spacing[a] * (double)(1 << l) : 1.0;
        }
        // FP16 coarse levels hold x / h_l^2 (operator scaled by h_l^2) so the
        // solution, of size h^2 |b|, stays clear of FP16's subnormal range
        v.xscale = (o.storage == FLUX_STORE_F16 && (l > 0 || mg->fine16)) ? v.h[0] * v.h[0] : 1.0;
        int64_t cells = cells_of(v.n);
        if (l == 0 && mg->fine16) {
            v.x16 = dmalloc(cells * es);
            v.b16 = dmalloc(cells * es);
            v.buf16 = dmalloc(cells * es);
            ok = ok && v.x16 && v.b16 && v.buf16;
        }
        size_t vs = (l == 0) ? 4 : es;  // finest vectors stay FP32
        v.buf = dmalloc(cells * vs);
        ok = ok && v.buf;
        const bool brick = o.smoother == FLUX_SMOOTH_BRICK && ndim == 3 && l + 1 < o.levels;
        if (ndim == 3 && (mg->fused == 1 || (mg->fused == 2 && l == 0) || brick)) {
            v.alt = dmalloc(cells * vs);
            ok = ok && v.alt;
        }
        if (mg->has_masks || mask) {
            v.nb = (uint8_t*)dmalloc(cells);
            ok = ok && v.nb;
        }
        if (l == 0) {
            v.mask = mask ? (uint8_t*)mask->d : nullptr;
            for (int a = 0; a < ndim; ++a) {
                if (!w) continue;
                if (o.storage == FLUX_STORE_F32) {
                    v.w[a] = w[a]->d;
                } else {
                    v.own_w = true;
                    v.w[a] = dmalloc(faces_of(v.n, a) * es);
                    ok = ok && v.w[a];
                }
            }
        } else {
            v.own = true;
            v.x = dmalloc(cells * vs);
            v.b = dmalloc(cells * vs);
            ok = ok && v.x && v.b;
            if (mg->has_masks) {
                v.mask = (uint8_t*)dmalloc(cells);
                ok = ok && v.mask;
                if (v.mask) fill_device(v.mask, FLUX_U8, cells, 0);
            }
            if (mg->has_w)
                for (int a = 0; a < ndim; ++a) {
                    v.w[a] = dmalloc(faces_of(v.n, a) * es);
                    ok = ok && v.w[a];
                }
        }
    }
    if (!ok) {
        flux_multigrid_destroy(mg);
        return set_error(FLUX_ERR_NOMEM, "multigrid: device allocation failed");
    }
    // coarse fields start at the reference's unrefreshed defaults (x, b 0; weights 1)
    for (int l = 1; l < o.levels; ++l) {
        MGLevel& v = mg->lv[l];
        int64_t cells = cells_of(v.n);
        for (void* p : {v.x, v.b}) {
            if (o.storage == FLUX_STORE_F32) fill_t<float>(p, cells, 0.0f);
            else if (o.storage == FLUX_STORE_F16) fill_t<__half>(p, cells, 0.0f);
            else fill_t<__hip_bfloat16>(p, cells, 0.0f);
        }
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
    g_march_toggle = 0;
    FLUX_TRY(mg_apply_raw(mg, (const float*)in->d, (float*)out->d, nullptr, nullptr, nullptr));
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
    if (mask_out && v.mask) {
        FLUX_HIP(hipMemcpyAsync(mask_out, v.mask, cells_of(v.n), hipMemcpyDeviceToHost, g_stream));
        FLUX_HIP(hipStreamSynchronize(g_stream));
    }
    if (w_out)
        for (int a = 0; a < mg->ndim; ++a) {
            if (!w_out[a] || !v.w[a]) continue;
            int64_t nf = faces_of(v.n, a);
            if (level == 0 && mg->o.storage == FLUX_STORE_F32) {
                FLUX_HIP(hipMemcpyAsync(w_out[a], v.w[a], nf * 4, hipMemcpyDeviceToHost, g_stream));
                FLUX_HIP(hipStreamSynchronize(g_stream));
                continue;
            }
            float* tmp = (float*)dmalloc(nf * 4);
            FLUX_REQUIRE(tmp, "multigrid: allocation failed");
            if (mg->o.storage == FLUX_STORE_F16)
                k_convert<__half, float><<<blocks_for(nf, MTPB), MTPB, 0, g_stream>>>((const __half*)v.w[a], tmp, (int)nf);
            else if (mg->o.storage == FLUX_STORE_BF16)
                k_convert<__hip_bfloat16, float><<<blocks_for(nf, MTPB), MTPB, 0, g_stream>>>(
                    (const __hip_bfloat16*)v.w[a], tmp, (int)nf);
            else
                k_convert<float, float><<<blocks_for(nf, MTPB), MTPB, 0, g_stream>>>((const float*)v.w[a], tmp, (int)nf);
            hipError_t e = hipMemcpyAsync(w_out[a], tmp, nf * 4, hipMemcpyDeviceToHost, g_stream);
            if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
            dfree(tmp);
            FLUX_HIP(e);
        }
    return FLUX_OK;
}

}  // extern "C"
