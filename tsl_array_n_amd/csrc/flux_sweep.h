// Fused red-black Gauss-Seidel sweep: both colours in one pass over HBM.
//
// Semantics are exactly two launches of buildRelaxKernel
// (packages/fluxflow/src/linalg/multigrid.js:626-640), colour A then colour B
// (A = 0 for pre-smoothing, A = 1 for the reversed post-smoothing,
// multigrid.js:1589-1607, 1640): every colour-A cell is updated from the old
// field, then every colour-B cell from the colour-A-updated field.
//
// The pass is out of place (xin -> xout), so no block ever reads a value
// another block has already overwritten.  A block owns a TX x TY column tile
// and a chunk of z-planes and marches along z with a one-plane lag:
//   step s: stage old plane s+1 (tile + 2-cell ring) in LDS,
//           colour A on plane s over the tile + 1-cell ring (the ring is
//           recomputed redundantly so no neighbour block is needed),
//           colour B on plane s-1 over the tile, written out with the
//           colour-A values of plane s-1.
// Rolling LDS buffers hold four planes of old x and mask and four planes of
// the colour-A state.
#pragma once
#include "flux_stencil.h"

namespace flux {

template <int TX, int TY>
struct SweepTile {
    static constexpr int R2X = TX + 4, R2Y = TY + 4, R2 = R2X * R2Y;  // tile + 2 ring
    static constexpr int R1X = TX + 2, R1Y = TY + 2, R1 = R1X * R1Y;  // tile + 1 ring
};

template <int FINE, int TX, int TY, bool ZERO, typename XT, typename WT>
__global__ void __launch_bounds__(TX* TY) k_rbgs_fused(LevelDevT<WT> L, const XT* __restrict__ xin,
                                                       const XT* __restrict__ b, XT* __restrict__ xout,
                                                       float omega, int colorA, int kchunk,
                                                       const int* __restrict__ stop) {
    using T = SweepTile<TX, TY>;
    if (stop && *stop) return;
    __shared__ float s_xo[4][T::R2];
    __shared__ uint8_t s_m[4][T::R2];
    __shared__ float s_xa[4][T::R1];

    const int n0 = L.n0, n1 = L.n1, n2 = L.n2;
    const int gx = (n0 + TX - 1) / TX, gy = (n1 + TY - 1) / TY;
    const unsigned nb = gridDim.x;
    const unsigned lb = xcd_remap(blockIdx.x, nb);
    const int bx = lb % gx;
    const int by = (lb / gx) % gy;
    const int bz = lb / (gx * gy);
    const int i0 = bx * TX, j0 = by * TY;
    const int k0 = bz * kchunk;
    const int k1 = min(k0 + kchunk, n2);
    const int tid = threadIdx.x;
    const int plane = n0 * n1;
    const uint8_t* __restrict__ mask = L.mask;
    const WT* __restrict__ w0 = L.w0;
    const WT* __restrict__ w1 = L.w1;
    const WT* __restrict__ w2 = L.w2;
    const float ihx = L.ih2[0], ihy = L.ih2[1], ihz = L.ih2[2];
    const int colorB = colorA ^ 1;  // ZERO: the input field is zero (xin unused)

    auto load_plane = [&](int k) {
        const int slot = k & 3;
        for (int t = tid; t < T::R2; t += TX * TY) {
            int ry = t / T::R2X, rx = t - ry * T::R2X;
            int i = i0 - 2 + rx, j = j0 - 2 + ry;
            float v = 0.0f;
            uint8_t mv = 0;
            if (k >= 0 && k < n2 && i >= 0 && i < n0 && j >= 0 && j < n1) {
                int c = i + n0 * (j + n1 * k);
                if constexpr (!ZERO) v = to_f32(xin[c]);
                if (mask) mv = mask[c];
            }
            s_xo[slot][t] = v;
            s_m[slot][t] = mv;
        }
    };

    // one relaxation of cell (i,j,k) reading neighbour values through get(dx,dy,dz)
    // and masks through mget(dx,dy,dz) (offsets in R2 coordinates)
    auto relax_cell = [&](int i, int j, int k, float center, uint8_t mc, auto get, auto mget) -> float {
        int c = i + n0 * (j + n1 * k);
        if (mc) {
            // Dirichlet row: -x = b  ->  x += ((b + x) / -1) * omega
            return center + ((to_f32(b[c]) + center) / -1.0f) * omega;
        }
        float sum = 0.0f, dsum = 0.0f;
        // x axis
        {
            float dLo = 0.0f, dUp = 0.0f, hLo = 0.0f, hUp = 0.0f;
            int fl = i + (n0 + 1) * (j + n1 * k);
            if (i > 0) { dLo = mget(-1, 0, 0) ? center : center - get(-1, 0, 0); hLo = 1.0f; }
            if (i < n0 - 1) { dUp = mget(1, 0, 0) ? -center : get(1, 0, 0) - center; hUp = 1.0f; }
            if (w0) { float wl = to_f32(w0[fl]), wu = to_f32(w0[fl + 1]); dLo *= wl; dUp *= wu; hLo *= wl; hUp *= wu; }
            sum = sum + (dUp - dLo) * ihx;
            dsum = dsum - (hLo + hUp) * ihx;
        }
        {
            float dLo = 0.0f, dUp = 0.0f, hLo = 0.0f, hUp = 0.0f;
            int fl = i + n0 * (j + (n1 + 1) * k);
            if (j > 0) { dLo = mget(0, -1, 0) ? center : center - get(0, -1, 0); hLo = 1.0f; }
            if (j < n1 - 1) { dUp = mget(0, 1, 0) ? -center : get(0, 1, 0) - center; hUp = 1.0f; }
            if (w1) { float wl = to_f32(w1[fl]), wu = to_f32(w1[fl + n0]); dLo *= wl; dUp *= wu; hLo *= wl; hUp *= wu; }
            sum = sum + (dUp - dLo) * ihy;
            dsum = dsum - (hLo + hUp) * ihy;
        }
        {
            float dLo = 0.0f, dUp = 0.0f, hLo = 0.0f, hUp = 0.0f;
            if (k > 0) { dLo = mget(0, 0, -1) ? center : center - get(0, 0, -1); hLo = 1.0f; }
            if (k < n2 - 1) { dUp = mget(0, 0, 1) ? -center : get(0, 0, 1) - center; hUp = 1.0f; }
            if (w2) { float wl = to_f32(w2[c]), wu = to_f32(w2[c + plane]); dLo *= wl; dUp *= wu; hLo *= wl; hUp *= wu; }
            sum = sum + (dUp - dLo) * ihz;
            dsum = dsum - (hLo + hUp) * ihz;
        }
        return center + ((to_f32(b[c]) - sum) / dsum) * omega;
    };

    // prologue: old planes k0-2, k0-1, k0 (out-of-range planes read as zero)
    load_plane(k0 - 2);
    load_plane(k0 - 1);
    load_plane(k0);
    for (int s = k0 - 1; s <= k1; ++s) {
        load_plane(s + 1);
        __syncthreads();
        // colour A on plane s, tile + 1 ring -> s_xa[s & 3] (colour-B cells keep the old value)
        if (s >= 0 && s < n2) {
            const int so = s & 3;
            for (int t = tid; t < T::R1; t += TX * TY) {
                int ry = t / T::R1X, rx = t - ry * T::R1X;
                int i = i0 - 1 + rx, j = j0 - 1 + ry;
                float v = 0.0f;
                if (i >= 0 && i < n0 && j >= 0 && j < n1) {
                    int r2 = (rx + 1) + T::R2X * (ry + 1);
                    float center = s_xo[so][r2];
                    v = center;
                    if (((i + j + s) & 1) == colorA) {
                        auto get = [&](int dx, int dy, int dz) {
                            return s_xo[(s + dz) & 3][r2 + dx + T::R2X * dy];
                        };
                        auto mget = [&](int dx, int dy, int dz) {
                            return s_m[(s + dz) & 3][r2 + dx + T::R2X * dy];
                        };
                        v = relax_cell(i, j, s, center, s_m[so][r2], get, mget);
                    }
                }
                s_xa[so][t] = v;
            }
        }
        __syncthreads();
        // colour B on plane s-1 over the tile; write the plane out
        const int kb = s - 1;
        if (kb >= k0 && kb < k1) {
            const int sb = kb & 3;
            for (int t = tid; t < TX * TY; t += TX * TY) {
                int ty = t / TX, tx = t - ty * TX;
                int i = i0 + tx, j = j0 + ty;
                if (i < n0 && j < n1) {
                    int r1 = (tx + 1) + T::R1X * (ty + 1);
                    int r2 = (tx + 2) + T::R2X * (ty + 2);
                    float center = s_xa[sb][r1];
                    float v = center;
                    if (((i + j + kb) & 1) == colorB) {
                        auto get = [&](int dx, int dy, int dz) {
                            return s_xa[(kb + dz) & 3][r1 + dx + T::R1X * dy];
                        };
                        auto mget = [&](int dx, int dy, int dz) {
                            return s_m[(kb + dz) & 3][r2 + dx + T::R2X * dy];
                        };
                        v = relax_cell(i, j, kb, center, s_m[sb][r2], get, mget);
                    }
                    xout[i + n0 * (j + n1 * kb)] = to_store<XT>(v);
                }
            }
        }
        __syncthreads();
    }
}

}  // namespace flux
