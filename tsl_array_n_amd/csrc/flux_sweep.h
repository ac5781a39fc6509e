// Fused red-black Gauss-Seidel sweep: both colours in one pass over HBM.
//
// Semantics are exactly two launches of buildRelaxKernel
// (packages/fluxflow/src/linalg/multigrid.js:626-640), colour A then colour B
// (A = 0 for pre-smoothing, A = 1 for the reversed post-smoothing,
// multigrid.js:1589-1607, 1640): every colour-A cell is updated from the old
// field, then every colour-B cell from the colour-A-updated field.  The
// arithmetic per cell is row_terms', in the same order as the two-launch
// kernels, so the result is bit-identical to them.
//
// The pass is out of place (xin -> xout), so no block ever reads a value
// another block has already overwritten.  A block owns a 64 x 8 column tile
// and a chunk of z-planes and marches along z with a one-plane lag:
//   step s: store the prefetched old plane s+1 (tile + 2-cell ring) and its
//           Dirichlet bits in LDS and start loading plane s+2 into registers;
//           colour A on plane s over the tile + 1-cell ring (the ring is
//           recomputed redundantly, so no neighbour block is needed);
//           colour B on plane s-1 over the tile, written out with the
//           colour-A values of plane s-1.
// Old planes rotate through 5 LDS slots and colour-A planes through 4, so a
// step needs two barriers: a thread still finishing step s-1's colour-B
// phase never reads a slot that step s overwrites.  Face weights, b and the
// centre's Dirichlet bit come from global memory (L2: the ring's belong to
// the neighbouring tiles, which the XCD-aware block order keeps on the same
// XCD at the same time).
#pragma once
#include "flux_stencil.h"
#include "flux_stencil3.h"

namespace flux {

struct Sweep3 {
    static constexpr int TX = 64, TY = 8, NT = TX * TY / 2;               // 256 threads
    static constexpr int R1X = TX + 2, R1Y = TY + 2, R1 = R1X * R1Y;     // tile + 1 ring
    static constexpr int R2X = TX + 4, R2Y = TY + 4, R2 = R2X * R2Y;     // tile + 2 ring
    static constexpr int Q2 = (R2 + NT - 1) / NT;                        // R2 slots per thread
    static constexpr int Q1 = (R1 + NT - 1) / NT;                        // R1 slots per thread
    static constexpr int QT = TX * TY / NT;                              // tile cells per thread
    static constexpr int SX = 5, SA = 4;                                 // LDS plane slots
};

// One relaxation of a cell from its (already Dirichlet-masked, 0 where
// absent) neighbour values.
template <typename WT>
__device__ __forceinline__ float relax_value(const LevelDevT<WT>& L, int i, int j, int k, int c, float xc, bool dir,
                                             const float xn[6], float bc, float omega) {
    float w[6];
    face_w3(L, i, j, k, c, w);
    const float ih2[3] = {L.ih2[0], L.ih2[1], L.ih2[2]};
    float sum, dsum;
    row_terms<float>(xc, xn, w, ih2, sum, dsum);
    const float d = dir ? -1.0f : dsum;
    const float ax = dir ? -xc : sum;
    return xc + ((bc - ax) / d) * omega;
}

template <int FINE, bool ZERO, bool DOT, typename XT, typename WT>
__global__ void __launch_bounds__(Sweep3::NT) k_sweep3(LevelDevT<WT> L, const XT* __restrict__ xin,
                                                       const XT* __restrict__ b, XT* __restrict__ xout, float omega,
                                                       int colorA, int kchunk, const int* __restrict__ stop,
                                                       double* __restrict__ partials) {
    using S = Sweep3;
    if (stop && *stop) return;
    __shared__ float sx[S::SX][S::R2];
    __shared__ uint8_t sm[S::SX][S::R2];
    __shared__ float sa[S::SA][S::R1];

    const int n0 = L.n0, n1 = L.n1, n2 = L.n2;
    const int plane = n0 * n1;
    const int gx = (n0 + S::TX - 1) / S::TX, gy = (n1 + S::TY - 1) / S::TY;
    const unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
    const int bx = (int)(lb % (unsigned)gx);
    const int by = (int)((lb / (unsigned)gx) % (unsigned)gy);
    const int bz = (int)(lb / (unsigned)(gx * gy));
    const int i0 = bx * S::TX, j0 = by * S::TY;
    const int k0 = bz * kchunk;
    const int k1 = min(k0 + kchunk, n2);
    const int tid = threadIdx.x;
    const uint8_t* __restrict__ mask = L.mask;

    // this thread's slots of the 2-ring plane (old x, Dirichlet bits)
    int g2[S::Q2];
    bool in2[S::Q2];
#pragma unroll
    for (int q = 0; q < S::Q2; ++q) {
        const int r = tid + q * S::NT;
        const int ry = r / S::R2X, rx = r - ry * S::R2X;
        const int i = i0 - 2 + rx, j = j0 - 2 + ry;
        in2[q] = r < S::R2 && i >= 0 && i < n0 && j >= 0 && j < n1;
        g2[q] = i + n0 * j;
    }
    float px[S::Q2];
    uint8_t pm[S::Q2];
    auto fetch = [&](int kk) {
        const bool ok = kk >= 0 && kk < n2;
#pragma unroll
        for (int q = 0; q < S::Q2; ++q) {
            float v = 0.0f;
            uint8_t m = 0;
            if (ok && in2[q]) {
                const int c = g2[q] + plane * kk;
                if (!ZERO) v = to_f32(xin[c]);
                if (mask) m = mask[c];
            }
            px[q] = v;
            pm[q] = m;
        }
    };
    auto xs = [](int kk) { return (kk + 2 * Sweep3::SX) % Sweep3::SX; };
    auto as = [](int kk) { return (kk + 2 * Sweep3::SA) & (Sweep3::SA - 1); };
    auto store = [&](int kk) {
        const int slot = xs(kk);
#pragma unroll
        for (int q = 0; q < S::Q2; ++q) {
            const int r = tid + q * S::NT;
            if (r < S::R2) {
                sx[slot][r] = px[q];
                sm[slot][r] = pm[q];
            }
        }
    };

    // prologue: planes k0-2 and k0-1 in LDS, plane k0 in flight
    fetch(k0 - 2);
    store(k0 - 2);
    fetch(k0 - 1);
    store(k0 - 1);
    fetch(k0);
    double dot = 0.0;

    for (int s = k0 - 1; s <= k1; ++s) {
        store(s + 1);
        fetch(s + 2);
        __syncthreads();
        // colour A on plane s over the tile + 1 ring
        if (s >= 0 && s < n2) {
            const int xm = xs(s - 1), x0 = xs(s), xp = xs(s + 1), a0 = as(s);
#pragma unroll
            for (int u = 0; u < S::Q1; ++u) {
                const int r1 = tid + u * S::NT;
                if (r1 >= S::R1) continue;
                const int ry = r1 / S::R1X, rx = r1 - ry * S::R1X;
                const int i = i0 - 1 + rx, j = j0 - 1 + ry;
                if (i < 0 || i >= n0 || j < 0 || j >= n1 || ((i + j + s) & 1) != colorA) continue;
                const int r2 = (rx + 1) + S::R2X * (ry + 1);
                const int c = i + n0 * (j + n1 * s);
                auto nb = [&](int slot, int q, bool present) {
                    return (present && !sm[slot][q]) ? sx[slot][q] : 0.0f;
                };
                float xn[6];
                xn[0] = nb(x0, r2 - 1, i > 0);
                xn[1] = nb(x0, r2 + 1, i < n0 - 1);
                xn[2] = nb(x0, r2 - S::R2X, j > 0);
                xn[3] = nb(x0, r2 + S::R2X, j < n1 - 1);
                xn[4] = nb(xm, r2, s > 0);
                xn[5] = nb(xp, r2, s < n2 - 1);
                const float xc = sx[x0][r2];
                // rounded through the storage type, as the two-launch form stores it
                sa[a0][r1] = to_f32(to_store<XT>(relax_value(L, i, j, s, c, xc, sm[x0][r2] != 0, xn, to_f32(b[c]),
                                                             omega)));
            }
        }
        __syncthreads();
        // colour B on plane s-1 over the tile, written out with colour A
        const int kb = s - 1;
        if (kb >= k0 && kb < k1) {
            const int xb = xs(kb), xbm = xs(kb - 1), xbp = xs(kb + 1);
            const int am = as(kb - 1), a0 = as(kb), ap = as(kb + 1);
#pragma unroll
            for (int u = 0; u < S::QT; ++u) {
                const int t = tid + u * S::NT;
                const int ty = t / S::TX, tx = t - ty * S::TX;
                const int i = i0 + tx, j = j0 + ty;
                if (i >= n0 || j >= n1) continue;
                const int r1 = (tx + 1) + S::R1X * (ty + 1);
                const int r2 = (tx + 2) + S::R2X * (ty + 2);
                const int c = i + n0 * (j + n1 * kb);
                float v;
                if (((i + j + kb) & 1) == colorA) {
                    v = sa[a0][r1];
                } else {
                    auto nb = [&](int aslot, int q1, int xslot, int q2, bool present) {
                        return (present && !sm[xslot][q2]) ? sa[aslot][q1] : 0.0f;
                    };
                    float xn[6];
                    xn[0] = nb(a0, r1 - 1, xb, r2 - 1, i > 0);
                    xn[1] = nb(a0, r1 + 1, xb, r2 + 1, i < n0 - 1);
                    xn[2] = nb(a0, r1 - S::R1X, xb, r2 - S::R2X, j > 0);
                    xn[3] = nb(a0, r1 + S::R1X, xb, r2 + S::R2X, j < n1 - 1);
                    xn[4] = nb(am, r1, xbm, r2, kb > 0);
                    xn[5] = nb(ap, r1, xbp, r2, kb < n2 - 1);
                    const float xc = sx[xb][r2];
                    v = relax_value(L, i, j, kb, c, xc, sm[xb][r2] != 0, xn, to_f32(b[c]), omega);
                }
                const XT o = to_store<XT>(v);
                xout[c] = o;
                if (DOT) dot += (double)to_f32(b[c]) * (double)to_f32(o);
            }
        }
    }
    if (DOT) store_block_partial(dot, partials);
}

}  // namespace flux
