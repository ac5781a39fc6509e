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
// and a chunk of z-planes and marches along z with a one-plane lag (n0
// even, so the tile's cell pairs are whole):
//   step s: colour A on plane s over the tile + 1-cell ring (the ring is
//           recomputed redundantly, so no neighbour block is needed);
//           colour B on plane s-1 over the tile, written out as aligned cell
//           pairs together with the colour-A values of plane s-1.
// Every operand of step s+1 is loaded during step s: the old plane (tile +
// 2-cell ring) and its Dirichlet bits go to LDS; b and the six face weights
// of the cells a thread updates stay in its registers (each lane updates one
// colour-A cell of the ring region per pass, 330 of them in 2 passes, and
// exactly one colour-B cell of the tile).  Old planes rotate through 5 LDS
// slots and colour-A planes through 4, so a step needs two barriers: a
// thread still finishing step s-1's colour-B phase never reads a slot that
// step s overwrites.
#pragma once
#include "flux_stencil.h"
#include "flux_stencil3.h"

namespace flux {

struct Sweep3 {
    static constexpr int TX = 64, TY = 8, NT = TX * TY / 2;               // 256 threads
    static constexpr int R1X = TX + 2, R1Y = TY + 2, R1 = R1X * R1Y;     // tile + 1 ring
    static constexpr int R2X = TX + 4, R2Y = TY + 4, R2 = R2X * R2Y;     // tile + 2 ring
    static constexpr int Q2 = (R2 + NT - 1) / NT;                        // R2 slots per thread
    static constexpr int AROW = R1X / 2;                                 // colour-A cells per R1 row
    static constexpr int NA = AROW * R1Y;                                // colour-A cells of R1
    static constexpr int QA = (NA + NT - 1) / NT;                        // colour-A passes
    static constexpr int SX = 5, SA = 4;                                 // LDS plane slots
};

// One relaxation from the centre, the (Dirichlet-masked, 0 where absent)
// neighbour values and the six face weights (outside faces zero).
__device__ __forceinline__ float relax_with(const float ih2[3], const float w[6], float xc, bool dir,
                                            const float xn[6], float bc, float omega) {
    float sum, dsum;
    row_terms<float>(xc, xn, w, ih2, sum, dsum);
    const float d = dir ? -1.0f : dsum;
    const float ax = dir ? -xc : sum;
    return xc + ((bc - ax) / d) * omega;
}

// b and the six face weights of one cell (zeros where the cell is absent)
struct CellOps {
    float b;
    float w[6];
};

template <typename XT, typename WT>
__device__ __forceinline__ void load_ops(const LevelDevT<WT>& L, const XT* __restrict__ b, bool ok, int i, int j,
                                         int k, CellOps& o) {
    if (ok) {
        const int c = i + L.n0 * (j + L.n1 * k);
        o.b = to_f32(b[c]);
        face_w3(L, i, j, k, c, o.w);
    } else {
        o.b = 0.0f;
#pragma unroll
        for (int a = 0; a < 6; ++a) o.w[a] = 0.0f;
    }
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
    const int i0 = bx * S::TX, j0 = by * S::TY;  // both even
    const int k0 = bz * kchunk;
    const int k1 = min(k0 + kchunk, n2);
    const int tid = threadIdx.x;
    const uint8_t* __restrict__ mask = L.mask;
    const float ih2[3] = {L.ih2[0], L.ih2[1], L.ih2[2]};
    const int colorB = colorA ^ 1;

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
    auto fetch_x = [&](int kk) {
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
    auto store_x = [&](int kk) {
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
    // the colour-A cell of pass u at plane kk (R1 coordinates), the colour-B
    // cell of plane kk (tile coordinates)
    auto a_cell = [&](int u, int kk, int& rx, int& ry) {
        const int q = tid + u * S::NT;
        ry = q / S::AROW;
        rx = 2 * (q - ry * S::AROW) + ((colorA + ry + kk) & 1);  // R1 origin (i0-1, j0-1): parity of i+j unchanged
    };
    auto b_cell = [&](int kk, int& rx, int& ry) {
        ry = tid / (S::TX / 2);
        rx = 2 * (tid - ry * (S::TX / 2)) + ((colorB + ry + kk) & 1);
    };
    CellOps nA[S::QA], nB;
    float nBp = 0.0f;  // b of the colour-B cell's pair partner (the r.z of the final field)
    auto fetch_ops = [&](int kk) {
        // colour A at plane kk, colour B at plane kk-1
#pragma unroll
        for (int u = 0; u < S::QA; ++u) {
            int rx, ry;
            a_cell(u, kk, rx, ry);
            const int i = i0 - 1 + rx, j = j0 - 1 + ry;
            const bool ok = tid + u * S::NT < S::NA && kk >= 0 && kk < n2 && i >= 0 && i < n0 && j >= 0 && j < n1;
            load_ops<XT, WT>(L, b, ok, i, j, kk, nA[u]);
        }
        const int kb = kk - 1;
        int rx, ry;
        b_cell(kb, rx, ry);
        const int i = i0 + rx, j = j0 + ry;
        const bool ok = kb >= k0 && kb < k1 && i < n0 && j < n1;
        load_ops<XT, WT>(L, b, ok, i, j, kb, nB);
        if (DOT) nBp = ok ? to_f32(b[(i ^ 1) + n0 * (j + n1 * kb)]) : 0.0f;
    };

    // prologue: planes k0-2 and k0-1 in LDS, plane k0 and the first step's operands in flight
    fetch_x(k0 - 2);
    store_x(k0 - 2);
    fetch_x(k0 - 1);
    store_x(k0 - 1);
    fetch_x(k0);
    fetch_ops(k0 - 1);
    double dot = 0.0;

    for (int s = k0 - 1; s <= k1; ++s) {
        CellOps cA[S::QA], cB;
#pragma unroll
        for (int u = 0; u < S::QA; ++u) cA[u] = nA[u];
        cB = nB;
        const float cBp = nBp;
        store_x(s + 1);
        fetch_x(s + 2);
        if (s + 1 <= k1) fetch_ops(s + 1);
        __syncthreads();
        // colour A on plane s over the tile + 1 ring
        if (s >= 0 && s < n2) {
            const int xm = xs(s - 1), x0 = xs(s), xp = xs(s + 1), a0 = as(s);
#pragma unroll
            for (int u = 0; u < S::QA; ++u) {
                if (tid + u * S::NT >= S::NA) continue;
                int rx, ry;
                a_cell(u, s, rx, ry);
                const int i = i0 - 1 + rx, j = j0 - 1 + ry;
                if (i < 0 || i >= n0 || j < 0 || j >= n1) continue;
                const int r2 = (rx + 1) + S::R2X * (ry + 1);
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
                sa[a0][rx + S::R1X * ry] =
                    to_f32(to_store<XT>(relax_with(ih2, cA[u].w, xc, sm[x0][r2] != 0, xn, cA[u].b, omega)));
            }
        }
        __syncthreads();
        // colour B on plane s-1 over the tile, written out with its colour-A partner
        const int kb = s - 1;
        if (kb >= k0 && kb < k1) {
            int rx, ry;
            b_cell(kb, rx, ry);
            const int i = i0 + rx, j = j0 + ry;
            if (i < n0 && j < n1) {
                const int xb = xs(kb), xbm = xs(kb - 1), xbp = xs(kb + 1);
                const int am = as(kb - 1), a0 = as(kb), ap = as(kb + 1);
                const int r1 = (rx + 1) + S::R1X * (ry + 1);
                const int r2 = (rx + 2) + S::R2X * (ry + 2);
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
                const float vb = relax_with(ih2, cB.w, xc, sm[xb][r2] != 0, xn, cB.b, omega);
                const XT ob = to_store<XT>(vb);
                const bool odd = (rx & 1) != 0;
                const int pr1 = odd ? r1 - 1 : r1 + 1;  // the colour-A partner in the pair
                const XT oa = to_store<XT>(sa[a0][pr1]);
                // n0 is even: the pair (i & ~1, i | 1) lies inside the row
                const int cp = (i & ~1) + n0 * (j + n1 * kb);
                Pair2<XT> o;
                o.lo = odd ? oa : ob;
                o.hi = odd ? ob : oa;
                *reinterpret_cast<Pair2<XT>*>(xout + cp) = o;
                if (DOT) {
                    const float blo = odd ? cBp : cB.b, bhi = odd ? cB.b : cBp;
                    dot += (double)blo * (double)to_f32(o.lo);
                    dot += (double)bhi * (double)to_f32(o.hi);
                }
            }
        }
    }
    if (DOT) store_block_partial(dot, partials);
}

}  // namespace flux
