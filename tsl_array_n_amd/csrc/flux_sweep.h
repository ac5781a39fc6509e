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


// ---------------------------------------------------------------------------
// Fused red-black sweep on cell pairs, rows whole (n0 <= 256, even).
//
// The same semantics and per-cell arithmetic as two k_relax3 launches on the
// finest level (colour A, then colour B from the colour-A field; with ZERO
// the field starts at 0): bit-identical to them.  Out of place, xin -> xout.
// A block owns a tile of whole rows (2 groups of RG rows, 128 cell pairs per
// row, one thread per pair and group) and a chunk of z-planes, and marches
// in z with a one-plane lag:
//   step s: colour A on plane s over its rows plus one ring row (group 0 the
//           row above the tile, group 1 the row below: recomputed, so no
//           neighbour tile is waited for), each value into an LDS plane of
//           colour-A results;
//           colour B on plane s-1 over the tile's rows, from the old field
//           and the colour-A planes s-2, s-1 and s in LDS, written out as
//           aligned pairs together with their colour-A partners.
// The old field comes from registers: each thread keeps the cell pairs of
// 7 rows (its colour-A rows and their y-neighbours) for planes s-1, s, s+1
// and has plane s+2's in flight.  Colour-A planes rotate through 4 LDS
// slots, so one barrier per step suffices (a thread still in step s-1's
// colour-B phase reads slots s-3..s-1; step s writes slot s).
struct SweepR {
    static constexpr int PX = 128;        // cell pairs per row
    static constexpr int NT = 2 * PX;     // threads per block
    static constexpr int RG = 4;          // colour-B rows per group
    static constexpr int TY = 2 * RG;     // rows per tile
    static constexpr int RA = RG + 1;     // colour-A rows per group (one ring row)
    static constexpr int WR = RA + 2;     // old-field rows per thread
    static constexpr int LR = TY + 2;     // colour-A rows of a plane in LDS
    static constexpr int SL = 4;          // LDS plane slots
};

// the old field's cell pairs of one plane for the WR rows of a thread's
// window, as separate low / high registers (no dynamic indexing)
struct PairRows {
    float lo[SweepR::WR];
    float hi[SweepR::WR];
};

template <bool ZERO, typename WT>
__device__ __forceinline__ void sweepr_load(const LevelDevT<WT>& L, const float* __restrict__ xin, int p, int jw,
                                            int k, PairRows& X) {
    const int n0 = L.n0, n1 = L.n1, n2 = L.n2;
#pragma unroll
    for (int q = 0; q < SweepR::WR; ++q) {
        const int j = jw + q;
        float2 v = make_float2(0.0f, 0.0f);
        if (!ZERO && 2 * p < n0 && j >= 0 && j < n1 && k >= 0 && k < n2)
            v = *reinterpret_cast<const float2*>(xin + 2 * p + n0 * (j + n1 * k));
        X.lo[q] = v.x;
        X.hi[q] = v.y;
    }
}

// colour B on plane kb for the group's RG rows; OFF = row index of its first
// colour-B row in the register window (2 for group 0, 1 for group 1)
template <int OFF, bool DOT, typename WT>
__device__ __forceinline__ void sweepr_colour_b(const LevelDevT<WT>& L, const float* __restrict__ b,
                                                float* __restrict__ xout, float omega, int colorA, int p, int j0,
                                                int jb, int kb, int sm, int s0, int sp,
                                                const float (&red)[SweepR::SL][SweepR::LR][SweepR::PX],
                                                const PairRows& Xo, double& dot) {
    const int n0 = L.n0, n1 = L.n1, n2 = L.n2;
    const uint8_t* __restrict__ nbits = L.nb;
    const float ih2[3] = {L.ih2[0], L.ih2[1], L.ih2[2]};
    // every operand of the group's rows first, so the loads of all rows are
    // in flight together (the stores below could otherwise hold them back)
    float w[SweepR::RG][6];
    float2 bp[SweepR::RG];
    unsigned bits[SweepR::RG];
#pragma unroll
    for (int t = 0; t < SweepR::RG; ++t) {
        const int j = jb + t;
        bits[t] = 0u;
        bp[t] = make_float2(0.0f, 0.0f);
#pragma unroll
        for (int a = 0; a < 6; ++a) w[t][a] = 0.0f;
        if (j >= n1 || 2 * p >= n0) continue;
        const int sb = (j + kb + colorA + 1) & 1;
        const int i = 2 * p + sb;
        const int cp = 2 * p + n0 * (j + n1 * kb);
        bits[t] = nbits ? nbits[cp + sb] : 0u;
        face_w3(L, i, j, kb, cp + sb, w[t]);
        bp[t] = *reinterpret_cast<const float2*>(b + cp);
    }
#pragma unroll
    for (int t = 0; t < SweepR::RG; ++t) {
        const int j = jb + t;
        if (j >= n1 || 2 * p >= n0) continue;
        const int lr = j - (j0 - 1);
        const int sb = (j + kb + colorA + 1) & 1;  // the colour-B cell of the pair
        const int i = 2 * p + sb;
        const int cp = 2 * p + n0 * (j + n1 * kb);
        const float xc = sb ? Xo.hi[OFF + t] : Xo.lo[OFF + t];
        const float part = red[s0][lr][p];  // the pair's colour-A cell (updated)
        float xo;                           // the x-neighbour outside the pair (colour A, updated)
        if (sb) xo = (i < n0 - 1) ? red[s0][lr][p + 1] : xc;
        else xo = (i > 0) ? red[s0][lr][p - 1] : xc;
        float xn[6];
        xn[0] = sb ? part : xo;
        xn[1] = sb ? xo : part;
        xn[2] = j > 0 ? red[s0][lr - 1][p] : xc;
        xn[3] = j < n1 - 1 ? red[s0][lr + 1][p] : xc;
        xn[4] = kb > 0 ? red[sm][lr][p] : xc;
        xn[5] = kb + 1 < n2 ? red[sp][lr][p] : xc;
#pragma unroll
        for (int a = 0; a < 6; ++a) xn[a] = ((bits[t] >> a) & 1u) ? 0.0f : xn[a];
        const float bc = sb ? bp[t].y : bp[t].x;
        float sum, dsum;
        row_terms<float>(xc, xn, w[t], ih2, sum, dsum);
        const bool dir = (bits[t] & 64u) != 0;
        const float d = dir ? -1.0f : dsum;
        const float ax = dir ? -xc : sum;
        const float nv = xc + ((bc - ax) / d) * omega;
        float2 o;
        o.x = sb ? part : nv;
        o.y = sb ? nv : part;
        *reinterpret_cast<float2*>(xout + cp) = o;
        if (DOT) {
            dot += (double)bp[t].x * (double)o.x;
            dot += (double)bp[t].y * (double)o.y;
        }
    }
}

template <bool ZERO, bool DOT, typename WT>
__global__ void __launch_bounds__(SweepR::NT) k_sweep3r(LevelDevT<WT> L, const float* __restrict__ xin,
                                                        const float* __restrict__ b, float* __restrict__ xout,
                                                        float omega, int colorA, int kchunk,
                                                        const int* __restrict__ stop, double* __restrict__ partials) {
    using S = SweepR;
    if (stop && *stop) return;
    __shared__ float red[S::SL][S::LR][S::PX];
    const int n0 = L.n0, n1 = L.n1, n2 = L.n2;
    const int gy = (n1 + S::TY - 1) / S::TY;
    const unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
    const int by = (int)(lb % (unsigned)gy), bz = (int)(lb / (unsigned)gy);
    const int j0 = by * S::TY;
    const int k0 = bz * kchunk, k1 = min(k0 + kchunk, n2);
    const int p = threadIdx.x % S::PX;
    const int grp = threadIdx.x / S::PX;  // uniform per wavefront
    const int jb = j0 + grp * S::RG;      // first colour-B row of the group
    const int ja = grp ? jb : jb - 1;     // first colour-A row
    const int jw = ja - 1;                // first row of the register window
    const int la = ja - (j0 - 1);         // LDS row of the first colour-A row
    const uint8_t* __restrict__ nbits = L.nb;
    const float ih2[3] = {L.ih2[0], L.ih2[1], L.ih2[2]};
    // old field: planes s-1, s, s+1 and s+2 (in flight)
    PairRows Xm, X0, Xp, Xq;
    sweepr_load<ZERO>(L, xin, p, jw, k0 - 2, Xm);
    sweepr_load<ZERO>(L, xin, p, jw, k0 - 1, X0);
    sweepr_load<ZERO>(L, xin, p, jw, k0, Xp);
    double dot = 0.0;
    for (int s = k0 - 1; s <= k1; ++s) {
        sweepr_load<ZERO>(L, xin, p, jw, s + 2, Xq);
        const int slot = (s - (k0 - 1)) & (S::SL - 1);
        // colour A on plane s
        if (s >= 0 && s < n2) {
#pragma unroll
            for (int r = 0; r < S::RA; ++r) {
                const int j = ja + r;
                if (j < 0 || j >= n1 || 2 * p >= n0) continue;
                const int sa = (j + s + colorA) & 1;  // the colour-A cell of the pair
                const int i = 2 * p + sa;
                const int cp = 2 * p + n0 * (j + n1 * s);
                const int c = cp + sa;
                const float xc = sa ? X0.hi[r + 1] : X0.lo[r + 1];
                const float part = sa ? X0.lo[r + 1] : X0.hi[r + 1];
                float xo;  // the x-neighbour outside the pair (colour B, old)
                if (ZERO) xo = 0.0f;
                else if (sa) xo = (i < n0 - 1) ? xin[c + 1] : xc;
                else xo = (i > 0) ? xin[c - 1] : xc;
                float xn[6];
                xn[0] = sa ? part : xo;
                xn[1] = sa ? xo : part;
                xn[2] = j > 0 ? (sa ? X0.hi[r] : X0.lo[r]) : xc;
                xn[3] = j < n1 - 1 ? (sa ? X0.hi[r + 2] : X0.lo[r + 2]) : xc;
                xn[4] = s > 0 ? (sa ? Xm.hi[r + 1] : Xm.lo[r + 1]) : xc;
                xn[5] = s + 1 < n2 ? (sa ? Xp.hi[r + 1] : Xp.lo[r + 1]) : xc;
                const unsigned bits = nbits ? nbits[c] : 0u;
#pragma unroll
                for (int a = 0; a < 6; ++a) xn[a] = ((bits >> a) & 1u) ? 0.0f : xn[a];
                float w[6];
                face_w3(L, i, j, s, c, w);
                const float bc = b[c];
                float sum, dsum;
                row_terms<float>(xc, xn, w, ih2, sum, dsum);
                const bool dir = (bits & 64u) != 0;
                const float d = dir ? -1.0f : dsum;
                const float ax = dir ? -xc : sum;
                red[slot][la + r][p] = xc + ((bc - ax) / d) * omega;
            }
        }
        __syncthreads();
        // colour B on plane s-1
        const int kb = s - 1;
        if (kb >= k0 && kb < k1) {
            const int sm = (slot + S::SL - 2) & (S::SL - 1), s0 = (slot + S::SL - 1) & (S::SL - 1);
            if (grp == 0)
                sweepr_colour_b<2, DOT>(L, b, xout, omega, colorA, p, j0, jb, kb, sm, s0, slot, red, Xm, dot);
            else
                sweepr_colour_b<1, DOT>(L, b, xout, omega, colorA, p, j0, jb, kb, sm, s0, slot, red, Xm, dot);
        }
        Xm = X0;
        X0 = Xp;
        Xp = Xq;
    }
    if (DOT) store_block_partial(dot, partials);
}

}  // namespace flux
