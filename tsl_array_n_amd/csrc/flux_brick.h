// Brick-wise red-black Gauss-Seidel: the V-cycle's one-pass smoother.
//
// A level is cut into bricks of 64 x 8 x 8 cells (BRICK_X/Y/Z), optionally
// shifted by half a brick on every axis.  One workgroup owns one brick: it
// reads the brick's right-hand side, face weights and neighbour-mask bits
// once, turns them into per-cell coefficients, reads the field into
// registers and LDS once, runs `sweeps` colour pairs of the reference's
// relaxation (buildRelaxKernel, multigrid.js:626-640: the same rows, the same
// colours) on the brick's cells, and stores the field once.  A neighbour
// outside the brick is read from the input field and held at that value for
// the whole pass (block Gauss-Seidel across bricks, red-black inside), so the
// input and output fields are distinct buffers.  With `reversed` colour 1
// runs before colour 0: a post-smoothing pass is then the adjoint of the
// pre-smoothing pass on the same bricks, and the V-cycle stays symmetric.
//
// The row is the reference's, rearranged once per pass: with c_n = w_n/h_a^2
// for a free neighbour (0 for a Dirichlet or outside one) and the diagonal
// d = -sum of w/h^2 over the present faces (laplacianDiagonalAt),
//   x_c + ((b - A x)_c / d) omega = x_c + omega (b/d - sum_n (c_n/d) x_n - x_c),
// so an update is six multiply-adds on b' = b/d and e_n = c_n/d; a Dirichlet
// row has e = 0 and d = -1.  Cells outside the grid carry e = 0, b' = 0 and
// x = 0, which the update keeps at 0, so the sweep has no branches.
//
// Thread layout: 512 threads, thread (tx, ty, tz) = (t % 32, (t / 32) % 8,
// t / 256) owns the cell pair (2 tx, 2 tx + 1) of brick row ty in planes
// 4 tz .. 4 tz + 3: every lane updates one cell per plane in either colour.
// LDS holds the brick's field with a one-cell halo, even and odd cells in
// separate halves ([parity][plane][row][pair]), so each neighbour is one read
// of consecutive words across a wavefront.
#pragma once
#include <cstdlib>
#include <type_traits>

#include "flux_stencil3.h"

namespace flux {

static constexpr int BRICK_X = 64, BRICK_Y = 8, BRICK_Z = 8;
static constexpr int BRICK_TPB = 512;
static constexpr int BRICK_KT = 4;               // planes per thread
static constexpr int BRICK_P = BRICK_X / 2 + 2;  // pairs per LDS row: 32 + the two x-halo slots
static constexpr int BRICK_R = BRICK_Y + 2;      // LDS rows per plane: 8 + the two y-halo rows
static constexpr int BRICK_L = BRICK_Z + 2;      // LDS planes: 8 + the two z-halo planes

struct BrickGrid {
    int nbx, nby, nbz;  // bricks per axis
    int ox, oy, oz;     // shift of the brick origin (0 or half a brick)
    int order;          // block -> brick: 0 x fastest, 1 z fastest; +2: no XCD remap
    int wrap;           // shifted bricks wrap around the grid (see brick_grid)
};

// Shifted bricks on a grid of whole bricks wrap around it (wrap_ok): the
// brick at the origin takes the last half brick of each axis with the first
// half brick, so the shifted pass has as many full bricks as the unshifted
// one instead of partial ones at every face.  Its two (four, eight) parts
// are not coupled: a face on the grid's boundary has weight 0, so the
// coefficient of the neighbour across the seam is exactly 0 and every cell
// is updated as in its partial brick, bit for bit.  FLUX_BRICK_WRAP=0 keeps
// the partial bricks (read at every call).
inline BrickGrid brick_grid(int n0, int n1, int n2, bool shifted) {
    BrickGrid g;
    const char* we = getenv("FLUX_BRICK_WRAP");
    const bool wrap_ok = !(we && we[0] == '0');
    g.wrap = shifted && wrap_ok && n0 % BRICK_X == 0 && n1 % BRICK_Y == 0 && n2 % BRICK_Z == 0 ? 1 : 0;
    g.ox = shifted ? BRICK_X / 2 : 0;
    g.oy = shifted ? BRICK_Y / 2 : 0;
    g.oz = shifted ? BRICK_Z / 2 : 0;
    const int ex = g.wrap ? 0 : g.ox, ey = g.wrap ? 0 : g.oy, ez = g.wrap ? 0 : g.oz;
    g.nbx = (n0 + ex + BRICK_X - 1) / BRICK_X;
    g.nby = (n1 + ey + BRICK_Y - 1) / BRICK_Y;
    g.nbz = (n2 + ez + BRICK_Z - 1) / BRICK_Z;
    static const int order = [] {
        const char* e = getenv("FLUX_BRICK_ORDER");
        return e ? atoi(e) & 3 : 0;
    }();
    g.order = order;
    return g;
}

// FROM_ZERO: the input field is zero (no loads).  DOT: each block also
// stores its partial of b.x over its cells (the PCG's r.z).  FULL: the level
// has face weights on every axis and neighbour-mask bits (the pressure
// solver's levels), so every load is unconditional and all of them are in
// flight together; otherwise a missing array stands for unit weights or no
// Dirichlet cells.  XT: storage type of x and b (computation in FP32); WT:
// storage type of the face weights.
template <int FROM_ZERO, int DOT, int FULL, typename XT, typename WT>
__global__ void __launch_bounds__(BRICK_TPB) k_brick3(LevelDevT<WT> L, BrickGrid G, const XT* __restrict__ xin,
                                                      XT* __restrict__ xout, const XT* __restrict__ b, float omega,
                                                      int sweeps, int reversed, const int* __restrict__ stop,
                                                      double* __restrict__ partials) {
    if (stop && *stop) return;
    __shared__ float xs[2][BRICK_L][BRICK_R][BRICK_P];
    const int n0 = L.n0, n1 = L.n1, n2 = L.n2;
    const int64_t plane = (int64_t)n0 * n1;
    const unsigned q = (G.order & 2) ? blockIdx.x : xcd_remap(blockIdx.x, gridDim.x);
    int bi, bj, bk;
    if (G.order & 1) {
        bk = (int)(q % (unsigned)G.nbz);
        bi = (int)((q / (unsigned)G.nbz) % (unsigned)G.nbx);
        bj = (int)(q / ((unsigned)G.nbz * (unsigned)G.nbx));
    } else {
        bi = (int)(q % (unsigned)G.nbx);
        bj = (int)((q / (unsigned)G.nbx) % (unsigned)G.nby);
        bk = (int)(q / ((unsigned)G.nbx * (unsigned)G.nby));
    }
    const int lo0 = bi * BRICK_X - G.ox, lo1 = bj * BRICK_Y - G.oy, lo2 = bk * BRICK_Z - G.oz;
    const int t = threadIdx.x;
    const int tx = t & 31, ty = (t >> 5) & 7, tz = t >> 8;
    // grid coordinates of a brick position: modulo the grid when wrapping
    // (every position then lies inside it)
    const int wr = G.wrap;
    auto wx = [&](int i) { return wr ? (i + n0) % n0 : i; };
    auto wy = [&](int jv) { return wr ? (jv + n1) % n1 : jv; };
    auto wz = [&](int k) { return wr ? (k + n2) % n2 : k; };
    const int i0 = wx(lo0 + 2 * tx), j = wy(lo1 + ty);
    // first plane of this thread: its 4 planes never straddle the seam (a
    // multiple of 4, the seam a multiple of 8)
    const int kt0 = wz(lo2 + BRICK_KT * tz);
    // n0 is even, so a pair is either wholly inside the grid or wholly outside
    const bool col = i0 >= 0 && i0 < n0 && j >= 0 && j < n1;
    const int pp = tx + 1, jj = ty + 1;
    const int64_t c0 = (int64_t)i0 + (int64_t)n0 * j;  // pair index in plane 0

    // ---- every global load first, from clamped (always valid) addresses, so
    // all of them are in flight together; validity is applied afterwards
    const int i0c = min(max(i0, 0), n0 - 2);
    const int jc = min(max(j, 0), n1 - 1);
    // halo of the field: the z-halo plane of this thread's column (below the
    // brick for tz = 0, above it for tz = 1), one y-halo pair (2 rows x 32
    // pairs x 8 planes over the 512 threads) and, for t < 128, one x-halo
    // cell (cells lo0 - 1 and lo0 + 64, 8 rows x 8 planes each)
    const int kzh = wz(tz ? lo2 + BRICK_Z : lo2 - 1);
    const bool zok = !FROM_ZERO && col && kzh >= 0 && kzh < n2;
    const int yside = t >> 8, yp = t & 31, yk = (t >> 5) & 7;
    const int yj = wy(yside ? lo1 + BRICK_Y : lo1 - 1);
    const int yi0 = wx(lo0 + 2 * yp);
    const int ykg = wz(lo2 + yk);
    const bool yok = !FROM_ZERO && yj >= 0 && yj < n1 && yi0 >= 0 && yi0 < n0 && ykg >= 0 && ykg < n2;
    const int xsd = (t >> 6) & 1, xr = (t >> 3) & 7, xk = t & 7;
    const int xi = wx(xsd ? lo0 + BRICK_X : lo0 - 1);
    const int xjg = wy(lo1 + xr), xkg = wz(lo2 + xk);
    const bool xok = !FROM_ZERO && t < 128 && xi >= 0 && xi < n0 && xjg >= 0 && xjg < n1 && xkg >= 0 && xkg < n2;
    Pair2<XT> zh, yh;
    zh.lo = zh.hi = yh.lo = yh.hi = to_store<XT>(0.0f);
    XT xh1 = to_store<XT>(0.0f);
    if (!FROM_ZERO) {
        const int kc = min(max(kzh, 0), n2 - 1);
        zh = *reinterpret_cast<const Pair2<XT>*>(xin + (int64_t)i0c + (int64_t)n0 * (jc + (int64_t)n1 * kc));
        const int yjc = min(max(yj, 0), n1 - 1), yic = min(max(yi0, 0), n0 - 2);
        const int ykc = min(max(ykg, 0), n2 - 1);
        yh = *reinterpret_cast<const Pair2<XT>*>(xin + (int64_t)yic + (int64_t)n0 * (yjc + (int64_t)n1 * ykc));
        const int xic = min(max(xi, 0), n0 - 1), xjc = min(max(xjg, 0), n1 - 1);
        const int xkc = min(max(xkg, 0), n2 - 1);
        xh1 = xin[(int64_t)xic + (int64_t)n0 * (xjc + (int64_t)n1 * xkc)];
    }

    Pair2<XT> xv[BRICK_KT], bv[BRICK_KT];
    Pair2<uint8_t> mv[BRICK_KT];
    WT wa0[BRICK_KT], wa1[BRICK_KT], wa2[BRICK_KT];
    Pair2<WT> wyl[BRICK_KT], wyu[BRICK_KT], wzl[BRICK_KT], wzu[BRICK_KT];
    bool kin[BRICK_KT];
#pragma unroll
    for (int kk = 0; kk < BRICK_KT; ++kk) {
        const int k = kt0 + kk;
        kin[kk] = col && k >= 0 && k < n2;
        const int kc = min(max(k, 0), n2 - 1);
        const int64_t cp = (int64_t)i0c + (int64_t)n0 * (jc + (int64_t)n1 * kc);
        if (!FROM_ZERO) xv[kk] = *reinterpret_cast<const Pair2<XT>*>(xin + cp);
        bv[kk] = *reinterpret_cast<const Pair2<XT>*>(b + cp);
        if (FULL || L.nb) mv[kk] = *reinterpret_cast<const Pair2<uint8_t>*>(L.nb + cp);
        if (FULL || L.w0) {
            const int64_t f = (int64_t)i0c + (int64_t)(n0 + 1) * (jc + (int64_t)n1 * kc);
            wa0[kk] = L.w0[f];
            wa1[kk] = L.w0[f + 1];
            wa2[kk] = L.w0[f + 2];
        }
        if (FULL || L.w1) {
            const int64_t f = (int64_t)i0c + (int64_t)n0 * (jc + (int64_t)(n1 + 1) * kc);
            wyl[kk] = *reinterpret_cast<const Pair2<WT>*>(L.w1 + f);
            wyu[kk] = *reinterpret_cast<const Pair2<WT>*>(L.w1 + f + n0);
        }
        if (FULL || L.w2) {
            wzl[kk] = *reinterpret_cast<const Pair2<WT>*>(L.w2 + cp);
            wzu[kk] = *reinterpret_cast<const Pair2<WT>*>(L.w2 + cp + plane);
        }
    }
    // ---- the field in registers and LDS
    float xl[BRICK_KT], xh[BRICK_KT];
#pragma unroll
    for (int kk = 0; kk < BRICK_KT; ++kk) {
        xl[kk] = (!FROM_ZERO && kin[kk]) ? to_f32(xv[kk].lo) : 0.f;
        xh[kk] = (!FROM_ZERO && kin[kk]) ? to_f32(xv[kk].hi) : 0.f;
        xs[0][BRICK_KT * tz + kk + 1][jj][pp] = xl[kk];
        xs[1][BRICK_KT * tz + kk + 1][jj][pp] = xh[kk];
    }
    xs[0][tz ? BRICK_L - 1 : 0][jj][pp] = zok ? to_f32(zh.lo) : 0.f;
    xs[1][tz ? BRICK_L - 1 : 0][jj][pp] = zok ? to_f32(zh.hi) : 0.f;
    xs[0][yk + 1][yside ? BRICK_R - 1 : 0][yp + 1] = yok ? to_f32(yh.lo) : 0.f;
    xs[1][yk + 1][yside ? BRICK_R - 1 : 0][yp + 1] = yok ? to_f32(yh.hi) : 0.f;
    if (t < 128) {
        const float v = xok ? to_f32(xh1) : 0.f;
        if (xsd) xs[0][xk + 1][xr + 1][BRICK_P - 1] = v;
        else xs[1][xk + 1][xr + 1][0] = v;
    }

    // ---- per-cell coefficients: e[cell][n] = c_n / d (n = -x, +x, -y, +y,
    // -z, +z) and b' = b / d; cell 0 the even cell of the pair, 1 the odd
    float e0[BRICK_KT][6], e1[BRICK_KT][6];
    float bp0[BRICK_KT], bp1[BRICK_KT];
    float bo0[BRICK_KT], bo1[BRICK_KT];  // b itself (the r.z partials of a DOT pass)
    const float ih2[3] = {L.ih2[0], L.ih2[1], L.ih2[2]};
#pragma unroll
    for (int kk = 0; kk < BRICK_KT; ++kk) {
        const int k = kt0 + kk;
        const bool hnb = FULL || L.nb, h0 = FULL || L.w0, h1 = FULL || L.w1, h2 = FULL || L.w2;
        const unsigned m0 = hnb ? (unsigned)mv[kk].lo : 0u, m1 = hnb ? (unsigned)mv[kk].hi : 0u;
        float w0[6], w1[6];  // face weights of the even and odd cell, outside faces 0
        const float a0 = h0 ? to_f32(wa0[kk]) : 1.f, a1 = h0 ? to_f32(wa1[kk]) : 1.f;
        const float a2 = h0 ? to_f32(wa2[kk]) : 1.f;
        w0[0] = i0 == 0 ? 0.f : a0;
        w0[1] = a1;
        w1[0] = a1;
        w1[1] = i0 + 1 == n0 - 1 ? 0.f : a2;
        w0[2] = j == 0 ? 0.f : (h1 ? to_f32(wyl[kk].lo) : 1.f);
        w1[2] = j == 0 ? 0.f : (h1 ? to_f32(wyl[kk].hi) : 1.f);
        w0[3] = j == n1 - 1 ? 0.f : (h1 ? to_f32(wyu[kk].lo) : 1.f);
        w1[3] = j == n1 - 1 ? 0.f : (h1 ? to_f32(wyu[kk].hi) : 1.f);
        w0[4] = k == 0 ? 0.f : (h2 ? to_f32(wzl[kk].lo) : 1.f);
        w1[4] = k == 0 ? 0.f : (h2 ? to_f32(wzl[kk].hi) : 1.f);
        w0[5] = k == n2 - 1 ? 0.f : (h2 ? to_f32(wzu[kk].lo) : 1.f);
        w1[5] = k == n2 - 1 ? 0.f : (h2 ? to_f32(wzu[kk].hi) : 1.f);
        float d0 = 0.f, d1 = 0.f;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            d0 = d0 - (w0[2 * a] + w0[2 * a + 1]) * ih2[a];
            d1 = d1 - (w1[2 * a] + w1[2 * a + 1]) * ih2[a];
        }
        const bool dir0 = (m0 & 64u) != 0, dir1 = (m1 & 64u) != 0;
        const float id0 = dir0 ? -1.0f : 1.0f / d0;
        const float id1 = dir1 ? -1.0f : 1.0f / d1;
        const bool in = kin[kk];
#pragma unroll
        for (int a = 0; a < 6; ++a) {
            e0[kk][a] = (!in || dir0 || ((m0 >> a) & 1u)) ? 0.f : w0[a] * ih2[a >> 1] * id0;
            e1[kk][a] = (!in || dir1 || ((m1 >> a) & 1u)) ? 0.f : w1[a] * ih2[a >> 1] * id1;
        }
        bo0[kk] = in ? to_f32(bv[kk].lo) : 0.f;
        bo1[kk] = in ? to_f32(bv[kk].hi) : 0.f;
        bp0[kk] = in ? bo0[kk] * id0 : 0.f;
        bp1[kk] = in ? bo1[kk] * id1 : 0.f;
    }
    __syncthreads();

    // ---- the sweeps.  A cell's colour is (i + j + k) & 1 and i0 is even, so
    // on plane kk the colour-0 cell of the pair sits at position
    // sa = (j + k) & 1 and the colour-1 cell at 1 - sa, alternating from
    // plane to plane.  The pair is relabelled once with its coefficients
    // (A: colour 0, B: colour 1), so a colour phase updates A (or B) on all
    // four planes with no per-update selection: the pair partner and the
    // z-neighbours of A on plane kk are B on planes kk, kk - 1 and kk + 1
    // (registers; at the thread's end planes the z-halo or the next
    // thread's plane from LDS), the y- and side neighbours come from LDS.  A
    // phase reads every LDS neighbour of its four updates first, then
    // updates, then writes, so its reads are in flight together; it reads
    // only the other colour and writes only its own.
    const int kl0 = BRICK_KT * tz + 1;  // LDS plane slot of this thread's first plane
    float xa[BRICK_KT], xb[BRICK_KT], ba[BRICK_KT], bb[BRICK_KT];
    float ea[BRICK_KT][6], eb[BRICK_KT][6];
    int sa[BRICK_KT];
#pragma unroll
    for (int kk = 0; kk < BRICK_KT; ++kk) {
        sa[kk] = (j + kt0 + kk) & 1;
        const bool s = sa[kk] != 0;
        xa[kk] = s ? xh[kk] : xl[kk];
        xb[kk] = s ? xl[kk] : xh[kk];
        ba[kk] = s ? bp1[kk] : bp0[kk];
        bb[kk] = s ? bp0[kk] : bp1[kk];
#pragma unroll
        for (int a = 0; a < 6; ++a) {
            ea[kk][a] = s ? e1[kk][a] : e0[kk][a];
            eb[kk][a] = s ? e0[kk][a] : e1[kk][a];
        }
    }
    // one colour phase: C = 0 updates A, C = 1 updates B
    auto phase = [&](auto cc) {
        constexpr int C = decltype(cc)::value;
        float* xu = C ? xb : xa;  // the updated cells
        const float* xq = C ? xa : xb;  // their pair partners (the other colour)
        const float* bu = C ? bb : ba;
        float side[BRICK_KT], ym[BRICK_KT], yp[BRICK_KT];
#pragma unroll
        for (int kk = 0; kk < BRICK_KT; ++kk) {
            const int s = C ? 1 - sa[kk] : sa[kk];  // position of the updated cell
            const int kl = kl0 + kk;
            // the x-neighbour outside the pair: the odd cell of pair tx-1 for
            // the even cell, the even cell of pair tx+1 for the odd one
            side[kk] = xs[1 - s][kl][jj][pp + 2 * s - 1];
            ym[kk] = xs[s][kl][jj - 1][pp];
            yp[kk] = xs[s][kl][jj + 1][pp];
        }
        const int s0 = C ? 1 - sa[0] : sa[0], s3 = C ? 1 - sa[BRICK_KT - 1] : sa[BRICK_KT - 1];
        const float zlo = xs[s0][kl0 - 1][jj][pp];
        const float zhi = xs[s3][kl0 + BRICK_KT][jj][pp];
        float nv[BRICK_KT];
#pragma unroll
        for (int kk = 0; kk < BRICK_KT; ++kk) {
            const bool s = (C ? 1 - sa[kk] : sa[kk]) != 0;
            const float* e = C ? eb[kk] : ea[kk];
            const float xc = xu[kk], xp = xq[kk];
            const float zm = kk > 0 ? xq[kk - 1] : zlo;
            const float zp = kk + 1 < BRICK_KT ? xq[kk + 1] : zhi;
            const float xm = s ? xp : side[kk];
            const float xr = s ? side[kk] : xp;
            float acc = bu[kk];
            acc -= e[0] * xm;
            acc -= e[1] * xr;
            acc -= e[2] * ym[kk];
            acc -= e[3] * yp[kk];
            acc -= e[4] * zm;
            acc -= e[5] * zp;
            nv[kk] = xc + omega * (acc - xc);
        }
#pragma unroll
        for (int kk = 0; kk < BRICK_KT; ++kk) {
            const int s = C ? 1 - sa[kk] : sa[kk];
            xu[kk] = nv[kk];
            xs[s][kl0 + kk][jj][pp] = nv[kk];
        }
        __syncthreads();
    };
    for (int sw = 0; sw < sweeps; ++sw) {
        if (reversed) {
            phase(std::integral_constant<int, 1>());
            phase(std::integral_constant<int, 0>());
        } else {
            phase(std::integral_constant<int, 0>());
            phase(std::integral_constant<int, 1>());
        }
    }

    // ---- store the brick's cells (rounded to XT once)
    double dot = 0.0;
#pragma unroll
    for (int kk = 0; kk < BRICK_KT; ++kk) {
        const int k = kt0 + kk;
        if (!(col && k >= 0 && k < n2)) continue;
        const bool s = sa[kk] != 0;
        Pair2<XT> o;
        o.lo = to_store<XT>(s ? xb[kk] : xa[kk]);
        o.hi = to_store<XT>(s ? xa[kk] : xb[kk]);
        *reinterpret_cast<Pair2<XT>*>(xout + c0 + plane * k) = o;
        if (DOT) {
            dot += (double)bo0[kk] * (double)to_f32(o.lo);
            dot += (double)bo1[kk] * (double)to_f32(o.hi);
        }
    }
    if (DOT) store_block_partial(dot, partials);
}

}  // namespace flux
