// 3-D kernels' shared pieces: column-march geometry, pair loads and the
// operator row evaluated branch-free.
//
// The row is the one of stencil_row (laplacianAt / laplacianDiagonalAt,
// multigrid.js:273-339, 399-437) with the same operations in the same order:
// per axis dLo = (x_c - x_lo') * w_lo', dUp = (x_up' - x_c) * w_up',
// sum += (dUp - dLo) / h^2, diag -= (w_lo' + w_up') / h^2, where x_n' is 0
// for a Dirichlet neighbour and w' is 0 for a face outside the grid (so an
// outside face contributes an exact 0).  Neighbour reads use clamped,
// always-valid indices, so the loads are issued unconditionally.
//
// Launch shape: a 1-D grid; block b (after the XCD remap) covers `tpb`
// consecutive threads of one z-chunk; thread t of the chunk owns position
// t of the (x-run, y) plane and marches over the chunk's planes, so the
// cell index needs one division per thread instead of two per cell.
#pragma once
#include <algorithm>

#include "flux_internal.h"
#include "flux_stencil.h"

namespace flux {

template <typename T>
struct alignas(2 * sizeof(T)) Pair2 {
    T lo, hi;
};

struct March {
    int run;        // threads per row of the plane
    int rows;       // rows per plane (n1)
    int kz;         // planes per chunk
    int n2;         // planes
    unsigned bxy;   // blocks per chunk
    unsigned nb;    // blocks in the grid
    unsigned rev;   // walk the z-chunks from the last one down
};

// `want` is the block count aimed at (several waves of the 256 CUs).
inline March march_shape(int run, int n1, int n2, int tpb, unsigned want) {
    March m;
    m.run = run;
    m.rows = n1;
    m.n2 = n2;
    const int64_t per_plane = (int64_t)run * n1;
    m.bxy = (unsigned)((per_plane + tpb - 1) / tpb);
    unsigned bz = (want + m.bxy - 1) / m.bxy;
    bz = std::max(1u, std::min<unsigned>(bz, (unsigned)n2));
    m.kz = (n2 + (int)bz - 1) / (int)bz;
    bz = (unsigned)((n2 + m.kz - 1) / m.kz);
    m.nb = m.bxy * bz;
    m.rev = 0;
    return m;
}

// The next finest-level launch's march: the z-chunk order flips every call.
inline March flip_march(March m) {
    m.rev = g_march_flip ? (g_march_toggle ^= 1u) : 0u;
    return m;
}

// Thread position in the plane and its plane range; false when the thread
// lies past the plane's end.
__device__ __forceinline__ bool march_begin(const March& M, int& t, int& k0, int& k1) {
    const unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
    unsigned bz = lb / M.bxy;
    const unsigned bxy = lb - bz * M.bxy;
    if (M.rev) bz = M.nb / M.bxy - 1u - bz;
    t = (int)(bxy * blockDim.x + threadIdx.x);
    k0 = (int)bz * M.kz;
    k1 = min(k0 + M.kz, M.n2);
    return t < M.run * M.rows;
}

// Face weights of cell (i,j,k) with outside faces zeroed: w[2a] lower and
// w[2a+1] upper face on axis a.
template <typename WT>
__device__ __forceinline__ void face_w3(const LevelDevT<WT>& L, int i, int j, int k, int c, float w[6]) {
    const int n0 = L.n0, n1 = L.n1, n2 = L.n2;
    const int plane = n0 * n1;
    w[0] = w[1] = w[2] = w[3] = w[4] = w[5] = 1.0f;
    if (L.w0) {
        const int f = i + (n0 + 1) * (j + n1 * k);
        w[0] = to_f32(L.w0[f]);
        w[1] = to_f32(L.w0[f + 1]);
    }
    if (L.w1) {
        const int f = i + n0 * (j + (n1 + 1) * k);
        w[2] = to_f32(L.w1[f]);
        w[3] = to_f32(L.w1[f + n0]);
    }
    if (L.w2) {
        w[4] = to_f32(L.w2[c]);
        w[5] = to_f32(L.w2[c + plane]);
    }
    if (i == 0) w[0] = 0.0f;
    if (i == n0 - 1) w[1] = 0.0f;
    if (j == 0) w[2] = 0.0f;
    if (j == n1 - 1) w[3] = 0.0f;
    if (k == 0) w[4] = 0.0f;
    if (k == n2 - 1) w[5] = 0.0f;
}

// Clamped neighbour indices of cell c: q[2a] lower, q[2a+1] upper on axis a
// (the cell itself where the neighbour is outside).
__device__ __forceinline__ void nbr_index3(int n0, int n1, int n2, int i, int j, int k, int c, int q[6]) {
    const int plane = n0 * n1;
    q[0] = i > 0 ? c - 1 : c;
    q[1] = i < n0 - 1 ? c + 1 : c;
    q[2] = j > 0 ? c - n0 : c;
    q[3] = j < n1 - 1 ? c + n0 : c;
    q[4] = k > 0 ? c - plane : c;
    q[5] = k < n2 - 1 ? c + plane : c;
}

// The neighbour-mask bits of every cell (LevelDevT::nb) from L.mask.
template <typename WT>
__global__ void k_nbits3(LevelDevT<WT> L, uint8_t* __restrict__ nb) {
    const int n0 = L.n0, n1 = L.n1, n2 = L.n2;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n0 * n1 * n2) return;
    const int row = c / n0;
    const int i = c - row * n0;
    const int k = row / n1;
    const int j = row - k * n1;
    int q[6];
    nbr_index3(n0, n1, n2, i, j, k, c, q);
    unsigned v = 0;
#pragma unroll
    for (int a = 0; a < 6; ++a) v |= L.mask[q[a]] ? (1u << a) : 0u;
    v |= L.mask[c] ? 64u : 0u;
    nb[c] = (uint8_t)v;
}

// Row sum and diagonal from the centre value, the six Dirichlet-masked
// neighbour values and the six face weights (outside faces zero).
template <typename T>
__device__ __forceinline__ void row_terms(T xc, const T xn[6], const float w[6], const T ih2[3], T& sum, T& dsum) {
    sum = T(0);
    dsum = T(0);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const T wl = (T)w[2 * a], wu = (T)w[2 * a + 1];
        const T dLo = (xc - xn[2 * a]) * wl;
        const T dUp = (xn[2 * a + 1] - xc) * wu;
        sum = sum + (dUp - dLo) * ih2[a];
        dsum = dsum - (wl + wu) * ih2[a];
    }
}

}  // namespace flux
