// Device-side stencil helpers shared by the solver kernels.
//
// The operator is the reference's laplacianAt / laplacianDiagonalAt
// (packages/fluxflow/src/linalg/multigrid.js:273-339, 399-437):
//   (A x)_c = sum_axis w_up*dUp - w_lo*dLo, scaled by 1/h_axis^2, with
//   dLo = x_c - x_lo (x_c if lo is Dirichlet, 0 if lo is outside),
//   dUp = x_up - x_c (-x_c if up is Dirichlet, 0 if up is outside);
//   a Dirichlet row is -x_c with diagonal -1; the diagonal of a free row is
//   -sum over present faces of w/h^2 (Dirichlet neighbours included).
#pragma once
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace flux {

// storage conversions: every kernel computes in FP32 (or FP64) and stores in
// the level's storage type
template <typename T>
__device__ __forceinline__ T to_store(float v);
template <>
__device__ __forceinline__ float to_store<float>(float v) { return v; }
template <>
__device__ __forceinline__ __half to_store<__half>(float v) { return __float2half(v); }
template <>
__device__ __forceinline__ __hip_bfloat16 to_store<__hip_bfloat16>(float v) { return __float2bfloat16(v); }

template <typename T>
__device__ __forceinline__ float to_f32(T v) { return static_cast<float>(v); }

// One multigrid level (or the operator's only level) as the kernels see it;
// WT is the storage type of the face weights.
template <typename WT>
struct LevelDevT {
    int n0, n1, n2;      // cells per axis (n2 == 1 in 2-D)
    int ndim;
    float ih2[3];        // 1/h^2 per axis
    double ih2d[3];
    const uint8_t* mask; // 1 = Dirichlet row; may be null
    const WT* w0;        // face weights per axis; all null = unit weights
    const WT* w1;
    const WT* w2;
    // column-march kernels: per cell, bit a (0..5) = mask of the clamped
    // neighbour a (-x, +x, -y, +y, -z, +z; the cell itself where the
    // neighbour is outside), bit 6 = the cell's own mask; null when mask is
    const uint8_t* nb = nullptr;
};
using LevelDev = LevelDevT<float>;

template <typename WT>
__device__ __forceinline__ void cell_coords(const LevelDevT<WT>& L, int c, int& i, int& j, int& k) {
    int row = c / L.n0;
    i = c - row * L.n0;
    k = row / L.n1;
    j = row - k * L.n1;
}

// lower-face index of cell (i,j,k) on each axis; upper = lower + fstride
template <typename WT>
__device__ __forceinline__ int face_lo(const LevelDevT<WT>& L, int axis, int i, int j, int k, int c) {
    if (axis == 0) return i + (L.n0 + 1) * (j + L.n1 * k);
    if (axis == 1) return i + L.n0 * (j + (L.n1 + 1) * k);
    return c;
}
template <typename WT>
__device__ __forceinline__ int face_up_stride(const LevelDevT<WT>& L, int axis) {
    return axis == 0 ? 1 : (axis == 1 ? L.n0 : L.n0 * L.n1);
}

// Row of A applied to a field read through `ld(index)`, arithmetic in T.
// Writes the row's diagonal to *diag when diag is non-null.
template <typename T, typename WT, typename LD>
__device__ __forceinline__ T stencil_row(const LevelDevT<WT>& L, int i, int j, int k, int c, LD ld, T* diag) {
    const uint8_t* m = L.mask;
    T center = ld(c);
    if (m && m[c]) {
        if (diag) *diag = T(-1);
        return -center;
    }
    const int n[3] = {L.n0, L.n1, L.n2};
    const int I[3] = {i, j, k};
    const int st[3] = {1, L.n0, L.n0 * L.n1};
    const WT* w[3] = {L.w0, L.w1, L.w2};
    T sum = T(0), dsum = T(0);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (a < L.ndim) {
            T dLo = T(0), dUp = T(0), hLo = T(0), hUp = T(0);
            if (I[a] > 0) {
                int l = c - st[a];
                dLo = (m && m[l]) ? center : (center - ld(l));
                hLo = T(1);
            }
            if (I[a] < n[a] - 1) {
                int u = c + st[a];
                dUp = (m && m[u]) ? -center : (ld(u) - center);
                hUp = T(1);
            }
            if (w[a]) {
                int fl = face_lo(L, a, i, j, k, c);
                T wl = (T)to_f32(w[a][fl]);
                T wu = (T)to_f32(w[a][fl + face_up_stride(L, a)]);
                dLo = dLo * wl;
                dUp = dUp * wu;
                hLo = hLo * wl;
                hUp = hUp * wu;
            }
            T ih2 = sizeof(T) == 8 ? (T)L.ih2d[a] : (T)L.ih2[a];
            sum = sum + (dUp - dLo) * ih2;
            dsum = dsum - (hLo + hUp) * ih2;
        }
    }
    if (diag) *diag = dsum;
    return sum;
}

// XCD-aware block order: the hardware hands consecutive block ids to the 8
// XCDs round robin; this bijection gives each XCD a contiguous run of logical
// blocks (neighbouring tiles share their halos in that XCD's L2).  XCD x
// receives hardware ids x, x+8, ...: ceil((nb-x)/8) of them.
__device__ __forceinline__ unsigned xcd_remap(unsigned b, unsigned nb) {
    const unsigned q = nb >> 3, r = nb & 7u;
    const unsigned xcd = b & 7u, slot = b >> 3;
    return xcd < r ? xcd * (q + 1) + slot : r * (q + 1) + (xcd - r) * q + slot;
}

// Deterministic FP64 block sum (blockDim.x a multiple of 64); the result is
// valid in thread 0.  All threads of the block must call it.
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double sm[16];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sm[wid] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int q = 0; q < nw; ++q) s += sm[q];
    }
    __syncthreads();
    return s;
}

// Two-stage deterministic FP64 sum.  Stage 1: every block of the producing
// kernel stores its block sum at partials[blockIdx.x] (all threads call).
__device__ __forceinline__ void store_block_partial(double v, double* partials) {
    const double s = block_sum(v);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// Stage 2, in a one-block kernel launched after the producer: the sum of the
// nb partials in a fixed order, valid in thread 0 (all threads call).
__device__ __forceinline__ double sum_partials(const double* partials, unsigned nb) {
    double acc = 0;
    for (unsigned q = threadIdx.x; q < nb; q += blockDim.x) acc += partials[q];
    return block_sum(acc);
}

}  // namespace flux
