// createGridFlipSolver3 for gfx950: the FLIP time step around the pressure
// projection (packages/fluxflow/src/grid/grid_flip_solver3.js:135-1155),
// free-surface form with reduced pressure, static collider, deterministic
// particle resampling (640-793).  Stages are composed from the library's
// own transfer, boundary, extrapolation and projection objects; this unit
// adds the resampling, the Dirichlet fields of the free surface and gravity.
// Built with -ffp-contract=off.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>

#include "flux_grid3.h"
#include "flux_internal.h"
#include "flux_solver.h"

namespace flux {

static constexpr int STPB = 256;

__device__ __forceinline__ int cell_of(const G3& g, const float* __restrict__ pos, int p) {
    const float x = (pos[3 * p + 0] - g.o0) / g.h0;
    const float y = (pos[3 * p + 1] - g.o1) / g.h1;
    const float z = (pos[3 * p + 2] - g.o2) / g.h2;
    const int i = clampi((int)floorf(x), 0, g.n0 - 1);
    const int j = clampi((int)floorf(y), 0, g.n1 - 1);
    const int k = clampi((int)floorf(z), 0, g.n2 - 1);
    return i + g.n0 * (j + g.n1 * k);
}

// countPerCellKernel (grid_flip_solver3.js:675-680): integer atomics, exact
__global__ void k_count_per_cell(G3 g, const float* __restrict__ pos, int count, int* __restrict__ cnt) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < count) atomicAdd(&cnt[cell_of(g, pos, p)], 1);
}

// markDonorsKernel (684-690)
__global__ void k_mark_donors(G3 g, const float* __restrict__ pos, int count, const int* __restrict__ cnt, int maxpc,
                              int* __restrict__ flag) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < count) flag[p] = cnt[cell_of(g, pos, p)] > maxpc ? 1 : 0;
}

// scatterDonorsKernel (693-698): rank among the donors before it
__global__ void k_scatter_donors(const int* __restrict__ flag, const int* __restrict__ scan, int count,
                                 int* __restrict__ pool) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < count && flag[p] == 1) pool[scan[p] - 1] = p;
}

// neighborFluidCount + markNeedKernel (703-727)
__global__ void k_mark_need(G3 g, const int* __restrict__ cnt, int minpc, int* __restrict__ need) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int n0 = g.n0, n1 = g.n1, n2 = g.n2;
    if (c >= n0 * n1 * n2) return;
    const int i = c % n0, j = (c / n0) % n1, k = c / (n0 * n1);
    auto at = [&](int ii, int jj, int kk) { return cnt[ii + n0 * (jj + n1 * kk)]; };
    const int nb = (at(max(0, i - 1), j, k) >= minpc) + (at(min(n0 - 1, i + 1), j, k) >= minpc) +
                   (at(i, max(0, j - 1), k) >= minpc) + (at(i, min(n1 - 1, j + 1), k) >= minpc) +
                   (at(i, j, max(0, k - 1)) >= minpc) + (at(i, j, min(n2 - 1, k + 1)) >= minpc);
    const int count = cnt[c];
    const bool wants = count < minpc && nb >= 3;  // MIN_FLUID_NEIGHBORS
    need[c] = wants ? minpc - count : 0;
}

// claimDonorsKernel (728-760): a cell's slice of the donor pool starts after
// everything earlier cells asked for; a claimed donor moves to the cell
// centre and takes the grid velocity there.
__global__ void k_claim_donors(G3 g, Mac m, const int* __restrict__ need, const int* __restrict__ need_scan,
                               const int* __restrict__ donor_scan, int count, const int* __restrict__ pool, int minpc,
                               float* __restrict__ pos, float* __restrict__ vel) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.n0 * g.n1 * g.n2) return;
    const int needed = need[c];
    if (needed <= 0) return;
    const int i = c % g.n0, j = (c / g.n0) % g.n1, k = c / (g.n0 * g.n1);
    const int base = need_scan[c] - needed;
    const int total = count > 0 ? donor_scan[count - 1] : 0;
    const float co[3] = {g.o0 + g.h0 * 0.5f, g.o1 + g.h1 * 0.5f, g.o2 + g.h2 * 0.5f};
    const float np[3] = {co[0] + (float)i * g.h0, co[1] + (float)j * g.h1, co[2] + (float)k * g.h2};
    int claimed = 0;
    for (int q = 0; q < minpc; ++q) {
        if (claimed >= needed) break;
        const int slot = base + claimed;
        if (slot >= total) break;
        const int d = pool[slot];
        pos[3 * d + 0] = np[0];
        pos[3 * d + 1] = np[1];
        pos[3 * d + 2] = np[2];
        float v[3];
        sample_mac(g, m, np, v);
        vel[3 * d + 0] = v[0];
        vel[3 * d + 1] = v[1];
        vel[3 * d + 2] = v[2];
        ++claimed;
    }
}

// dirichlet(pos) of the FLIP solver (grid_flip_solver3.js:252-259) evaluated
// by updateDirichletFields at cell centres (grid_pressure_solver3.js:171-206):
// air cells (fluid mask < 0.5) are Dirichlet with the reduced-pressure target
// -dt (g . pos); the mask sampled exactly at a cell centre is the cell's own.
__global__ void k_free_surface(G3 g, const float* __restrict__ mask, float gx, float gy, float gz, float dt,
                               int reduced, uint8_t* __restrict__ active, float* __restrict__ target) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.n0 * g.n1 * g.n2) return;
    const int i = c % g.n0, j = (c / g.n0) % g.n1, k = c / (g.n0 * g.n1);
    const float px = (g.o0 + g.h0 * 0.5f) + (float)i * g.h0;
    const float py = (g.o1 + g.h1 * 0.5f) + (float)j * g.h1;
    const float pz = (g.o2 + g.h2 * 0.5f) + (float)k * g.h2;
    active[c] = mask[c] < 0.5f ? 1 : 0;
    target[c] = reduced ? -(((gx * px + gy * py) + gz * pz) * dt) : 0.0f;
}

// applyGravityU/V/W (grid_flip_solver3.js:798-808), the non-reduced form
__global__ void k_add_const(float* __restrict__ d, int n, float v) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f < n) d[f] = d[f] + v;
}

// ---------------------------------------------- carried concentration, density
static constexpr float MIN_DENSITY = 1e-4f;

__device__ __forceinline__ float clamp01(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }

__device__ __forceinline__ int f2i_sat(float v) {
    if (!(v == v)) return 0;
    if (v >= 2147483520.0f) return 2147483647;
    if (v <= -2147483648.0f) return (int)0x80000000u;
    return (int)v;
}

// pushOutOfCollider (grid_flip_solver3.js:945-964): a particle where the
// collider's SDF is below `thresh` moves along the SDF's normalised
// (trilinear) gradient by (thresh - sdf) + shift
__global__ void k_push_out(Sdf3 col, float thresh, float shift, float* __restrict__ pos, int count) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= count) return;
    const float pp[3] = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
    const float v = sdf_sample(col, pp);
    if (!(v < thresh)) return;
    const int n[3] = {col.n0, col.n1, col.n2};
    const float o[3] = {col.o0, col.o1, col.o2}, h[3] = {col.h0, col.h1, col.h2};
    float g[3];
    sdf_gradient(col.d, n, o, h, pp, g);
    const float len = sqrtf((g[0] * g[0] + g[1] * g[1]) + g[2] * g[2]);
    if (!(len > 0.0f)) return;
    const float dist = (thresh - v) + shift;
#pragma unroll
    for (int a = 0; a < 3; ++a) pos[3 * p + a] = pp[a] + (g[a] / len) * dist;
}

// accumulateConcentration (grid_flip_solver3.js:992-998): count and the
// fixed-point sum of clamp(c, 0, 1) per cell (integer atomics, order-free).
// A block takes kCellPPT * STPB consecutive particles; when their cells lie in
// a box of at most kCellBox cells (particles in spatial order) both sums are
// formed in LDS and added to the grid once per touched cell, otherwise every
// particle adds to the grid directly.  Integer sums in any order give the
// same bits.
static constexpr int kCellPPT = 4;
static constexpr int kCellBox = 4096;

__device__ __forceinline__ void cell_coords(const G3& g, const float* __restrict__ pos, int64_t p, int c[3]) {
    const float x = (pos[3 * p + 0] - g.o0) / g.h0;
    const float y = (pos[3 * p + 1] - g.o1) / g.h1;
    const float z = (pos[3 * p + 2] - g.o2) / g.h2;
    c[0] = clampi((int)floorf(x), 0, g.n0 - 1);
    c[1] = clampi((int)floorf(y), 0, g.n1 - 1);
    c[2] = clampi((int)floorf(z), 0, g.n2 - 1);
}

__global__ void __launch_bounds__(STPB) k_conc_accumulate(G3 g, const float* __restrict__ pos,
                                                         const float* __restrict__ conc, int count, float scale,
                                                         int* __restrict__ accum, int* __restrict__ cnt) {
    __shared__ int s_acc[kCellBox];
    __shared__ int s_cnt[kCellBox];
    __shared__ int s_lo[3], s_hi[3];
    const int64_t first = (int64_t)blockIdx.x * (STPB * kCellPPT);
    if (threadIdx.x < 3) {
        s_lo[threadIdx.x] = 0x7fffffff;
        s_hi[threadIdx.x] = -1;
    }
    int lo[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff}, hi[3] = {-1, -1, -1};
#pragma unroll
    for (int q = 0; q < kCellPPT; ++q) {
        const int64_t p = first + q * STPB + threadIdx.x;
        if (p < count) {
            int c[3];
            cell_coords(g, pos, p, c);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                lo[a] = min(lo[a], c[a]);
                hi[a] = max(hi[a], c[a]);
            }
        }
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            lo[a] = min(lo[a], __shfl_xor(lo[a], d));
            hi[a] = max(hi[a], __shfl_xor(hi[a], d));
        }
    __syncthreads();
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            atomicMin(&s_lo[a], lo[a]);
            atomicMax(&s_hi[a], hi[a]);
        }
    __syncthreads();
    const int b0 = s_lo[0], b1 = s_lo[1], b2 = s_lo[2];
    const int bx = s_hi[0] - b0 + 1, by = s_hi[1] - b1 + 1, bz = s_hi[2] - b2 + 1;
    if (bx <= 0) return;  // no particle in this block (uniform)
    const int64_t vol = (int64_t)bx * by * bz;
    const bool lds = vol <= kCellBox;  // uniform over the block
    if (lds) {
        for (int e = threadIdx.x; e < (int)vol; e += STPB) {
            s_acc[e] = 0;
            s_cnt[e] = 0;
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < kCellPPT; ++q) {
        const int64_t p = first + q * STPB + threadIdx.x;
        if (p >= count) continue;
        int c[3];
        cell_coords(g, pos, p, c);
        const int v = f2i_sat(rintf(clamp01(conc[p]) * scale));
        if (lds) {
            const int e = (c[0] - b0) + bx * ((c[1] - b1) + by * (c[2] - b2));
            atomicAdd(&s_cnt[e], 1);
            atomicAdd(&s_acc[e], v);
        } else {
            const int cell = c[0] + g.n0 * (c[1] + g.n1 * c[2]);
            atomicAdd(&cnt[cell], 1);
            atomicAdd(&accum[cell], v);
        }
    }
    if (!lds) return;
    __syncthreads();
    for (int e = threadIdx.x; e < (int)vol; e += STPB) {
        const int nc = s_cnt[e], v = s_acc[e];
        if (nc == 0) continue;  // no particle: both sums stay 0
        const int cell = (b0 + e % bx) + g.n0 * ((b1 + (e / bx) % by) + g.n1 * (b2 + e / (bx * by)));
        atomicAdd(&cnt[cell], nc);
        if (v) atomicAdd(&accum[cell], v);
    }
}

// resolveCellConcentration (1000-1012)
__global__ void k_conc_resolve(int cells, const int* __restrict__ accum, const int* __restrict__ cnt, float scale,
                               float d0, float d1, float* __restrict__ cell_conc, float* __restrict__ cell_density) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cells) return;
    const float total = (float)cnt[c];
    const float sum = (float)accum[c] / scale;
    const float mean = sum / fmaxf(total, 1.0f);
    const float cc = total > 0.5f ? clamp01(mean) : 0.0f;
    cell_conc[c] = cc;
    if (cell_density) cell_density[c] = d0 + (d1 - d0) * cc;
}

// computeBetaU/V/W with faceDensity (313-357), all three axes in one launch
__global__ void k_density_beta(G3 g, const float* __restrict__ rho, const float* __restrict__ mask, float ref,
                               float* __restrict__ bu, float* __restrict__ bv, float* __restrict__ bw) {
    const int nu = (g.n0 + 1) * g.n1 * g.n2, nv = g.n0 * (g.n1 + 1) * g.n2, nw = g.n0 * g.n1 * (g.n2 + 1);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nu + nv + nw) return;
    const int ax = t < nu ? 0 : (t < nu + nv ? 1 : 2);
    const int f = ax == 0 ? t : (ax == 1 ? t - nu : t - nu - nv);
    const int s0 = g.n0 + (ax == 0), s1 = g.n1 + (ax == 1);
    const int i = f % s0, j = (f / s0) % s1, k = f / (s0 * s1);
    int A[3] = {i, j, k}, B[3] = {i, j, k};
    const int n[3] = {g.n0, g.n1, g.n2};
    A[ax] = clampi(A[ax] - 1, 0, n[ax] - 1);
    B[ax] = clampi(B[ax], 0, n[ax] - 1);
    const int ca = A[0] + g.n0 * (A[1] + g.n1 * A[2]), cb = B[0] + g.n0 * (B[1] + g.n1 * B[2]);
    const float ra = rho[ca], rb = rho[cb];
    const bool fa = mask[ca] > 0.5f, fb = mask[cb] > 0.5f;
    const float rf = (fa && fb) ? (ra + rb) * 0.5f : ((fa || fb) ? (fa ? ra : rb) : ref);
    (ax == 0 ? bu : (ax == 1 ? bv : bw))[f] = ref / fmaxf(rf, MIN_DENSITY);
}

// applyReducedGravityU/V/W (848-880): u -= beta dt / rho_ref (g . x_face) d(rho)
__global__ void k_reduced_gravity(G3 g, const float* __restrict__ rho, const float* __restrict__ bu,
                                  const float* __restrict__ bv, const float* __restrict__ bw, float gx, float gy,
                                  float gz, float dt, float ref, float* __restrict__ u, float* __restrict__ v,
                                  float* __restrict__ w) {
    const int nu = (g.n0 + 1) * g.n1 * g.n2, nv = g.n0 * (g.n1 + 1) * g.n2, nw = g.n0 * g.n1 * (g.n2 + 1);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nu + nv + nw) return;
    const int ax = t < nu ? 0 : (t < nu + nv ? 1 : 2);
    const int f = ax == 0 ? t : (ax == 1 ? t - nu : t - nu - nv);
    const int s0 = g.n0 + (ax == 0), s1 = g.n1 + (ax == 1);
    const int I[3] = {f % s0, (f / s0) % s1, f / (s0 * s1)};
    const float h[3] = {g.h0, g.h1, g.h2}, o[3] = {g.o0, g.o1, g.o2};
    const int n[3] = {g.n0, g.n1, g.n2};
    float pos[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) pos[q] = o[q] + (q == ax ? (float)I[q] * h[q] : ((float)I[q] + 0.5f) * h[q]);
    int A[3] = {I[0], I[1], I[2]}, B[3] = {I[0], I[1], I[2]};
    A[ax] = clampi(A[ax] - 1, 0, n[ax] - 1);
    B[ax] = clampi(B[ax], 0, n[ax] - 1);
    const float drho = (rho[B[0] + g.n0 * (B[1] + g.n1 * B[2])] - rho[A[0] + g.n0 * (A[1] + g.n1 * A[2])]) / h[ax];
    const float gdot = (gx * pos[0] + gy * pos[1]) + gz * pos[2];
    const float beta = (ax == 0 ? bu : (ax == 1 ? bv : bw))[f];
    const float term = (((beta * dt) / ref) * gdot) * drho;
    float* vel = ax == 0 ? u : (ax == 1 ? v : w);
    vel[f] = vel[f] - term;
}

// mixAndFade (1018-1027)
__global__ void k_mix_fade(G3 g, const float* __restrict__ pos, float* __restrict__ conc, int count,
                           const float* __restrict__ cell_conc, float mixing, float fade) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= count) return;
    const int n[3] = {g.n0, g.n1, g.n2};
    const float h[3] = {g.h0, g.h1, g.h2};
    const float co[3] = {g.o0 + g.h0 * 0.5f, g.o1 + g.h1 * 0.5f, g.o2 + g.h2 * 0.5f};
    const float pp[3] = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
    const float sampled = sample_tri(cell_conc, n, co, h, pp);
    const float c = clamp01(conc[p]);
    const float local = isfinite(sampled) ? sampled : c;
    const float blended = c + (local - c) * mixing;
    conc[p] = clamp01(blended * (1.0f - fade));
}

// pressurePin: flag any Dirichlet cell, then pin the given cell to 0 when there is none
__global__ void k_any_active(int cells, const uint8_t* __restrict__ active, int* __restrict__ flag) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < cells && active[c] && *flag == 0) atomicOr(flag, 1);
}

__global__ void k_pin(int cell, const int* __restrict__ flag, uint8_t* __restrict__ active,
                      float* __restrict__ target) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && *flag == 0) {
        active[cell] = 1;
        target[cell] = 0.0f;
    }
}

__global__ void k_set_int(int* __restrict__ p, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// ------------------------------------------------ phase-aware resampling
// (grid_two_phase_flip_solver2.js:1296-1478, in 3-D with the 3-of-6 rule)
__global__ void k_need_phase(G3 g, const int* __restrict__ cnt, const int* __restrict__ accum, int minpc, float scale,
                             int* __restrict__ need_l, int* __restrict__ need_g) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int n0 = g.n0, n1 = g.n1, n2 = g.n2;
    if (c >= n0 * n1 * n2) return;
    const int i = c % n0, j = (c / n0) % n1, k = c / (n0 * n1);
    const int nb[6] = {max(0, i - 1) + n0 * (j + n1 * k), min(n0 - 1, i + 1) + n0 * (j + n1 * k),
                       i + n0 * (max(0, j - 1) + n1 * k), i + n0 * (min(n1 - 1, j + 1) + n1 * k),
                       i + n0 * (j + n1 * max(0, k - 1)), i + n0 * (j + n1 * min(n2 - 1, k + 1))};
    int well = 0;
    unsigned ntot = 0, nsum = 0;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        well += cnt[nb[q]] >= minpc;
        ntot += (unsigned)cnt[nb[q]];
        nsum += (unsigned)accum[nb[q]];
    }
    const float total = (float)cnt[c];
    const float own = ((float)accum[c] / scale) / fmaxf(total, 1.0f);
    const float ntf = (float)(int)ntot;
    const float nmean = ((float)(int)nsum / scale) / fmaxf(ntf, 1.0f);
    const bool liquid = total > 0.5f ? own > 0.5f : nmean > 0.5f;
    const bool eligible = cnt[c] < minpc && well >= 3;
    const int needed = minpc - cnt[c];
    need_l[c] = (eligible && liquid) ? needed : 0;
    need_g[c] = (eligible && !liquid) ? needed : 0;
}

__global__ void k_mark_donors_phase(G3 g, const float* __restrict__ pos, const float* __restrict__ conc, int count,
                                    const int* __restrict__ cnt, int maxpc, int* __restrict__ flag_l,
                                    int* __restrict__ flag_g) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= count) return;
    const bool over = cnt[cell_of(g, pos, p)] > maxpc;
    const bool liquid = conc[p] > 0.5f;
    flag_l[p] = (over && liquid) ? 1 : 0;
    flag_g[p] = (over && !liquid) ? 1 : 0;
}

__global__ void k_scatter_donors_phase(const int* __restrict__ flag_l, const int* __restrict__ scan_l,
                                       const int* __restrict__ flag_g, const int* __restrict__ scan_g, int count,
                                       int half, int* __restrict__ pool) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= count) return;
    if (flag_l[p] == 1 && scan_l[p] - 1 < half) pool[scan_l[p] - 1] = p;
    if (flag_g[p] == 1 && scan_g[p] - 1 < half) pool[(count - 1) - (scan_g[p] - 1)] = p;
}

__global__ void k_claim_donors_phase(G3 g, Mac m, const int* __restrict__ need_l, const int* __restrict__ need_g,
                                     const int* __restrict__ scan_nl, const int* __restrict__ scan_ng,
                                     const int* __restrict__ scan_l, const int* __restrict__ scan_g, int count,
                                     int half, const int* __restrict__ pool, int minpc, float* __restrict__ pos,
                                     float* __restrict__ vel) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.n0 * g.n1 * g.n2) return;
    const bool liquid = need_l[c] > 0;
    const int needed = liquid ? need_l[c] : need_g[c];
    if (needed <= 0) return;
    const int i = c % g.n0, j = (c / g.n0) % g.n1, k = c / (g.n0 * g.n1);
    const int base = (liquid ? scan_nl[c] : scan_ng[c]) - needed;
    const int donors = count > 0 ? (liquid ? scan_l[count - 1] : scan_g[count - 1]) : 0;
    const int avail = donors < half ? donors : half;
    const float co[3] = {g.o0 + g.h0 * 0.5f, g.o1 + g.h1 * 0.5f, g.o2 + g.h2 * 0.5f};
    const float np[3] = {co[0] + (float)i * g.h0, co[1] + (float)j * g.h1, co[2] + (float)k * g.h2};
    int claimed = 0;
    for (int q = 0; q < minpc; ++q) {
        if (claimed >= needed) break;
        const int slot = base + claimed;
        if (slot >= avail) break;
        const int d = pool[liquid ? slot : (count - 1) - slot];
        pos[3 * d + 0] = np[0];
        pos[3 * d + 1] = np[1];
        pos[3 * d + 2] = np[2];
        float v[3];
        sample_mac(g, m, np, v);
        vel[3 * d + 0] = v[0];
        vel[3 * d + 1] = v[1];
        vel[3 * d + 2] = v[2];
        ++claimed;
    }
}

}  // namespace flux

using namespace flux;

struct flux_flip3 {
    flux_flip3_desc d;
    flux_grid3 g;
    int64_t np = 0;
    int64_t cells = 0;
    flux_field* positions = nullptr;
    flux_field* velocities = nullptr;
    flux_field* fluid_mask = nullptr;
    flux_field* vel[3] = {nullptr, nullptr, nullptr};
    flux_field* old[3] = {nullptr, nullptr, nullptr};
    flux_field* numer[3] = {nullptr, nullptr, nullptr};
    flux_field* denom[3] = {nullptr, nullptr, nullptr};
    flux_field* weight_valid[3] = {nullptr, nullptr, nullptr};
    flux_field* fluid_adj[3] = {nullptr, nullptr, nullptr};
    flux_field* active = nullptr;
    flux_field* target = nullptr;
    flux_extrapolator* ex_weight[3] = {nullptr, nullptr, nullptr};
    flux_extrapolator* ex_pressure[3] = {nullptr, nullptr, nullptr};
    flux_boundary3* boundary = nullptr;
    flux_pressure3* pressure = nullptr;
    // resampling
    int* cell_count = nullptr;
    int* donor_flag = nullptr;
    int* donor_scan = nullptr;
    int* donor_pool = nullptr;
    int* need = nullptr;
    int* need_scan = nullptr;
    void* scan_temp = nullptr;
    size_t scan_temp_bytes = 0;
    // carried concentration and density coupling
    bool coupling = false;
    float d0 = 1.0f, d1 = 1.0f, ref = 1.0f;  // floored densities and the heavier one
    flux_field* conc = nullptr;               // per particle
    flux_field* cell_conc = nullptr;
    flux_field* cell_density = nullptr;
    flux_field* beta[3] = {nullptr, nullptr, nullptr};
    flux_field* captured[3] = {nullptr, nullptr, nullptr};  // the last projection's input velocity, when kept
    int* conc_accum = nullptr;
    int* conc_count = nullptr;
    int* pin_flag = nullptr;
    int pin_index = -1;
    // phase-aware resampling (second flag/scan/need arrays)
    int* donor_flag_g = nullptr;
    int* donor_scan_g = nullptr;
    int* need_g = nullptr;
    int* need_scan_g = nullptr;
};

namespace {

int make_face(const int n[3], int axis, int dtype, flux_field** f) {
    int64_t s[3] = {n[0] + (axis == 0), n[1] + (axis == 1), n[2] + (axis == 2)};
    return flux_field_create(dtype, 3, s, f);
}

// the resample pass (grid_flip_solver3.js:762-793)
int resample(flux_flip3* s) {
    const G3 g = make_g3(&s->g);
    const int np = (int)s->np, nc = (int)s->cells;
    const float* pos = (const float*)s->positions->d;
    FLUX_HIP(hipMemsetAsync(s->cell_count, 0, sizeof(int) * nc, g_stream));
    if (np > 0) {
        k_count_per_cell<<<blocks_for(np, STPB), STPB, 0, g_stream>>>(g, pos, np, s->cell_count);
        k_mark_donors<<<blocks_for(np, STPB), STPB, 0, g_stream>>>(g, pos, np, s->cell_count,
                                                                   s->d.max_particles_per_cell, s->donor_flag);
        FLUX_HIP(hipcub::DeviceScan::InclusiveSum(s->scan_temp, s->scan_temp_bytes, s->donor_flag, s->donor_scan, np,
                                                  g_stream));
        k_scatter_donors<<<blocks_for(np, STPB), STPB, 0, g_stream>>>(s->donor_flag, s->donor_scan, np, s->donor_pool);
    }
    k_mark_need<<<blocks_for(nc, STPB), STPB, 0, g_stream>>>(g, s->cell_count, s->d.min_particles_per_cell, s->need);
    FLUX_HIP(hipcub::DeviceScan::InclusiveSum(s->scan_temp, s->scan_temp_bytes, s->need, s->need_scan, nc, g_stream));
    Mac m{(const float*)s->vel[0]->d, (const float*)s->vel[1]->d, (const float*)s->vel[2]->d};
    k_claim_donors<<<blocks_for(nc, STPB), STPB, 0, g_stream>>>(g, m, s->need, s->need_scan, s->donor_scan, np,
                                                                s->donor_pool, s->d.min_particles_per_cell,
                                                                (float*)s->positions->d, (float*)s->velocities->d);
    FLUX_LAUNCH_CHECK();
    return FLUX_OK;
}

int mark_fluid_and_adjacency(flux_flip3* s) {
    FLUX_TRY(flux_mark_fluid_cells(&s->g, s->positions, s->np, s->fluid_mask));
    for (int a = 0; a < 3; ++a) FLUX_TRY(flux_fluid_adjacent(&s->g, s->fluid_mask, a, s->fluid_adj[a]));
    return FLUX_OK;
}

// per-cell particle count and fixed-point concentration sum
int accumulate_concentration(flux_flip3* s) {
    const int np = (int)s->np, nc = (int)s->cells;
    FLUX_TRY(fill_device(s->conc_accum, FLUX_I32, nc, 0));
    FLUX_TRY(fill_device(s->conc_count, FLUX_I32, nc, 0));
    if (np > 0)
        k_conc_accumulate<<<blocks_for(np, STPB * kCellPPT), STPB, 0, g_stream>>>(
            make_g3(&s->g), (const float*)s->positions->d, (const float*)s->conc->d, np,
            s->d.concentration_atomic_scale, s->conc_accum, s->conc_count);
    FLUX_LAUNCH_CHECK();
    return FLUX_OK;
}

// the phase-aware resample pass (grid_two_phase_flip_solver2.js:1480-1503)
int resample_phase(flux_flip3* s) {
    const G3 g = make_g3(&s->g);
    const int np = (int)s->np, nc = (int)s->cells;
    const int half = std::max(1, np / 2);
    FLUX_TRY(accumulate_concentration(s));
    k_need_phase<<<blocks_for(nc, STPB), STPB, 0, g_stream>>>(g, s->conc_count, s->conc_accum,
                                                              s->d.min_particles_per_cell,
                                                              s->d.concentration_atomic_scale, s->need, s->need_g);
    if (np > 0) {
        k_mark_donors_phase<<<blocks_for(np, STPB), STPB, 0, g_stream>>>(
            g, (const float*)s->positions->d, (const float*)s->conc->d, np, s->conc_count,
            s->d.max_particles_per_cell, s->donor_flag, s->donor_flag_g);
        FLUX_HIP(hipcub::DeviceScan::InclusiveSum(s->scan_temp, s->scan_temp_bytes, s->donor_flag, s->donor_scan, np,
                                                  g_stream));
        FLUX_HIP(hipcub::DeviceScan::InclusiveSum(s->scan_temp, s->scan_temp_bytes, s->donor_flag_g, s->donor_scan_g,
                                                  np, g_stream));
        k_scatter_donors_phase<<<blocks_for(np, STPB), STPB, 0, g_stream>>>(
            s->donor_flag, s->donor_scan, s->donor_flag_g, s->donor_scan_g, np, half, s->donor_pool);
    }
    FLUX_HIP(hipcub::DeviceScan::InclusiveSum(s->scan_temp, s->scan_temp_bytes, s->need, s->need_scan, nc, g_stream));
    FLUX_HIP(hipcub::DeviceScan::InclusiveSum(s->scan_temp, s->scan_temp_bytes, s->need_g, s->need_scan_g, nc,
                                              g_stream));
    Mac m{(const float*)s->vel[0]->d, (const float*)s->vel[1]->d, (const float*)s->vel[2]->d};
    k_claim_donors_phase<<<blocks_for(nc, STPB), STPB, 0, g_stream>>>(
        g, m, s->need, s->need_g, s->need_scan, s->need_scan_g, s->donor_scan, s->donor_scan_g, np, half,
        s->donor_pool, s->d.min_particles_per_cell, (float*)s->positions->d, (float*)s->velocities->d);
    FLUX_LAUNCH_CHECK();
    return FLUX_OK;
}

// concentrationGridPass (grid_flip_solver3.js:1029-1045): accumulate, resolve
// the cell concentration (and density), then the face weights
int concentration_grid_pass(flux_flip3* s) {
    const int nc = (int)s->cells;
    FLUX_TRY(accumulate_concentration(s));
    k_conc_resolve<<<blocks_for(nc, STPB), STPB, 0, g_stream>>>(
        nc, s->conc_accum, s->conc_count, s->d.concentration_atomic_scale, s->d0, s->d1, (float*)s->cell_conc->d,
        s->coupling ? (float*)s->cell_density->d : nullptr);
    if (s->coupling) {
        const int64_t nf = faces_of(s->g.n, 0) + faces_of(s->g.n, 1) + faces_of(s->g.n, 2);
        k_density_beta<<<blocks_for(nf, STPB), STPB, 0, g_stream>>>(
            make_g3(&s->g), (const float*)s->cell_density->d, (const float*)s->fluid_mask->d, s->ref,
            (float*)s->beta[0]->d, (float*)s->beta[1]->d, (float*)s->beta[2]->d);
    }
    FLUX_LAUNCH_CHECK();
    return FLUX_OK;
}

}  // namespace

extern "C" {

void flux_flip3_desc_default(flux_flip3_desc* d) {
    memset(d, 0, sizeof(*d));
    d->resolution[0] = d->resolution[1] = d->resolution[2] = 1;
    d->spacing[0] = d->spacing[1] = d->spacing[2] = 1.0;
    d->gravity[0] = 0.0f;
    d->gravity[1] = -9.81f;
    d->gravity[2] = 0.0f;
    d->reduced_pressure = 1;
    d->flip_ratio = 0.97f;
    d->velocity_damping = 0.02f;
    d->p2g_atomic_scale = 65536.0f;
    d->weight_epsilon = 1e-4f;
    d->closed_domain_flags = 63;
    d->resample = 1;
    d->min_particles_per_cell = 4;
    d->max_particles_per_cell = 8;
    d->friction = 0.5f;
    d->extrapolation_depth = 5;
    d->max_substeps = 32;
    flux_pressure3_desc_default(&d->pressure);
    d->pressure.max_plausible_pressure = 0.0;  // derive from the scene
    d->concentration_atomic_scale = 4096.0f;
}

int flux_flip3_create(const flux_flip3_desc* d, flux_flip3** out) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(d && out, "createGridFlipSolver3: NULL argument");
    FLUX_REQUIRE(d->max_particles > 0 && d->max_particles < (1ll << 31),
                 "createGridFlipSolver3: options.maxParticles is required.");
    for (int a = 0; a < 3; ++a)
        FLUX_REQUIRE(d->resolution[a] > 0 && d->spacing[a] > 0, "createGridFlipSolver3: bad grid");
    FLUX_REQUIRE(d->min_particles_per_cell >= 0 && d->max_particles_per_cell >= 0, "createGridFlipSolver3: bad resample counts");
    const bool coupling = d->ambient_density != 0.0f || d->component_density != 0.0f;
    FLUX_REQUIRE(!coupling || d->carry_concentration,
                 "createGridFlipSolver3: ambientDensity/componentDensity require carryConcentration: true -- there is "
                 "no concentration field to build a density from otherwise.");
    FLUX_REQUIRE(!coupling || d->ambient_density > 0.0f, "createGridFlipSolver3: ambientDensity must be > 0, got %g.",
                 (double)d->ambient_density);
    FLUX_REQUIRE(!coupling || d->component_density > 0.0f,
                 "createGridFlipSolver3: componentDensity must be > 0, got %g.", (double)d->component_density);
    FLUX_REQUIRE(!d->phase_resample || d->carry_concentration,
                 "createGridFlipSolver3: phase-aware resampling needs carryConcentration: true");
    FLUX_REQUIRE(!d->carry_concentration || d->concentration_atomic_scale > 0.0f,
                 "createGridFlipSolver3: concentrationAtomicScale must be > 0");
    FLUX_REQUIRE(d->pressure_pin >= 0 && d->pressure_pin <= 2, "createGridFlipSolver3: bad pressurePin mode");
    if (d->pressure_pin == 2)
        for (int a = 0; a < 3; ++a)
            FLUX_REQUIRE(d->pin_cell[a] >= 0 && d->pin_cell[a] < d->resolution[a],
                         "createGridFlipSolver3: pressurePin must be 'topCenter', null, or an [i, j, k] cell index");
    FLUX_REQUIRE(d->pressure.face_weights[0] == nullptr || !coupling,
                 "createGridFlipSolver3: with density coupling the face weights are the solver's own");
    flux_flip3* s = new flux_flip3();
    s->d = *d;
    s->coupling = coupling;
    if (coupling) {
        s->d0 = std::max(d->ambient_density, MIN_DENSITY);
        s->d1 = std::max(d->component_density, MIN_DENSITY);
        s->ref = std::max(s->d0, s->d1);
    }
    if (d->pressure_pin == 1)
        s->pin_index = d->resolution[0] / 2 + d->resolution[0] * ((d->resolution[1] - 1) + d->resolution[1] * (d->resolution[2] / 2));
    else if (d->pressure_pin == 2)
        s->pin_index = d->pin_cell[0] + d->resolution[0] * (d->pin_cell[1] + d->resolution[1] * d->pin_cell[2]);
    for (int a = 0; a < 3; ++a) {
        s->g.n[a] = d->resolution[a];
        s->g.h[a] = (float)d->spacing[a];
        s->g.origin[a] = (float)d->origin[a];
    }
    s->np = d->max_particles;
    s->cells = (int64_t)d->resolution[0] * d->resolution[1] * d->resolution[2];
    int rc = FLUX_OK;
    auto ok = [&](int r) {
        if (rc == FLUX_OK && r != FLUX_OK) rc = r;
        return rc == FLUX_OK;
    };
    int64_t pshape[1] = {s->np};
    int64_t cshape[3] = {d->resolution[0], d->resolution[1], d->resolution[2]};
    ok(flux_field_create(FLUX_VEC3, 1, pshape, &s->positions));
    ok(flux_field_create(FLUX_VEC3, 1, pshape, &s->velocities));
    ok(flux_field_create(FLUX_F32, 3, cshape, &s->fluid_mask));
    ok(flux_field_create(FLUX_U8, 3, cshape, &s->active));
    ok(flux_field_create(FLUX_F32, 3, cshape, &s->target));
    for (int a = 0; a < 3 && rc == FLUX_OK; ++a) {
        ok(make_face(d->resolution, a, FLUX_F32, &s->vel[a]));
        ok(make_face(d->resolution, a, FLUX_F32, &s->old[a]));
        ok(make_face(d->resolution, a, FLUX_I32, &s->numer[a]));
        ok(make_face(d->resolution, a, FLUX_I32, &s->denom[a]));
        ok(make_face(d->resolution, a, FLUX_I32, &s->weight_valid[a]));
        ok(make_face(d->resolution, a, FLUX_I32, &s->fluid_adj[a]));
    }
    if (d->carry_concentration) {
        ok(flux_field_create(FLUX_F32, 1, pshape, &s->conc));
        ok(flux_field_create(FLUX_F32, 3, cshape, &s->cell_conc));
        if (rc == FLUX_OK) ok(fill_device(s->conc->d, FLUX_F32, s->np, 0.0));
        if (rc == FLUX_OK) ok(fill_device(s->cell_conc->d, FLUX_F32, s->cells, 0.0));
        s->conc_accum = (int*)dmalloc(sizeof(int) * s->cells);
        s->conc_count = (int*)dmalloc(sizeof(int) * s->cells);
        if (rc == FLUX_OK && (!s->conc_accum || !s->conc_count))
            rc = set_error(FLUX_ERR_NOMEM, "createGridFlipSolver3: allocation failed");
    }
    if (coupling) {
        ok(flux_field_create(FLUX_F32, 3, cshape, &s->cell_density));
        if (rc == FLUX_OK) ok(fill_device(s->cell_density->d, FLUX_F32, s->cells, 0.0));
        for (int a = 0; a < 3 && rc == FLUX_OK; ++a) {
            ok(make_face(d->resolution, a, FLUX_F32, &s->beta[a]));
            if (rc == FLUX_OK) ok(fill_device(s->beta[a]->d, FLUX_F32, s->beta[a]->count, 1.0));
        }
    }
    if (d->pressure_pin) {
        s->pin_flag = (int*)dmalloc(sizeof(int));
        if (rc == FLUX_OK && !s->pin_flag) rc = set_error(FLUX_ERR_NOMEM, "createGridFlipSolver3: allocation failed");
    }
    for (int a = 0; a < 3 && rc == FLUX_OK; ++a) {
        ok(flux_extrapolator3_create(s->vel[a], s->weight_valid[a], s->vel[a], &s->ex_weight[a]));
        ok(flux_extrapolator3_create(s->vel[a], s->fluid_adj[a], s->vel[a], &s->ex_pressure[a]));
    }
    if (rc == FLUX_OK) {
        ok(flux_boundary3_create(&s->g, s->vel, d->collider_sdf, d->collider_weights[0] ? d->collider_weights : nullptr,
                                 d->friction, &s->boundary));
        if (rc == FLUX_OK) ok(flux_boundary3_set_closed(s->boundary, d->closed_domain_flags));
    }
    if (rc == FLUX_OK) {
        flux_pressure3_desc pd = d->pressure;
        for (int a = 0; a < 3; ++a) {
            pd.resolution[a] = d->resolution[a];
            pd.spacing[a] = d->spacing[a];
            pd.origin[a] = d->origin[a];
        }
        pd.dirichlet_active = s->active;
        pd.dirichlet_target = s->target;
        if (coupling)
            for (int a = 0; a < 3; ++a) pd.face_weights[a] = s->beta[a];
        if (!(pd.max_plausible_pressure > 0)) {
            // derivedMaxPlausiblePressure (grid_flip_solver3.js:425-432)
            const double extent = d->resolution[0] * d->spacing[0] + d->resolution[1] * d->spacing[1] +
                                  d->resolution[2] * d->spacing[2];
            const double gmag = std::hypot(std::hypot((double)d->gravity[0], (double)d->gravity[1]),
                                           (double)d->gravity[2]);
            pd.max_plausible_pressure = (d->reduced_pressure && gmag > 0) ? 8.0 * d->dt * gmag * extent : 1e6;
        }
        ok(flux_pressure3_create(&pd, &s->pressure));
    }
    if (rc == FLUX_OK) {
        const int np = (int)s->np, nc = (int)s->cells;
        s->cell_count = (int*)dmalloc(sizeof(int) * nc);
        s->donor_flag = (int*)dmalloc(sizeof(int) * np);
        s->donor_scan = (int*)dmalloc(sizeof(int) * np);
        s->donor_pool = (int*)dmalloc(sizeof(int) * np);
        s->need = (int*)dmalloc(sizeof(int) * nc);
        s->need_scan = (int*)dmalloc(sizeof(int) * nc);
        size_t t1 = 0, t2 = 0;
        (void)hipcub::DeviceScan::InclusiveSum(nullptr, t1, (const int*)nullptr, (int*)nullptr, np, g_stream);
        (void)hipcub::DeviceScan::InclusiveSum(nullptr, t2, (const int*)nullptr, (int*)nullptr, nc, g_stream);
        s->scan_temp_bytes = std::max<size_t>(std::max(t1, t2), 16);
        s->scan_temp = dmalloc(s->scan_temp_bytes);
        if (!s->cell_count || !s->donor_flag || !s->donor_scan || !s->donor_pool || !s->need || !s->need_scan ||
            !s->scan_temp)
            rc = set_error(FLUX_ERR_NOMEM, "createGridFlipSolver3: allocation failed");
        if (rc == FLUX_OK && d->phase_resample) {
            s->donor_flag_g = (int*)dmalloc(sizeof(int) * np);
            s->donor_scan_g = (int*)dmalloc(sizeof(int) * np);
            s->need_g = (int*)dmalloc(sizeof(int) * nc);
            s->need_scan_g = (int*)dmalloc(sizeof(int) * nc);
            if (!s->donor_flag_g || !s->donor_scan_g || !s->need_g || !s->need_scan_g)
                rc = set_error(FLUX_ERR_NOMEM, "createGridFlipSolver3: allocation failed");
        }
    }
    if (rc != FLUX_OK) {
        std::string msg = flux_last_error();
        flux_flip3_destroy(s);
        return set_error(rc, "%s", msg.c_str());
    }
    *out = s;
    return FLUX_OK;
}

// onAdvanceTimeStep (grid_flip_solver3.js:1120-1132)
int flux_flip3_step(flux_flip3* s, flux_pressure_diag* diag) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(s, "onAdvanceTimeStep: NULL solver");
    const flux_flip3_desc& d = s->d;
    // preP2G: advectParticles (933-940)
    FLUX_TRY(flux_advect_particles_collider(&s->g, s->positions, s->np, s->vel, d.collider_sdf, d.dt,
                                            d.max_substeps));
    if (d.collider_sdf && s->np > 0) {
        k_push_out<<<blocks_for(s->np, STPB), STPB, 0, g_stream>>>(
            sdf3_on(make_g3(&s->g), (const float*)d.collider_sdf->d), d.collider_push_thresh, d.collider_push_shift,
            (float*)s->positions->d, (int)s->np);
        FLUX_LAUNCH_CHECK();
    }
    // afterP2G (1083-1114): scatter + finalize (mass-weighted with density
    // coupling), velOld snapshot, weight extrapolation, constrainVelocity,
    // fluid mask, adjacency, resampling, the concentration pass
    const flux_field* mass = (d.mass_weighted_transfer && s->coupling) ? s->conc : nullptr;
    for (int a = 0; a < 3; ++a)
        FLUX_TRY(p2g_fixed_async(&s->g, s->positions, s->velocities, s->np, a, d.p2g_atomic_scale, d.weight_epsilon,
                                 s->numer[a], s->denom[a], s->vel[a], s->weight_valid[a], mass, s->d0, s->d1));
    for (int a = 0; a < 3; ++a) FLUX_TRY(flux_field_copy(s->old[a], s->vel[a]));
    for (int a = 0; a < 3; ++a) FLUX_TRY(flux_extrapolator3_run(s->ex_weight[a], d.extrapolation_depth));
    FLUX_TRY(flux_boundary3_constrain(s->boundary, d.extrapolation_depth));
    FLUX_TRY(mark_fluid_and_adjacency(s));
    if (d.resample) {
        FLUX_TRY(d.phase_resample ? resample_phase(s) : resample(s));
        FLUX_TRY(mark_fluid_and_adjacency(s));
    }
    if (d.carry_concentration) FLUX_TRY(concentration_grid_pass(s));
    const int64_t nf_all = faces_of(s->g.n, 0) + faces_of(s->g.n, 1) + faces_of(s->g.n, 2);
    if (!d.reduced_pressure) {
        // applyGravityU/V/W; with reduced pressure the body force enters via
        // the air target, plus the density-gradient term when coupled
        for (int a = 0; a < 3; ++a) {
            const int64_t nf = s->vel[a]->count;
            k_add_const<<<blocks_for(nf, STPB), STPB, 0, g_stream>>>((float*)s->vel[a]->d, (int)nf,
                                                                     d.gravity[a] * d.dt);
        }
    } else if (s->coupling) {
        k_reduced_gravity<<<blocks_for(nf_all, STPB), STPB, 0, g_stream>>>(
            make_g3(&s->g), (const float*)s->cell_density->d, (const float*)s->beta[0]->d,
            (const float*)s->beta[1]->d, (const float*)s->beta[2]->d, d.gravity[0], d.gravity[1], d.gravity[2], d.dt,
            s->ref, (float*)s->vel[0]->d, (float*)s->vel[1]->d, (float*)s->vel[2]->d);
    }
    FLUX_LAUNCH_CHECK();
    FLUX_TRY(flux_boundary3_constrain(s->boundary, d.extrapolation_depth));
    // the free-surface Dirichlet fields (and the pin), then the projection in place
    k_free_surface<<<blocks_for(s->cells, STPB), STPB, 0, g_stream>>>(
        make_g3(&s->g), (const float*)s->fluid_mask->d, d.gravity[0], d.gravity[1], d.gravity[2], d.dt,
        d.reduced_pressure, (uint8_t*)s->active->d, (float*)s->target->d);
    if (s->pin_index >= 0) {
        k_set_int<<<1, 64, 0, g_stream>>>(s->pin_flag, 0);
        k_any_active<<<blocks_for(s->cells, STPB), STPB, 0, g_stream>>>((int)s->cells, (const uint8_t*)s->active->d,
                                                                         s->pin_flag);
        k_pin<<<1, 64, 0, g_stream>>>(s->pin_index, s->pin_flag, (uint8_t*)s->active->d, (float*)s->target->d);
    }
    FLUX_LAUNCH_CHECK();
    if (s->captured[0])
        for (int a = 0; a < 3; ++a) FLUX_TRY(flux_field_copy(s->captured[a], s->vel[a]));
    flux_pressure_diag pdg;
    FLUX_TRY(flux_pressure3_project(s->pressure, s->vel[0], s->vel[1], s->vel[2], s->vel[0], s->vel[1], s->vel[2],
                                    &pdg));
    if (diag) *diag = pdg;
    // postProject (1115-1119): constrain, extrapolate into the fluid's
    // neighbourhood, G2P, then mixAndFade of the carried concentration
    FLUX_TRY(flux_boundary3_constrain(s->boundary, d.extrapolation_depth));
    for (int a = 0; a < 3; ++a) FLUX_TRY(flux_extrapolator3_run(s->ex_pressure[a], d.extrapolation_depth));
    FLUX_TRY(flux_g2p(&s->g, s->positions, s->velocities, s->np, s->vel, s->old, d.flip_ratio, d.velocity_damping));
    if (d.carry_concentration && s->np > 0) {
        k_mix_fade<<<blocks_for(s->np, STPB), STPB, 0, g_stream>>>(
            make_g3(&s->g), (const float*)s->positions->d, (float*)s->conc->d, (int)s->np,
            (const float*)s->cell_conc->d, std::min(std::max(d.mixing, 0.0f), 1.0f),
            std::min(std::max(d.fade, 0.0f), 1.0f));
        FLUX_LAUNCH_CHECK();
        FLUX_HIP(hipStreamSynchronize(g_stream));
    }
    return FLUX_OK;
}

flux_field* flux_flip3_field(flux_flip3* s, int which) {
    if (!s) return nullptr;
    switch (which) {
        case FLUX_FLIP3_POSITIONS: return s->positions;
        case FLUX_FLIP3_VELOCITIES: return s->velocities;
        case FLUX_FLIP3_FLUID_MASK: return s->fluid_mask;
        case FLUX_FLIP3_U: return s->vel[0];
        case FLUX_FLIP3_V: return s->vel[1];
        case FLUX_FLIP3_W: return s->vel[2];
        case FLUX_FLIP3_PRESSURE: return s->pressure ? flux_pressure3_field(s->pressure, FLUX_P3_PRESSURE) : nullptr;
        case FLUX_FLIP3_CONCENTRATION: return s->conc;
        case FLUX_FLIP3_CELL_CONCENTRATION: return s->cell_conc;
        case FLUX_FLIP3_CELL_DENSITY: return s->cell_density;
        case FLUX_FLIP3_DIRICHLET_ACTIVE: return s->active;
        case FLUX_FLIP3_DIRICHLET_TARGET: return s->target;
        case FLUX_FLIP3_BETA_U: return s->beta[0];
        case FLUX_FLIP3_BETA_V: return s->beta[1];
        case FLUX_FLIP3_BETA_W: return s->beta[2];
        case FLUX_FLIP3_PROJECTION_INPUT_U: return s->captured[0];
        case FLUX_FLIP3_PROJECTION_INPUT_V: return s->captured[1];
        case FLUX_FLIP3_PROJECTION_INPUT_W: return s->captured[2];
        default: return nullptr;
    }
}

flux_pressure3* flux_flip3_pressure(flux_flip3* s) { return s ? s->pressure : nullptr; }

int flux_flip3_set_dt(flux_flip3* s, float dt) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(s, "dt: NULL solver");
    FLUX_REQUIRE(dt > 0.0f && std::isfinite(dt), "dt must be positive and finite, got %g", (double)dt);
    s->d.dt = dt;
    return FLUX_OK;
}

int flux_flip3_capture_projection_input(flux_flip3* s, int on) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(s, "captureProjectionInput: NULL solver");
    for (int a = 0; a < 3; ++a) {
        if (on && !s->captured[a]) {
            int64_t fs[3] = {s->g.n[0] + (a == 0), s->g.n[1] + (a == 1), s->g.n[2] + (a == 2)};
            FLUX_TRY(flux_field_create(FLUX_F32, 3, fs, &s->captured[a]));
        } else if (!on && s->captured[a]) {
            (void)hipStreamSynchronize(g_stream);
            flux_field_destroy(s->captured[a]);
            s->captured[a] = nullptr;
        }
    }
    return FLUX_OK;
}

int flux_flip3_destroy(flux_flip3* s) {
    if (!s) return FLUX_OK;
    if (g_init) (void)hipStreamSynchronize(g_stream);
    for (int a = 0; a < 3; ++a)
        if (s->captured[a]) flux_field_destroy(s->captured[a]);
    if (s->pressure) flux_pressure3_destroy(s->pressure);
    if (s->boundary) flux_boundary3_destroy(s->boundary);
    for (int a = 0; a < 3; ++a) {
        flux_extrapolator3_destroy(s->ex_weight[a]);
        flux_extrapolator3_destroy(s->ex_pressure[a]);
        flux_field* fs[] = {s->vel[a], s->old[a], s->numer[a], s->denom[a], s->weight_valid[a], s->fluid_adj[a]};
        for (flux_field* f : fs)
            if (f) flux_field_destroy(f);
    }
    flux_field* fs[] = {s->positions, s->velocities, s->fluid_mask, s->active, s->target, s->conc, s->cell_conc,
                        s->cell_density, s->beta[0], s->beta[1], s->beta[2]};
    for (flux_field* f : fs)
        if (f) flux_field_destroy(f);
    dfree(s->conc_accum);
    dfree(s->conc_count);
    dfree(s->pin_flag);
    dfree(s->donor_flag_g);
    dfree(s->donor_scan_g);
    dfree(s->need_g);
    dfree(s->need_scan_g);
    dfree(s->cell_count);
    dfree(s->donor_flag);
    dfree(s->donor_scan);
    dfree(s->donor_pool);
    dfree(s->need);
    dfree(s->need_scan);
    dfree(s->scan_temp);
    delete s;
    return FLUX_OK;
}

}  // extern "C"
