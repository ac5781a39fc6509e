// Particle <-> grid transfers and semi-Lagrangian advection for gfx950.
//
// Reference: packages/fluxflow/src/grid/grid_flip_solver3.js (binning, P2G,
// G2P, particle advection), grid_math3.js (trilinear and monotone cubic
// sampling), advection_solver3.js (back-trace, face advection).
// Built with -ffp-contract=off: every product and sum rounds like the
// oracle's, so binning, weights and the gather sums agree bit for bit.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>

#include "flux_grid3.h"
#include "flux_internal.h"
#include "flux_solver.h"

namespace flux {

static constexpr int XTPB = 256;

// ------------------------------------------------------------------ binning
// cellIndexOf (grid_flip_solver3.js:227-235) with an outside counter
__global__ void k_cell_index(G3 g, const float* __restrict__ pos, int count, int* __restrict__ cell,
                             int* __restrict__ ids, unsigned* outside) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= count) return;
    const float o[3] = {g.o0, g.o1, g.o2}, h[3] = {g.h0, g.h1, g.h2};
    const int n[3] = {g.n0, g.n1, g.n2};
    int c[3];
    bool out = false;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float x = (pos[3 * p + a] - o[a]) / h[a];
        int f = (int)floorf(x);
        out = out || !(f >= 0 && f < n[a]);
        c[a] = clampi(f, 0, n[a] - 1);
    }
    cell[p] = c[0] + g.n0 * (c[1] + g.n1 * c[2]);
    ids[p] = p;
    if (out) atomicAdd(outside, 1u);
}

// cell_start[c] = first sorted rank whose key >= c (lower bound), c in [0, ncells]
__global__ void k_cell_start(const int* __restrict__ keys, int count, int ncells, int* __restrict__ start) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > ncells) return;
    int lo = 0, hi = count;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (keys[mid] < c) lo = mid + 1;
        else hi = mid;
    }
    start[c] = lo;
}

// ---------------------------------------------------------------------- P2G
// Deterministic gather: one thread per face of component `comp`; the cells
// that can reach the face are visited in (k, j, i) order, their particles in
// sorted order, each particle's corners in the reference order.
__global__ void k_p2g_gather(G3 g, int comp, const float* __restrict__ pos, const float* __restrict__ vel,
                             const int* __restrict__ perm, const int* __restrict__ start, float eps,
                             float* __restrict__ out, int* __restrict__ valid) {
    const int s[3] = {g.n0 + (comp == 0), g.n1 + (comp == 1), g.n2 + (comp == 2)};
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= s[0] * s[1] * s[2]) return;
    int row = t / s[0];
    int fi = t - row * s[0];
    int fk = row / s[1];
    int fj = row - fk * s[1];
    const int F[3] = {fi, fj, fk};
    const int n[3] = {g.n0, g.n1, g.n2};
    int lo[3], hi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        lo[a] = max(F[a] - 1, 0);
        hi[a] = min((a == comp) ? F[a] : F[a] + 1, n[a] - 1);
    }
    float o[3];
    mac_origin(g, comp, o);
    const float h[3] = {g.h0, g.h1, g.h2};
    float nu = 0.0f, de = 0.0f;
    for (int ck = lo[2]; ck <= hi[2]; ++ck)
        for (int cj = lo[1]; cj <= hi[1]; ++cj)
            for (int ci = lo[0]; ci <= hi[0]; ++ci) {
                int c = ci + g.n0 * (cj + g.n1 * ck);
                int e = start[c + 1];
                for (int q = start[c]; q < e; ++q) {
                    int p = perm[q];
                    const float pp[3] = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
                    Tri tr;
                    tri_weights(pp, o, h, s, tr);
                    float value = vel[3 * p + comp];
#pragma unroll
                    for (int cc = 0; cc < 8; ++cc) {
                        int i = (cc & 1) ? tr.i1 : tr.i0, j = (cc & 2) ? tr.j1 : tr.j0, k = (cc & 4) ? tr.k1 : tr.k0;
                        if (i == fi && j == fj && k == fk) {
                            nu = nu + value * tr.w[cc];
                            de = de + tr.w[cc];
                        }
                    }
                }
            }
    if (de > eps) {
        out[t] = nu / de;
        valid[t] = 1;
    } else {
        out[t] = 0.0f;
        valid[t] = 0;
    }
}

__device__ __forceinline__ int f32_to_i32_sat(float v) {
    if (!(v == v)) return 0;
    if (v >= 2147483520.0f) return 2147483647;
    if (v <= -2147483648.0f) return (int)0x80000000u;
    return (int)v;
}

// buildScatter (grid_flip_solver3.js:508-535): every particle adds its 8
// corner weights w and value * w, in int32 fixed point (rint(x * scale),
// saturated), to the face accumulators; with `conc` (massWeightedTransfer,
// grid_flip_solver3.js:493-531) every corner weight is scaled by
// rho(c) / rho_ref, rho(c) = ma + (mb - ma) c.
// Aggregated per block: a block takes kP2GPPT * XTPB
// consecutive particles; when every corner they reach lies in a box of at
// most `box_limit` faces (particles stored in spatial order: seeded, or
// sorted, as they are in a FLIP scene), the fixed-point contributions are
// first summed in LDS and the box is then added to the grid with one atomic
// per touched face and array, instead of 16 atomics per particle.  Integer
// addition is associative and commutative (wrapping like the atomics), so the
// grid ends with the same bits as the per-corner atomics in any order.  A
// block whose box is larger adds its corners to the grid directly.
static constexpr int kP2GPPT = 4;
static constexpr int kP2GBox = 4096;

__global__ void __launch_bounds__(XTPB) k_p2g_scatter_box(G3 g, int comp, const float* __restrict__ pos,
                                                          const float* __restrict__ vel, int count, float scale,
                                                          int* __restrict__ numer, int* __restrict__ denom,
                                                          const float* __restrict__ conc, float ma, float mb,
                                                          float mref, int box_limit) {
    __shared__ int s_num[kP2GBox];
    __shared__ int s_den[kP2GBox];
    __shared__ int s_lo[3], s_hi[3];
    const int s[3] = {g.n0 + (comp == 0), g.n1 + (comp == 1), g.n2 + (comp == 2)};
    float o[3];
    mac_origin(g, comp, o);
    const float h[3] = {g.h0, g.h1, g.h2};
    const int64_t first = (int64_t)blockIdx.x * (XTPB * kP2GPPT);
    if (threadIdx.x < 3) {
        s_lo[threadIdx.x] = 0x7fffffff;
        s_hi[threadIdx.x] = -1;
    }
    // pass 1: the box of corners the block's particles reach
    int lo[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff}, hi[3] = {-1, -1, -1};
#pragma unroll
    for (int q = 0; q < kP2GPPT; ++q) {
        const int64_t p = first + q * XTPB + threadIdx.x;
        if (p < count) {
            const float pp[3] = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
            Tri tr;
            tri_weights(pp, o, h, s, tr);
            lo[0] = min(lo[0], tr.i0);
            lo[1] = min(lo[1], tr.j0);
            lo[2] = min(lo[2], tr.k0);
            hi[0] = max(hi[0], tr.i1);
            hi[1] = max(hi[1], tr.j1);
            hi[2] = max(hi[2], tr.k1);
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
    const bool lds = vol <= (int64_t)min(box_limit, kP2GBox);  // uniform over the block
    if (lds) {
        for (int e = threadIdx.x; e < (int)vol; e += XTPB) {
            s_num[e] = 0;
            s_den[e] = 0;
        }
        __syncthreads();
    }
    // pass 2: the corners, into LDS or straight to the grid
#pragma unroll
    for (int q = 0; q < kP2GPPT; ++q) {
        const int64_t p = first + q * XTPB + threadIdx.x;
        if (p >= count) continue;
        float mass = 1.0f;
        if (conc) {
            const float c = fminf(fmaxf(conc[p], 0.0f), 1.0f);
            mass = (ma + (mb - ma) * c) / mref;
        }
        const float pp[3] = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
        Tri tr;
        tri_weights(pp, o, h, s, tr);
        const float value = vel[3 * p + comp];
#pragma unroll
        for (int cc = 0; cc < 8; ++cc) {
            const int i = (cc & 1) ? tr.i1 : tr.i0, j = (cc & 2) ? tr.j1 : tr.j0, k = (cc & 4) ? tr.k1 : tr.k0;
            const float w = conc ? tr.w[cc] * mass : tr.w[cc];
            const int nv = f32_to_i32_sat(rintf((value * w) * scale));
            const int dv = f32_to_i32_sat(rintf(w * scale));
            if (lds) {
                const int e = (i - b0) + bx * ((j - b1) + by * (k - b2));
                atomicAdd(&s_num[e], nv);
                atomicAdd(&s_den[e], dv);
            } else {
                const int f = i + s[0] * (j + s[1] * k);
                atomicAdd(&numer[f], nv);
                atomicAdd(&denom[f], dv);
            }
        }
    }
    if (!lds) return;
    __syncthreads();
    // adding 0 changes nothing: untouched and cancelled faces are skipped
    for (int e = threadIdx.x; e < (int)vol; e += XTPB) {
        const int nv = s_num[e], dv = s_den[e];
        if ((nv | dv) == 0) continue;
        const int i = b0 + e % bx, j = b1 + (e / bx) % by, k = b2 + e / (bx * by);
        const int f = i + s[0] * (j + s[1] * k);
        if (nv) atomicAdd(&numer[f], nv);
        if (dv) atomicAdd(&denom[f], dv);
    }
}

// buildFinalize (grid_flip_solver3.js:543-564)
__global__ void k_p2g_finalize_fixed(int nf, const int* __restrict__ numer, const int* __restrict__ denom, float scale,
                                     float eps, float* __restrict__ out, int* __restrict__ valid) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    float nu = (float)numer[f] / scale;
    float de = (float)denom[f] / scale;
    if (de > eps) {
        out[f] = nu / de;
        valid[f] = 1;
    } else {
        out[f] = 0.0f;
        valid[f] = 0;
    }
}

// ---------------------------------------------------------------- G2P etc.
// g2pUpdate (grid_flip_solver3.js:903-922)
__global__ void k_g2p(G3 g, Mac now, Mac old, const float* __restrict__ pos, float* __restrict__ vel, int count,
                      float flip_ratio, float damping) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= count) return;
    const float pp[3] = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
    float nv[3], ov[3];
    sample_mac(g, now, pp, nv);
    sample_mac(g, old, pp, ov);
    float dmp = damping < 0.0f ? 0.0f : (damping > 1.0f ? 1.0f : damping);
    float one_r = 1.0f - flip_ratio, one_d = 1.0f - dmp;
    float res[3];
    bool bad = false;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float delta = nv[a] - ov[a];
        float fl = vel[3 * p + a] + delta;
        float blended = fl * flip_ratio + nv[a] * one_r;
        res[a] = blended * one_d;
        bad = bad || !isfinite(res[a]);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float r = fminf(fmaxf(res[a], -500.0f), 500.0f);
        vel[3 * p + a] = bad ? 0.0f : r;
    }
}

// advectParticles (grid_flip_solver3.js:933-940)
__global__ void k_advect_particles(G3 g, Mac m, Sdf3 col, float* __restrict__ pos, int count, float dt,
                                   int max_sub) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= count) return;
    float hmin = fminf(g.h0, fminf(g.h1, g.h2));
    const float pp[3] = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
    float tr[3];
    back_trace_c(g, m, col, pp, dt, hmin, max_sub, -1.0f, tr);
    bool bad = !isfinite(tr[0]) || !isfinite(tr[1]) || !isfinite(tr[2]);
    const float o[3] = {g.o0, g.o1, g.o2}, h[3] = {g.h0, g.h1, g.h2};
    const int n[3] = {g.n0, g.n1, g.n2};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float lo = o[a] + 1e-4f;
        float hi = (o[a] + (float)n[a] * h[a]) - 1e-4f;
        float s = bad ? pp[a] : tr[a];
        pos[3 * p + a] = s < lo ? lo : (s > hi ? hi : s);
    }
}

__global__ void k_advect_faces(G3 g, Mac m, int comp, const float* __restrict__ src, float* __restrict__ dst, float dt,
                               int max_sub) {
    const int s[3] = {g.n0 + (comp == 0), g.n1 + (comp == 1), g.n2 + (comp == 2)};
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= s[0] * s[1] * s[2]) return;
    int row = t / s[0];
    int i = t - row * s[0];
    int k = row / s[1];
    int j = row - k * s[1];
    float o[3];
    mac_origin(g, comp, o);
    const float h[3] = {g.h0, g.h1, g.h2};
    const float pos[3] = {o[0] + (float)i * h[0], o[1] + (float)j * h[1], o[2] + (float)k * h[2]};
    float hmin = fminf(g.h0, fminf(g.h1, g.h2));
    float tr[3];
    back_trace(g, m, pos, dt, hmin, max_sub, 1.0f, tr);
    dst[t] = sample_cubic(src, s, o, h, tr);
}

// liquid fraction from sorted particles: a cell holding particles is liquid
__global__ void k_fraction_from_bins(const int* __restrict__ start, int ncells, float* __restrict__ frac) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < ncells) frac[c] = (start[c + 1] - start[c]) > 0 ? 1.0f : 0.0f;
}

// two-phase face beta (grid_two_phase_flip_solver2.js:693-762, in 3-D)
__global__ void k_two_phase_beta(G3 g, int comp, const float* __restrict__ frac, float r0, float r1, float ref,
                                 float* __restrict__ beta) {
    const int s[3] = {g.n0 + (comp == 0), g.n1 + (comp == 1), g.n2 + (comp == 2)};
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= s[0] * s[1] * s[2]) return;
    int row = t / s[0];
    int i = t - row * s[0];
    int k = row / s[1];
    int j = row - k * s[1];
    int A[3] = {i, j, k}, B[3] = {i, j, k};
    const int n[3] = {g.n0, g.n1, g.n2};
    A[comp] = clampi(A[comp] - 1, 0, n[comp] - 1);
    B[comp] = clampi(B[comp], 0, n[comp] - 1);
    float ra = r0 + (r1 - r0) * frac[A[0] + g.n0 * (A[1] + g.n1 * A[2])];
    float rb = r0 + (r1 - r0) * frac[B[0] + g.n0 * (B[1] + g.n1 * B[2])];
    float rf = (ra + rb) * 0.5f;
    beta[t] = ref / (rf > 1e-4f ? rf : 1e-4f);
}

// faceCenteredDivergenceAtCenter3 (grid_math3.js:397-410)
__global__ void k_divergence(G3 g, Mac m, float* __restrict__ out) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.n0 * g.n1 * g.n2) return;
    int row = c / g.n0;
    int i = c - row * g.n0;
    int k = row / g.n1;
    int j = row - k * g.n1;
    int ul = i + (g.n0 + 1) * (j + g.n1 * k);
    int vb = i + g.n0 * (j + (g.n1 + 1) * k);
    float d = (m.u[ul + 1] - m.u[ul]) / g.h0;
    d = d + (m.v[vb + g.n0] - m.v[vb]) / g.h1;
    d = d + (m.w[c + g.n0 * g.n1] - m.w[c]) / g.h2;
    out[c] = d;
}

// ------------------------------------------------------------- sort scratch
struct SortScratch {
    void* temp = nullptr;
    size_t temp_bytes = 0;
    int* keys_alt = nullptr;
    int* ids = nullptr;
    int64_t cap = 0;
    unsigned* outside = nullptr;
};
static SortScratch g_sort;

static int ensure_sort(int64_t count, size_t temp_bytes) {
    if (count > g_sort.cap) {
        dfree(g_sort.keys_alt);
        dfree(g_sort.ids);
        g_sort.keys_alt = (int*)dmalloc(count * 4);
        g_sort.ids = (int*)dmalloc(count * 4);
        g_sort.cap = (g_sort.keys_alt && g_sort.ids) ? count : 0;
        if (!g_sort.cap) return set_error(FLUX_ERR_NOMEM, "bin_particles: scratch allocation failed");
    }
    if (temp_bytes > g_sort.temp_bytes) {
        dfree(g_sort.temp);
        g_sort.temp = dmalloc(temp_bytes);
        g_sort.temp_bytes = g_sort.temp ? temp_bytes : 0;
        if (!g_sort.temp) return set_error(FLUX_ERR_NOMEM, "bin_particles: sort scratch allocation failed");
    }
    if (!g_sort.outside) {
        g_sort.outside = (unsigned*)dmalloc(sizeof(unsigned));
        if (!g_sort.outside) return set_error(FLUX_ERR_NOMEM, "bin_particles: allocation failed");
    }
    return FLUX_OK;
}

static int check_grid(const flux_grid3* g) {
    FLUX_REQUIRE(g, "grid is NULL");
    for (int a = 0; a < 3; ++a) {
        FLUX_REQUIRE(g->n[a] > 0, "grid: n[%d] must be positive", a);
        FLUX_REQUIRE(g->h[a] > 0, "grid: h[%d] must be positive", a);
    }
    FLUX_REQUIRE((int64_t)(g->n[0] + 1) * (g->n[1] + 1) * (g->n[2] + 1) < (1ll << 31), "grid: too many cells");
    return FLUX_OK;
}

static Mac mac_of(const flux_field* const* v) {
    return Mac{(const float*)v[0]->d, (const float*)v[1]->d, (const float*)v[2]->d};
}

static int check_mac(const flux_grid3* g, const flux_field* const* v, const char* what) {
    FLUX_REQUIRE(v, "%s: velocity fields are NULL", what);
    for (int a = 0; a < 3; ++a) FLUX_TRY(check_field(v[a], FLUX_F32, faces_of(g->n, a), what));
    return FLUX_OK;
}

}  // namespace flux

using namespace flux;

extern "C" {

int flux_bin_particles(const flux_grid3* g, const flux_field* positions, int64_t count, flux_field* cell,
                       flux_field* perm, flux_field* cell_start, int64_t* outside) {
    FLUX_REQUIRE_INIT();
    FLUX_TRY(check_grid(g));
    FLUX_REQUIRE(count >= 0 && count < (1ll << 31), "bin_particles: count out of range");
    FLUX_REQUIRE(positions && (positions->dtype == FLUX_VEC3 || positions->dtype == FLUX_F32),
                 "bin_particles: positions must be vec3 or f32x3");
    int64_t pos_elems = positions->dtype == FLUX_VEC3 ? positions->count : positions->count / 3;
    FLUX_REQUIRE(pos_elems >= count, "bin_particles: positions hold %lld particles, %lld requested",
                 (long long)pos_elems, (long long)count);
    int64_t ncells = cells_of(g->n);
    FLUX_REQUIRE(cell && cell->dtype == FLUX_I32 && cell->count >= count, "bin_particles: cell must be i32[count]");
    FLUX_REQUIRE(perm && perm->dtype == FLUX_I32 && perm->count >= count, "bin_particles: perm must be i32[count]");
    FLUX_TRY(check_field(cell_start, FLUX_I32, ncells + 1, "bin_particles: cell_start"));
    G3 gg = make_g3(g);
    size_t temp = 0;
    FLUX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, temp, (const int*)nullptr, (int*)nullptr, (const int*)nullptr,
                                                (int*)nullptr, (int)std::max<int64_t>(count, 1), 0, 32, g_stream));
    FLUX_TRY(ensure_sort(std::max<int64_t>(count, 1), temp));
    FLUX_HIP(hipMemsetAsync(g_sort.outside, 0, sizeof(unsigned), g_stream));
    if (count > 0) {
        k_cell_index<<<blocks_for(count, XTPB), XTPB, 0, g_stream>>>(gg, (const float*)positions->d, (int)count,
                                                                      g_sort.keys_alt, g_sort.ids, g_sort.outside);
        FLUX_LAUNCH_CHECK();
        // stable LSD radix sort on the cell key: ties keep particle order
        int bits = 1;
        while ((1ll << bits) < ncells) ++bits;
        FLUX_HIP(hipcub::DeviceRadixSort::SortPairs(g_sort.temp, temp, g_sort.keys_alt, (int*)cell->d, g_sort.ids,
                                                    (int*)perm->d, (int)count, 0, bits, g_stream));
    }
    k_cell_start<<<blocks_for(ncells + 1, XTPB), XTPB, 0, g_stream>>>((const int*)cell->d, (int)count, (int)ncells,
                                                                       (int*)cell_start->d);
    FLUX_LAUNCH_CHECK();
    unsigned out = 0;
    FLUX_HIP(hipMemcpyAsync(&out, g_sort.outside, sizeof out, hipMemcpyDeviceToHost, g_stream));
    FLUX_HIP(hipStreamSynchronize(g_stream));
    if (outside) *outside = out;
    return FLUX_OK;
}

int flux_p2g_gather(const flux_grid3* g, const flux_field* positions, const flux_field* velocities,
                    const flux_field* perm, const flux_field* cell_start, int component, float eps, flux_field* out,
                    flux_field* valid) {
    FLUX_REQUIRE_INIT();
    FLUX_TRY(check_grid(g));
    FLUX_REQUIRE(component >= 0 && component < 3, "p2g: component must be 0, 1 or 2");
    FLUX_REQUIRE(positions && velocities && perm && cell_start, "p2g: NULL field");
    FLUX_TRY(check_field(cell_start, FLUX_I32, cells_of(g->n) + 1, "p2g: cell_start"));
    int64_t nf = faces_of(g->n, component);
    FLUX_TRY(check_field(out, FLUX_F32, nf, "p2g: out"));
    FLUX_TRY(check_field(valid, FLUX_I32, nf, "p2g: valid"));
    k_p2g_gather<<<blocks_for(nf, XTPB), XTPB, 0, g_stream>>>(make_g3(g), component, (const float*)positions->d,
                                                               (const float*)velocities->d, (const int*)perm->d,
                                                               (const int*)cell_start->d, eps, (float*)out->d,
                                                               (int*)valid->d);
    FLUX_LAUNCH_CHECK();
    FLUX_HIP(hipStreamSynchronize(g_stream));
    return FLUX_OK;
}

}  // extern "C"

namespace flux {

// FLUX_P2G_BOX=0 sends every block's corners straight to the grid (the
// per-corner atomics; the same bits, for comparison)
static int p2g_box_limit() {
    const char* e = getenv("FLUX_P2G_BOX");
    return (e && e[0] == '0') ? 0 : kP2GBox;
}

int p2g_fixed_async(const flux_grid3* g, const flux_field* positions, const flux_field* velocities, int64_t count,
                    int component, float scale, float eps, flux_field* numer, flux_field* denom, flux_field* out,
                    flux_field* valid, const flux_field* conc, float d0, float d1) {
    FLUX_TRY(check_grid(g));
    FLUX_REQUIRE(component >= 0 && component < 3, "p2g: component must be 0, 1 or 2");
    FLUX_REQUIRE(count >= 0 && count < (1ll << 31), "p2g: count out of range");
    FLUX_REQUIRE(positions && velocities, "p2g: NULL field");
    FLUX_REQUIRE((int64_t)(positions->bytes / 12) >= count && (int64_t)(velocities->bytes / 12) >= count,
                 "p2g: particle fields hold fewer than %lld particles", (long long)count);
    if (conc) FLUX_TRY(check_field(conc, FLUX_F32, count, "p2g: concentration"));
    int64_t nf = faces_of(g->n, component);
    FLUX_TRY(check_field(numer, FLUX_I32, nf, "p2g: numer"));
    FLUX_TRY(check_field(denom, FLUX_I32, nf, "p2g: denom"));
    FLUX_TRY(check_field(out, FLUX_F32, nf, "p2g: out"));
    FLUX_TRY(check_field(valid, FLUX_I32, nf, "p2g: valid"));
    FLUX_TRY(fill_device(numer->d, FLUX_I32, nf, 0));
    FLUX_TRY(fill_device(denom->d, FLUX_I32, nf, 0));
    const float ma = std::max(d0, 1e-4f), mb = std::max(d1, 1e-4f);
    if (count > 0)
        k_p2g_scatter_box<<<blocks_for(count, XTPB * kP2GPPT), XTPB, 0, g_stream>>>(
            make_g3(g), component, (const float*)positions->d, (const float*)velocities->d, (int)count, scale,
            (int*)numer->d, (int*)denom->d, conc ? (const float*)conc->d : nullptr, ma, mb, std::max(ma, mb),
            p2g_box_limit());
    k_p2g_finalize_fixed<<<blocks_for(nf, XTPB), XTPB, 0, g_stream>>>((int)nf, (const int*)numer->d,
                                                                        (const int*)denom->d, scale, eps,
                                                                        (float*)out->d, (int*)valid->d);
    FLUX_LAUNCH_CHECK();
    return FLUX_OK;
}

}  // namespace flux

extern "C" {

int flux_p2g_fixed(const flux_grid3* g, const flux_field* positions, const flux_field* velocities, int64_t count,
                   int component, float scale, float eps, flux_field* numer, flux_field* denom, flux_field* out,
                   flux_field* valid) {
    FLUX_REQUIRE_INIT();
    FLUX_TRY(p2g_fixed_async(g, positions, velocities, count, component, scale, eps, numer, denom, out, valid,
                             nullptr, 0.0f, 0.0f));
    FLUX_HIP(hipStreamSynchronize(g_stream));
    return FLUX_OK;
}

int flux_g2p(const flux_grid3* g, const flux_field* positions, flux_field* velocities, int64_t count,
             const flux_field* const* vel_new, const flux_field* const* vel_old, float flip_ratio, float damping) {
    FLUX_REQUIRE_INIT();
    FLUX_TRY(check_grid(g));
    FLUX_TRY(check_mac(g, vel_new, "g2p: new velocity"));
    FLUX_TRY(check_mac(g, vel_old, "g2p: old velocity"));
    FLUX_REQUIRE(count >= 0 && count < (1ll << 31), "g2p: count out of range");
    FLUX_REQUIRE(positions && velocities && (int64_t)(positions->bytes / 12) >= count &&
                     (int64_t)(velocities->bytes / 12) >= count,
                 "g2p: particle fields hold fewer than %lld particles", (long long)count);
    if (count > 0)
        k_g2p<<<blocks_for(count, XTPB), XTPB, 0, g_stream>>>(make_g3(g), mac_of(vel_new), mac_of(vel_old),
                                                               (const float*)positions->d, (float*)velocities->d,
                                                               (int)count, flip_ratio, damping);
    FLUX_LAUNCH_CHECK();
    FLUX_HIP(hipStreamSynchronize(g_stream));
    return FLUX_OK;
}

int flux_advect_particles_collider(const flux_grid3* g, flux_field* positions, int64_t count,
                                   const flux_field* const* vel, const flux_field* collider_sdf, float dt,
                                   int max_substeps) {
    FLUX_REQUIRE_INIT();
    FLUX_TRY(check_grid(g));
    FLUX_TRY(check_mac(g, vel, "advectParticles: velocity"));
    FLUX_REQUIRE(count >= 0 && count < (1ll << 31), "advectParticles: count out of range");
    FLUX_REQUIRE(positions && (int64_t)(positions->bytes / 12) >= count, "advectParticles: positions too small");
    FLUX_REQUIRE(max_substeps >= 1 && max_substeps <= 1024, "advectParticles: maxSubsteps out of range");
    if (collider_sdf)
        FLUX_TRY(check_field(collider_sdf, FLUX_F32, (int64_t)g->n[0] * g->n[1] * g->n[2], "advectParticles: collider"));
    const G3 g3 = make_g3(g);
    const Sdf3 col = sdf3_on(g3, collider_sdf ? (const float*)collider_sdf->d : nullptr);
    if (count > 0)
        k_advect_particles<<<blocks_for(count, XTPB), XTPB, 0, g_stream>>>(g3, mac_of(vel), col, (float*)positions->d,
                                                                            (int)count, dt, max_substeps);
    FLUX_LAUNCH_CHECK();
    FLUX_HIP(hipStreamSynchronize(g_stream));
    return FLUX_OK;
}

int flux_advect_particles(const flux_grid3* g, flux_field* positions, int64_t count, const flux_field* const* vel,
                          float dt, int max_substeps) {
    return flux_advect_particles_collider(g, positions, count, vel, nullptr, dt, max_substeps);
}

int flux_advect_faces(const flux_grid3* g, const flux_field* const* vel_in, flux_field* const* vel_out, float dt,
                      int max_substeps) {
    FLUX_REQUIRE_INIT();
    FLUX_TRY(check_grid(g));
    FLUX_TRY(check_mac(g, vel_in, "advectFaceCentered3: input"));
    FLUX_TRY(check_mac(g, vel_out, "advectFaceCentered3: output"));
    FLUX_REQUIRE(max_substeps >= 1 && max_substeps <= 1024, "advectFaceCentered3: maxSubsteps out of range");
    for (int a = 0; a < 3; ++a)
        FLUX_REQUIRE(vel_out[a]->d != vel_in[0]->d && vel_out[a]->d != vel_in[1]->d && vel_out[a]->d != vel_in[2]->d,
                     "advectFaceCentered3: output must not alias the input");
    Mac m = mac_of(vel_in);
    for (int a = 0; a < 3; ++a) {
        int64_t nf = faces_of(g->n, a);
        k_advect_faces<<<blocks_for(nf, XTPB), XTPB, 0, g_stream>>>(make_g3(g), m, a, (const float*)vel_in[a]->d,
                                                                     (float*)vel_out[a]->d, dt, max_substeps);
    }
    FLUX_LAUNCH_CHECK();
    FLUX_HIP(hipStreamSynchronize(g_stream));
    return FLUX_OK;
}

int flux_two_phase_beta(const flux_grid3* g, const flux_field* cell_start, float rho_gas, float rho_liquid,
                        flux_field* fraction, flux_field* const* beta) {
    FLUX_REQUIRE_INIT();
    FLUX_TRY(check_grid(g));
    int64_t ncells = cells_of(g->n);
    FLUX_TRY(check_field(fraction, FLUX_F32, ncells, "two_phase_beta: fraction"));
    FLUX_TRY(check_mac(g, beta, "two_phase_beta: beta"));
    if (cell_start) {
        FLUX_TRY(check_field(cell_start, FLUX_I32, ncells + 1, "two_phase_beta: cell_start"));
        k_fraction_from_bins<<<blocks_for(ncells, XTPB), XTPB, 0, g_stream>>>((const int*)cell_start->d, (int)ncells,
                                                                               (float*)fraction->d);
    }
    float r0 = std::max(rho_gas, 1e-4f), r1 = std::max(rho_liquid, 1e-4f), ref = std::max(r0, r1);
    for (int a = 0; a < 3; ++a) {
        int64_t nf = faces_of(g->n, a);
        k_two_phase_beta<<<blocks_for(nf, XTPB), XTPB, 0, g_stream>>>(make_g3(g), a, (const float*)fraction->d, r0, r1,
                                                                       ref, (float*)beta[a]->d);
    }
    FLUX_LAUNCH_CHECK();
    FLUX_HIP(hipStreamSynchronize(g_stream));
    return FLUX_OK;
}

int flux_divergence(const flux_grid3* g, const flux_field* const* vel, flux_field* out) {
    FLUX_REQUIRE_INIT();
    FLUX_TRY(check_grid(g));
    FLUX_TRY(check_mac(g, vel, "divergence: velocity"));
    FLUX_TRY(check_field(out, FLUX_F32, cells_of(g->n), "divergence: out"));
    k_divergence<<<blocks_for(cells_of(g->n), XTPB), XTPB, 0, g_stream>>>(make_g3(g), mac_of(vel), (float*)out->d);
    FLUX_LAUNCH_CHECK();
    FLUX_HIP(hipStreamSynchronize(g_stream));
    return FLUX_OK;
}

}  // extern "C"
