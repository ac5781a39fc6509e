// FLIP grid stages around the projection for gfx950: extrapolation,
// fluid marking, fluid-adjacent faces, static-collider geometry and the
// velocity constraint.
//
// Reference: packages/fluxflow/src/grid/array_utils3.js (extrapolation),
// grid_flip_solver3.js (fluid mask and adjacency), sdf_collider3.js and
// level_set_utils.js (collider SDF and face weights),
// grid_blocked_boundary_condition_solver3.js (constrainVelocity).
// Built with -ffp-contract=off so float results round like the oracle's.
#include <cmath>
#include <vector>

#include "flux_grid3.h"
#include "flux_internal.h"
#include "flux_solver.h"

namespace flux {

static constexpr int FTPB = 256;

// ------------------------------------------------------------ extrapolation
// createExtrapolateStepKernel3 (array_utils3.js:88-162): one Jacobi step.
__global__ void k_extrap_step(int n0, int n1, int n2, const float* __restrict__ vs, float* __restrict__ vd,
                              const int* __restrict__ ms, int* __restrict__ md) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n0 * n1 * n2) return;
    if (ms[c] != 0) {
        vd[c] = vs[c];
        md[c] = 1;
        return;
    }
    const int i = c % n0, j = (c / n0) % n1, k = c / (n0 * n1);
    const int sy = n0, sz = n0 * n1;
    float total = 0.0f;
    int count = 0;
    if (i + 1 < n0 && ms[c + 1]) { total += vs[c + 1]; ++count; }
    if (i > 0 && ms[c - 1]) { total += vs[c - 1]; ++count; }
    if (j + 1 < n1 && ms[c + sy]) { total += vs[c + sy]; ++count; }
    if (j > 0 && ms[c - sy]) { total += vs[c - sy]; ++count; }
    if (k + 1 < n2 && ms[c + sz]) { total += vs[c + sz]; ++count; }
    if (k > 0 && ms[c - sz]) { total += vs[c - sz]; ++count; }
    if (count > 0) {
        vd[c] = total / (float)count;
        md[c] = 1;
    } else {
        vd[c] = vs[c];
        md[c] = 0;
    }
}

// valid flags of any stored type into the int ping-pong buffer (nonzero -> 1)
template <typename T>
__global__ void k_valid_to_int(const T* __restrict__ in, int* __restrict__ out, int n) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) out[c] = in[c] != T(0) ? 1 : 0;
}

__global__ void k_copy_floats(const float* __restrict__ a, float* __restrict__ b, int n) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) b[c] = a[c];
}

__global__ void k_fill_floats(float* __restrict__ a, int n, float v) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) a[c] = v;
}

// ---------------------------------------------------------- fluid marking
// markFluidCellsKernel (grid_flip_solver3.js:237-245) after clearFluidMask:
// the same-value race of the reference is benign (every writer stores 1).
__global__ void k_mark_fluid(G3 g, const float* __restrict__ pos, int count, float* __restrict__ mask) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= count) return;
    const float x = (pos[3 * p + 0] - g.o0) / g.h0;
    const float y = (pos[3 * p + 1] - g.o1) / g.h1;
    const float z = (pos[3 * p + 2] - g.o2) / g.h2;
    const int i = clampi((int)floorf(x), 0, g.n0 - 1);
    const int j = clampi((int)floorf(y), 0, g.n1 - 1);
    const int k = clampi((int)floorf(z), 0, g.n2 - 1);
    mask[i + g.n0 * (j + g.n1 * k)] = 1.0f;
}

// computeU/V/WFluidAdjacent (grid_flip_solver3.js:595-625)
__global__ void k_fluid_adjacent(G3 g, int axis, const float* __restrict__ mask, int* __restrict__ out) {
    const int s0 = g.n0 + (axis == 0), s1 = g.n1 + (axis == 1), s2 = g.n2 + (axis == 2);
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= s0 * s1 * s2) return;
    int I[3] = {f % s0, (f / s0) % s1, f / (s0 * s1)};
    const int n[3] = {g.n0, g.n1, g.n2};
    int lo[3] = {I[0], I[1], I[2]}, hi[3] = {I[0], I[1], I[2]};
    lo[axis] = clampi(I[axis] - 1, 0, n[axis] - 1);
    hi[axis] = clampi(I[axis], 0, n[axis] - 1);
    const float a = mask[lo[0] + g.n0 * (lo[1] + g.n1 * lo[2])];
    const float b = mask[hi[0] + g.n0 * (hi[1] + g.n1 * hi[2])];
    out[f] = (a > 0.5f || b > 0.5f) ? 1 : 0;
}

// ------------------------------------------------------- collider geometry
// primitiveDistanceCpu / primitivesDistanceCpu (primitive_sdf3.js:149-208) in FP64
__device__ double prim_distance(const flux_primitive3& sh, double x, double y, double z) {
    const double lx = x - sh.center[0], ly = y - sh.center[1], lz = z - sh.center[2];
    if (sh.type == 0) return hypot(hypot(lx, ly), lz) - sh.radius;
    double q0 = lx, q1 = ly, q2 = lz;
    if (sh.rotated) {
        // rotateAroundAxisCpu(v, normalizeAxis(axis), -angle) (Rodrigues)
        const double ax = sh.axis[0], ay = sh.axis[1], az = sh.axis[2];
        const double len = hypot(hypot(ax, ay), az);
        double kx = 0, ky = 0, kz = 1;
        if (len > 0) {
            kx = ax / len;
            ky = ay / len;
            kz = az / len;
        }
        const double ang = -sh.angle;
        const double c = cos(ang), s = sin(ang);
        const double cx = ky * lz - kz * ly, cy = kz * lx - kx * lz, cz = kx * ly - ky * lx;
        const double kd = kx * lx + ky * ly + kz * lz;
        q0 = lx * c + cx * s + kx * kd * (1 - c);
        q1 = ly * c + cy * s + ky * kd * (1 - c);
        q2 = lz * c + cz * s + kz * kd * (1 - c);
    }
    const double dx = fabs(q0) - sh.half[0], dy = fabs(q1) - sh.half[1], dz = fabs(q2) - sh.half[2];
    const double outside = hypot(hypot(fmax(dx, 0.0), fmax(dy, 0.0)), fmax(dz, 0.0));
    const double inside = fmin(fmax(dx, fmax(dy, dz)), 0.0);
    return outside + inside;
}

__device__ double prims_distance(const flux_primitive3* __restrict__ shapes, int count, double x, double y,
                                 double z) {
    double d = INFINITY;
    for (int q = 0; q < count; ++q) d = fmin(d, prim_distance(shapes[q], x, y, z));
    return d;
}

// fractionInsideSdfCpu / fractionInsideCpu (level_set_utils.js:51-161)
__device__ __forceinline__ double frac_sdf(double p0, double p1) {
    const bool in0 = p0 < 0, in1 = p1 < 0;
    if (in0 && in1) return 1;
    if (in0 && !in1) return p0 / (p0 - p1);
    if (!in0 && in1) return p1 / (p1 - p0);
    return 0;
}

__device__ __forceinline__ void cycle4(double* a) {
    const double t = a[0];
    a[0] = a[1];
    a[1] = a[2];
    a[2] = a[3];
    a[3] = t;
}

__device__ double fraction_inside(double bl, double br, double tl, double tr) {
    const int inside = (bl < 0) + (tl < 0) + (br < 0) + (tr < 0);
    double l[4] = {bl, br, tr, tl};
    if (inside == 4) return 1;
    if (inside == 3) {
        for (int g = 0; g < 4 && l[0] < 0; ++g) cycle4(l);
        const double s0 = 1 - frac_sdf(l[0], l[3]);
        const double s1 = 1 - frac_sdf(l[0], l[1]);
        return 1 - 0.5 * s0 * s1;
    }
    if (inside == 2) {
        for (int g = 0; g < 4 && (l[0] >= 0 || !(l[1] < 0 || l[2] < 0)); ++g) cycle4(l);
        if (l[1] < 0) return 0.5 * (frac_sdf(l[0], l[3]) + frac_sdf(l[1], l[2]));
        const double mid = 0.25 * (l[0] + l[1] + l[2] + l[3]);
        if (mid < 0) {
            double area = 0;
            const double s1 = 1 - frac_sdf(l[0], l[3]);
            const double s3 = 1 - frac_sdf(l[2], l[3]);
            area += 0.5 * s1 * s3;
            const double s2 = 1 - frac_sdf(l[2], l[1]);
            const double s0 = 1 - frac_sdf(l[0], l[1]);
            area += 0.5 * s0 * s2;
            return 1 - area;
        }
        double area = 0;
        const double s0 = frac_sdf(l[0], l[1]);
        const double s1 = frac_sdf(l[0], l[3]);
        area += 0.5 * s0 * s1;
        const double s2 = frac_sdf(l[2], l[1]);
        const double s3 = frac_sdf(l[2], l[3]);
        area += 0.5 * s2 * s3;
        return area;
    }
    if (inside == 1) {
        for (int g = 0; g < 4 && l[0] >= 0; ++g) cycle4(l);
        return 0.5 * frac_sdf(l[0], l[3]) * frac_sdf(l[0], l[1]);
    }
    return 0;
}

struct G3D {
    int n0, n1, n2;
    double h[3];
    double o[3];
};

// computeFaceWeights (sdf_collider3.js:131-262): the open fraction of each
// face's cross-section square, floored at 0.01 when open, 0 when all four
// corners lie on or inside the surface.
__global__ void k_face_weights(G3D g, int axis, const flux_primitive3* __restrict__ shapes, int count, double sign,
                               float* __restrict__ out) {
    const int s0 = g.n0 + (axis == 0), s1 = g.n1 + (axis == 1), s2 = g.n2 + (axis == 2);
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= s0 * s1 * s2) return;
    const int i = f % s0, j = (f / s0) % s1, k = f / (s0 * s1);
    const double hx = 0.5 * g.h[0], hy = 0.5 * g.h[1], hz = 0.5 * g.h[2];
    double o[3] = {g.o[0] + hx, g.o[1] + hy, g.o[2] + hz};
    o[axis] = g.o[axis];
    const double x = o[0] + i * g.h[0], y = o[1] + j * g.h[1], z = o[2] + k * g.h[2];
    double off[4][3];
    if (axis == 0) {
        const double t[4][3] = {{0, -hy, -hz}, {0, hy, -hz}, {0, -hy, hz}, {0, hy, hz}};
        for (int q = 0; q < 4; ++q)
            for (int a = 0; a < 3; ++a) off[q][a] = t[q][a];
    } else if (axis == 1) {
        const double t[4][3] = {{-hx, 0, -hz}, {hx, 0, -hz}, {-hx, 0, hz}, {hx, 0, hz}};
        for (int q = 0; q < 4; ++q)
            for (int a = 0; a < 3; ++a) off[q][a] = t[q][a];
    } else {
        const double t[4][3] = {{-hx, -hy, 0}, {hx, -hy, 0}, {-hx, hy, 0}, {hx, hy, 0}};
        for (int q = 0; q < 4; ++q)
            for (int a = 0; a < 3; ++a) off[q][a] = t[q][a];
    }
    double c[4];
    for (int q = 0; q < 4; ++q) c[q] = sign * prims_distance(shapes, count, x + off[q][0], y + off[q][1], z + off[q][2]);
    const double frac = fraction_inside(c[0], c[1], c[2], c[3]);
    double w = fmin(1.0, fmax(0.0, 1 - frac));
    if (c[0] <= 0 && c[1] <= 0 && c[2] <= 0 && c[3] <= 0) w = 0;
    else if (w > 0 && w < 0.01) w = 0.01;  // MIN_FACE_WEIGHT
    out[f] = (float)w;
}

// addShapes (sdf_collider3.js:71-110): SDF at cell centres
__global__ void k_collider_sdf(G3D g, const flux_primitive3* __restrict__ shapes, int count, double sign,
                               float* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.n0 * g.n1 * g.n2) return;
    const int i = c % g.n0, j = (c / g.n0) % g.n1, k = c / (g.n0 * g.n1);
    const double x = g.o[0] + 0.5 * g.h[0] + i * g.h[0];
    const double y = g.o[1] + 0.5 * g.h[1] + j * g.h[1];
    const double z = g.o[2] + 0.5 * g.h[2] + k * g.h[2];
    out[c] = (float)(sign * prims_distance(shapes, count, x, y, z));
}

// ------------------------------------------------------ velocity constraint
// fillMarker + markAndProject (…solver3.js:200-262) for a static collider:
// an open face is fluid; a closed one takes the collider's velocity (0).
__global__ void k_mark_and_project(G3 g, int axis, const float* __restrict__ sdf, const float* __restrict__ weight,
                                   float* __restrict__ data, int* __restrict__ marker) {
    const int s[3] = {g.n0 + (axis == 0), g.n1 + (axis == 1), g.n2 + (axis == 2)};
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= s[0] * s[1] * s[2]) return;
    float open;
    if (weight) {
        open = weight[f];
    } else {
        const int I[3] = {f % s[0], (f / s[0]) % s[1], f / (s[0] * s[1])};
        const float h[3] = {g.h0, g.h1, g.h2};
        float fo[3];
        mac_origin(g, axis, fo);
        float pt[3];
        for (int a = 0; a < 3; ++a) pt[a] = fo[a] + (float)I[a] * h[a];
        const int n[3] = {g.n0, g.n1, g.n2};
        const float so[3] = {g.o0 + 0.5f * g.h0, g.o1 + 0.5f * g.h1, g.o2 + 0.5f * g.h2};
        float pa[3] = {pt[0], pt[1], pt[2]}, pb[3] = {pt[0], pt[1], pt[2]};
        pa[axis] = pt[axis] - h[axis] * 0.5f;
        pb[axis] = pt[axis] + h[axis] * 0.5f;
        const float s0 = sample_tri(sdf, n, so, h, pa);
        const float s1 = sample_tri(sdf, n, so, h, pb);
        const bool in0 = s0 < 0, in1 = s1 < 0;
        float fr = (in0 && in1) ? 1.0f : (in0 ? s0 / (s0 - s1) : (in1 ? s1 / (s1 - s0) : 0.0f));
        fr = fr < 0.0f ? 0.0f : (fr > 1.0f ? 1.0f : fr);
        open = 1.0f - fr;
    }
    if (open > 0) {
        marker[f] = 1;
    } else {
        data[f] = 0.0f;
        marker[f] = 0;
    }
}

// noFluxProjection (…solver3.js:264-305) with friction
// (projectAndApplyFriction, 29-45) for a static collider, into `tmp`.
__global__ void k_no_flux(G3 g, int axis, Mac m, const float* __restrict__ sdf, const float* __restrict__ weight,
                          float friction, const float* __restrict__ data, float* __restrict__ tmp) {
    const int s[3] = {g.n0 + (axis == 0), g.n1 + (axis == 1), g.n2 + (axis == 2)};
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= s[0] * s[1] * s[2]) return;
    const int I[3] = {f % s[0], (f / s[0]) % s[1], f / (s[0] * s[1])};
    const float h[3] = {g.h0, g.h1, g.h2};
    const int n[3] = {g.n0, g.n1, g.n2};
    const float so[3] = {g.o0 + 0.5f * g.h0, g.o1 + 0.5f * g.h1, g.o2 + 0.5f * g.h2};
    float fo[3];
    mac_origin(g, axis, fo);
    float pt[3];
    for (int a = 0; a < 3; ++a) pt[a] = fo[a] + (float)I[a] * h[a];
    const bool blocked = weight ? (weight[f] <= 0.0f) : (sample_tri(sdf, n, so, h, pt) < 0.0f);
    if (!blocked) {
        tmp[f] = data[f];
        return;
    }
    float vel[3];
    sample_mac(g, m, pt, vel);
    float gr[3];
    sdf_gradient(sdf, n, so, h, pt, gr);
    const float gl = sqrtf(gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2]);
    if (!(gl > 0.0f)) {
        tmp[f] = 0.0f;
        return;
    }
    const float nn[3] = {gr[0] / gl, gr[1] / gl, gr[2] / gl};
    const float dn = vel[0] * nn[0] + vel[1] * nn[1] + vel[2] * nn[2];
    float vt[3] = {vel[0] - nn[0] * dn, vel[1] - nn[1] * dn, vel[2] - nn[2] * dn};
    const float tl = sqrtf(vt[0] * vt[0] + vt[1] * vt[1] + vt[2] * vt[2]);
    if (tl > 0.0f) {
        const float veln = fmaxf(-dn, 0.0f);
        const float scale = fmaxf(1.0f - (veln * friction) / tl, 0.0f);
        vt[0] *= scale;
        vt[1] *= scale;
        vt[2] *= scale;
    }
    tmp[f] = vt[axis] + 0.0f;
}

// blockedBoundary (…solver3.js:309-366): faces between a collider cell and a
// fluid cell take the collider's velocity (0).
__global__ void k_blocked_boundary(G3 g, const float* __restrict__ sdf, float* __restrict__ u, float* __restrict__ v,
                                   float* __restrict__ w) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int n0 = g.n0, n1 = g.n1, n2 = g.n2;
    if (c >= n0 * n1 * n2) return;
    if (!(sdf[c] < 0.0f)) return;
    const int i = c % n0, j = (c / n0) % n1, k = c / (n0 * n1);
    auto fluid = [&](int ii, int jj, int kk) { return !(sdf[ii + n0 * (jj + n1 * kk)] < 0.0f); };
    const int su0 = n0 + 1, sv1 = n1 + 1;
    if (i > 0 && fluid(i - 1, j, k)) u[i + su0 * (j + n1 * k)] = 0.0f;
    if (i < n0 - 1 && fluid(i + 1, j, k)) u[i + 1 + su0 * (j + n1 * k)] = 0.0f;
    if (j > 0 && fluid(i, j - 1, k)) v[i + n0 * (j + sv1 * k)] = 0.0f;
    if (j < n1 - 1 && fluid(i, j + 1, k)) v[i + n0 * (j + 1 + sv1 * k)] = 0.0f;
    if (k > 0 && fluid(i, j, k - 1)) w[i + n0 * (j + n1 * k)] = 0.0f;
    if (k < n2 - 1 && fluid(i, j, k + 1)) w[i + n0 * (j + n1 * (k + 1))] = 0.0f;
}

// zeroULeft ... zeroWFront (…solver3.js:118-125) by closedDomainBoundaryFlag,
// then clampVelocityU/V/W (±100, non-finite -> 0, 74-96), one face field per launch.
// inflow objects (createSDFInflow3): SDF cell-centred on the velocity grid,
// velocity, mode 0 'set' / 1 'add'
static constexpr int kMaxInflows3 = 4;
struct Inflows3 {
    const float* sdf[kMaxInflows3];
    float vel[kMaxInflows3][3];
    int mode[kMaxInflows3];
    int count;
};

// The walls by flag (zeroULeft ... zeroWFront), then every inflow in order
// (buildInflowKernels, grid_blocked_boundary_condition_solver3.js:130-176:
// a face whose position is inside the inflow's SDF takes or adds its
// velocity component), then the +-100 clamp: each is a per-face update, so
// one thread applies them to its face in the reference's order.
__global__ void k_walls_and_clamp(G3 g, int axis, int flags, Inflows3 inflows, float* __restrict__ d) {
    const int s[3] = {g.n0 + (axis == 0), g.n1 + (axis == 1), g.n2 + (axis == 2)};
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= s[0] * s[1] * s[2]) return;
    const int I[3] = {f % s[0], (f / s[0]) % s[1], f / (s[0] * s[1])};
    float v = d[f];
    const int lo_bit = 1 << (2 * axis), hi_bit = 1 << (2 * axis + 1);
    if ((flags & lo_bit) && I[axis] == 0) v = 0.0f;
    if ((flags & hi_bit) && I[axis] == s[axis] - 1) v = 0.0f;
    if (inflows.count > 0) {
        const int n[3] = {g.n0, g.n1, g.n2};
        const float h[3] = {g.h0, g.h1, g.h2};
        const float so[3] = {g.o0 + 0.5f * g.h0, g.o1 + 0.5f * g.h1, g.o2 + 0.5f * g.h2};
        float fo[3];
        mac_origin(g, axis, fo);
        const float pt[3] = {fo[0] + (float)I[0] * h[0], fo[1] + (float)I[1] * h[1], fo[2] + (float)I[2] * h[2]};
        for (int q = 0; q < inflows.count; ++q) {
            if (!(sample_tri(inflows.sdf[q], n, so, h, pt) < 0.0f)) continue;
            const float value = inflows.vel[q][axis];
            v = inflows.mode[q] == 1 ? v + value : value;
        }
    }
    d[f] = !isfinite(v) ? 0.0f : fminf(fmaxf(v, -100.0f), 100.0f);
}

// createOutflowPressureDirichlet3 + combineDirichlet3(outflow, user)
// (sdf_inflow_outflow3.js:44-94) at the cell centres: active = inside any
// outflow (its SDF, cell-centred on this grid, sampled < 0) or the user's
// mask; target = the user's target where the user's mask is set, else 0
static constexpr int kMaxOutflows3 = 4;
struct Outflows3 {
    const float* sdf[kMaxOutflows3];
    int count;
};

__global__ void k_outflow_dirichlet3(G3 g, Outflows3 o, const uint8_t* __restrict__ user_active,
                                     const float* __restrict__ user_target, uint8_t* __restrict__ active,
                                     float* __restrict__ target) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.n0 * g.n1 * g.n2) return;
    const int i = c % g.n0, j = (c / g.n0) % g.n1, k = c / (g.n0 * g.n1);
    const int n[3] = {g.n0, g.n1, g.n2};
    const float h[3] = {g.h0, g.h1, g.h2};
    const float so[3] = {g.o0 + g.h0 * 0.5f, g.o1 + g.h1 * 0.5f, g.o2 + g.h2 * 0.5f};
    const float p[3] = {so[0] + h[0] * (float)i, so[1] + h[1] * (float)j, so[2] + h[2] * (float)k};
    bool act = false;
    for (int q = 0; q < o.count; ++q) act = act || sample_tri(o.sdf[q], n, so, h, p) < 0.0f;
    float tg = 0.0f;
    if (user_active) {
        const bool ua = user_active[c] != 0;
        act = act || ua;
        tg = ua ? (user_target ? user_target[c] : 0.0f) : 0.0f;
    }
    active[c] = (uint8_t)act;
    target[c] = tg;
}

}  // namespace flux

using namespace flux;

struct flux_extrapolator {
    int n[3] = {1, 1, 1};
    int64_t cells = 0;
    const flux_field* value = nullptr;
    const flux_field* valid = nullptr;
    flux_field* out = nullptr;
    float* scratch = nullptr;
    int* va = nullptr;
    int* vb = nullptr;
};

namespace {

int field_shape3(const flux_field* f, int n[3]) {
    for (int a = 0; a < 3; ++a) n[a] = (int)(a < f->ndim ? f->shape[a] : 1);
    return FLUX_OK;
}

// createExtrapolateToRegion3.run (array_utils3.js:187-222)
int extrapolate_run(flux_extrapolator* e, int iterations) {
    const int n = (int)e->cells;
    const unsigned g = blocks_for(n, FTPB);
    float* out = (float*)e->out->d;
    if (e->out->d != e->value->d) k_copy_floats<<<g, FTPB, 0, g_stream>>>((const float*)e->value->d, out, n);
    switch (e->valid->dtype) {
        case FLUX_I32: k_valid_to_int<int><<<g, FTPB, 0, g_stream>>>((const int*)e->valid->d, e->va, n); break;
        case FLUX_U8: k_valid_to_int<uint8_t><<<g, FTPB, 0, g_stream>>>((const uint8_t*)e->valid->d, e->va, n); break;
        default: k_valid_to_int<float><<<g, FTPB, 0, g_stream>>>((const float*)e->valid->d, e->va, n); break;
    }
    for (int it = 0; it < iterations; ++it) {
        if ((it & 1) == 0)
            k_extrap_step<<<g, FTPB, 0, g_stream>>>(e->n[0], e->n[1], e->n[2], out, e->scratch, e->va, e->vb);
        else
            k_extrap_step<<<g, FTPB, 0, g_stream>>>(e->n[0], e->n[1], e->n[2], e->scratch, out, e->vb, e->va);
    }
    if (iterations & 1) k_copy_floats<<<g, FTPB, 0, g_stream>>>(e->scratch, out, n);
    FLUX_LAUNCH_CHECK();
    return FLUX_OK;
}

G3D make_g3d(const flux_grid3d* g) {
    G3D d;
    d.n0 = g->n[0];
    d.n1 = g->n[1];
    d.n2 = g->n[2];
    for (int a = 0; a < 3; ++a) {
        d.h[a] = g->h[a];
        d.o[a] = g->origin[a];
    }
    return d;
}

int check_grid3(const int n[3], const char* who) {
    int64_t cells = 1;
    for (int a = 0; a < 3; ++a) {
        FLUX_REQUIRE(n[a] > 0, "%s: grid resolution must be positive", who);
        cells *= n[a] + 1;
    }
    FLUX_REQUIRE(cells < (1ll << 31), "%s: grid too large", who);
    return FLUX_OK;
}

int64_t face_count(const int n[3], int axis) {
    return (int64_t)(n[0] + (axis == 0)) * (n[1] + (axis == 1)) * (n[2] + (axis == 2));
}

}  // namespace

struct flux_boundary3 {
    flux_grid3 g;
    flux_field* vel[3] = {nullptr, nullptr, nullptr};
    const flux_field* sdf = nullptr;
    const flux_field* weights[3] = {nullptr, nullptr, nullptr};
    float friction = 0.5f;
    int flags = 63;
    int* marker[3] = {nullptr, nullptr, nullptr};
    float* tmp[3] = {nullptr, nullptr, nullptr};
    flux_field* marker_field[3] = {nullptr, nullptr, nullptr};
    flux_extrapolator* ex[3] = {nullptr, nullptr, nullptr};
    Inflows3 inflows{};
};

extern "C" {

int flux_extrapolator3_create(const flux_field* value, const flux_field* valid, flux_field* out,
                              flux_extrapolator** e) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(value && valid && out && e, "extrapolate: NULL argument");
    FLUX_REQUIRE(value->dtype == FLUX_F32 && out->dtype == FLUX_F32, "extrapolate: value fields must be float");
    FLUX_REQUIRE(valid->dtype == FLUX_I32 || valid->dtype == FLUX_U8 || valid->dtype == FLUX_F32,
                 "extrapolate: valid must be int, u8 or float");
    FLUX_REQUIRE(value->count == valid->count && value->count == out->count, "extrapolate: shape mismatch");
    FLUX_REQUIRE(value->count < (1ll << 31), "extrapolate: field too large");
    flux_extrapolator* x = new flux_extrapolator();
    field_shape3(value, x->n);
    x->cells = value->count;
    x->value = value;
    x->valid = valid;
    x->out = out;
    x->scratch = (float*)dmalloc(sizeof(float) * x->cells);
    x->va = (int*)dmalloc(sizeof(int) * x->cells);
    x->vb = (int*)dmalloc(sizeof(int) * x->cells);
    if (!x->scratch || !x->va || !x->vb) {
        flux_extrapolator3_destroy(x);
        return set_error(FLUX_ERR_NOMEM, "extrapolate: allocation failed");
    }
    *e = x;
    return FLUX_OK;
}

int flux_extrapolator3_run(flux_extrapolator* e, int iterations) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(e && iterations >= 0, "extrapolate: bad arguments");
    return extrapolate_run(e, iterations);
}

int flux_extrapolator3_destroy(flux_extrapolator* e) {
    if (!e) return FLUX_OK;
    if (g_init) (void)hipStreamSynchronize(g_stream);
    dfree(e->scratch);
    dfree(e->va);
    dfree(e->vb);
    delete e;
    return FLUX_OK;
}

int flux_mark_fluid_cells(const flux_grid3* g, const flux_field* positions, int64_t count, flux_field* fluid_mask) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(g && positions && fluid_mask, "markFluidCells: NULL argument");
    FLUX_TRY(check_grid3(g->n, "markFluidCells"));
    const int64_t cells = (int64_t)g->n[0] * g->n[1] * g->n[2];
    FLUX_TRY(check_field(fluid_mask, FLUX_F32, cells, "fluid mask"));
    FLUX_REQUIRE(count >= 0 && count < (1ll << 31), "markFluidCells: bad particle count");
    FLUX_REQUIRE(positions->dtype == FLUX_VEC3 && positions->count >= count, "markFluidCells: positions must be vec3");
    k_fill_floats<<<blocks_for(cells, FTPB), FTPB, 0, g_stream>>>((float*)fluid_mask->d, (int)cells, 0.0f);
    if (count > 0)
        k_mark_fluid<<<blocks_for(count, FTPB), FTPB, 0, g_stream>>>(make_g3(g), (const float*)positions->d, (int)count,
                                                                     (float*)fluid_mask->d);
    FLUX_LAUNCH_CHECK();
    return FLUX_OK;
}

int flux_fluid_adjacent(const flux_grid3* g, const flux_field* fluid_mask, int axis, flux_field* out) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(g && fluid_mask && out, "fluidAdjacent: NULL argument");
    FLUX_REQUIRE(axis >= 0 && axis < 3, "fluidAdjacent: axis must be 0, 1 or 2");
    FLUX_TRY(check_grid3(g->n, "fluidAdjacent"));
    const int64_t cells = (int64_t)g->n[0] * g->n[1] * g->n[2];
    FLUX_TRY(check_field(fluid_mask, FLUX_F32, cells, "fluid mask"));
    const int64_t nf = face_count(g->n, axis);
    FLUX_TRY(check_field(out, FLUX_I32, nf, "fluid-adjacent faces"));
    k_fluid_adjacent<<<blocks_for(nf, FTPB), FTPB, 0, g_stream>>>(make_g3(g), axis, (const float*)fluid_mask->d,
                                                                  (int*)out->d);
    FLUX_LAUNCH_CHECK();
    return FLUX_OK;
}

int flux_collider_face_weights(const flux_grid3d* g, const flux_primitive3* shapes, int count, int invert,
                               flux_field* const* weights) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(g && weights && (shapes || count == 0) && count >= 0, "computeFaceWeights: bad arguments");
    FLUX_TRY(check_grid3(g->n, "computeFaceWeights"));
    for (int q = 0; q < count; ++q)
        FLUX_REQUIRE(shapes[q].type == 0 || shapes[q].type == 1, "primitive_sdf3: unknown shape type '%d'.",
                     shapes[q].type);
    for (int a = 0; a < 3; ++a) FLUX_TRY(check_field(weights[a], FLUX_F32, face_count(g->n, a), "face weights"));
    flux_primitive3* dsh = nullptr;
    if (count > 0) {
        dsh = (flux_primitive3*)dmalloc(sizeof(flux_primitive3) * count);
        FLUX_REQUIRE(dsh, "computeFaceWeights: allocation failed");
        FLUX_HIP(hipMemcpyAsync(dsh, shapes, sizeof(flux_primitive3) * count, hipMemcpyHostToDevice, g_stream));
    }
    const G3D gd = make_g3d(g);
    for (int a = 0; a < 3; ++a) {
        const int64_t nf = face_count(g->n, a);
        k_face_weights<<<blocks_for(nf, FTPB), FTPB, 0, g_stream>>>(gd, a, dsh, count, invert ? -1.0 : 1.0,
                                                                     (float*)weights[a]->d);
    }
    FLUX_LAUNCH_CHECK();
    FLUX_HIP(hipStreamSynchronize(g_stream));
    dfree(dsh);
    return FLUX_OK;
}

int flux_outflow3_dirichlet(const flux_grid3* g, const flux_field* const* outflow_sdf, int count,
                            const flux_field* user_active, const flux_field* user_target, flux_field* active,
                            flux_field* target) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(g && active && target && (count == 0 || outflow_sdf), "createOutflowPressureDirichlet3: NULL argument");
    FLUX_REQUIRE(count >= 0 && count <= kMaxOutflows3, "createOutflowPressureDirichlet3: at most %d outflows, got %d",
                 kMaxOutflows3, count);
    for (int a = 0; a < 3; ++a) FLUX_REQUIRE(g->n[a] > 0 && g->h[a] > 0, "createOutflowPressureDirichlet3: bad grid");
    const int64_t cells = (int64_t)g->n[0] * g->n[1] * g->n[2];
    FLUX_TRY(check_field(active, FLUX_U8, cells, "dirichlet active"));
    FLUX_TRY(check_field(target, FLUX_F32, cells, "dirichlet target"));
    Outflows3 o{};
    o.count = count;
    for (int q = 0; q < count; ++q) {
        FLUX_TRY(check_field(outflow_sdf[q], FLUX_F32, cells, "outflow sdf"));
        o.sdf[q] = (const float*)outflow_sdf[q]->d;
    }
    if (user_active) FLUX_TRY(check_field(user_active, FLUX_U8, cells, "dirichlet active"));
    if (user_target) FLUX_TRY(check_field(user_target, FLUX_F32, cells, "dirichlet target"));
    k_outflow_dirichlet3<<<blocks_for(cells, FTPB), FTPB, 0, g_stream>>>(
        make_g3(g), o, user_active ? (const uint8_t*)user_active->d : nullptr,
        user_target ? (const float*)user_target->d : nullptr, (uint8_t*)active->d, (float*)target->d);
    FLUX_LAUNCH_CHECK();
    return FLUX_OK;
}

int flux_collider_sdf(const flux_grid3d* g, const flux_primitive3* shapes, int count, int invert, flux_field* sdf) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(g && sdf && (shapes || count == 0) && count >= 0, "addShapes: bad arguments");
    FLUX_TRY(check_grid3(g->n, "addShapes"));
    for (int q = 0; q < count; ++q)
        FLUX_REQUIRE(shapes[q].type == 0 || shapes[q].type == 1, "primitive_sdf3: unknown shape type '%d'.",
                     shapes[q].type);
    const int64_t cells = (int64_t)g->n[0] * g->n[1] * g->n[2];
    FLUX_TRY(check_field(sdf, FLUX_F32, cells, "collider sdf"));
    flux_primitive3* dsh = nullptr;
    if (count > 0) {
        dsh = (flux_primitive3*)dmalloc(sizeof(flux_primitive3) * count);
        FLUX_REQUIRE(dsh, "addShapes: allocation failed");
        FLUX_HIP(hipMemcpyAsync(dsh, shapes, sizeof(flux_primitive3) * count, hipMemcpyHostToDevice, g_stream));
    }
    k_collider_sdf<<<blocks_for(cells, FTPB), FTPB, 0, g_stream>>>(make_g3d(g), dsh, count, invert ? -1.0 : 1.0,
                                                                   (float*)sdf->d);
    FLUX_LAUNCH_CHECK();
    FLUX_HIP(hipStreamSynchronize(g_stream));
    dfree(dsh);
    return FLUX_OK;
}

int flux_boundary3_create(const flux_grid3* g, flux_field* const* vel, const flux_field* sdf,
                          const flux_field* const* face_weights, float friction, flux_boundary3** out) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(g && vel && out, "boundary: NULL argument");
    FLUX_TRY(check_grid3(g->n, "boundary"));
    for (int a = 0; a < 3; ++a) FLUX_TRY(check_field(vel[a], FLUX_F32, face_count(g->n, a), "velocity"));
    const int64_t cells = (int64_t)g->n[0] * g->n[1] * g->n[2];
    if (sdf) FLUX_TRY(check_field(sdf, FLUX_F32, cells, "collider sdf"));
    if (face_weights)
        for (int a = 0; a < 3; ++a) FLUX_TRY(check_field(face_weights[a], FLUX_F32, face_count(g->n, a), "face weights"));
    flux_boundary3* b = new flux_boundary3();
    b->g = *g;
    b->sdf = sdf;
    b->friction = friction;
    bool ok = true;
    for (int a = 0; a < 3; ++a) {
        b->vel[a] = vel[a];
        if (face_weights) b->weights[a] = face_weights[a];
        if (sdf) {
            // the marker is an int field (the extrapolation's valid flags); tmp is scratch
            int64_t shape[3] = {g->n[0] + (a == 0), g->n[1] + (a == 1), g->n[2] + (a == 2)};
            ok = ok && flux_field_create(FLUX_I32, 3, shape, &b->marker_field[a]) == FLUX_OK;
            b->tmp[a] = (float*)dmalloc(sizeof(float) * face_count(g->n, a));
            ok = ok && b->tmp[a];
            if (ok) {
                b->marker[a] = (int*)b->marker_field[a]->d;
                ok = flux_extrapolator3_create(vel[a], b->marker_field[a], vel[a], &b->ex[a]) == FLUX_OK;
            }
        }
    }
    if (!ok) {
        flux_boundary3_destroy(b);
        return set_error(FLUX_ERR_NOMEM, "boundary: allocation failed");
    }
    *out = b;
    return FLUX_OK;
}

int flux_boundary3_set_closed(flux_boundary3* b, int flags) {
    FLUX_REQUIRE(b, "boundary: NULL argument");
    FLUX_REQUIRE(flags >= 0 && flags < 64, "boundary: closedDomainBoundaryFlag must be a 6-bit mask");
    b->flags = flags;
    return FLUX_OK;
}

// constrainVelocityPlan (…solver3.js:389-436)
int flux_boundary3_constrain(flux_boundary3* b, int depth) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(b && depth >= 0, "constrainVelocity: bad arguments");
    const G3 g = make_g3(&b->g);
    if (b->sdf) {
        for (int a = 0; a < 3; ++a) {
            const int64_t nf = face_count(b->g.n, a);
            k_mark_and_project<<<blocks_for(nf, FTPB), FTPB, 0, g_stream>>>(
                g, a, (const float*)b->sdf->d, b->weights[a] ? (const float*)b->weights[a]->d : nullptr,
                (float*)b->vel[a]->d, b->marker[a]);
        }
        for (int a = 0; a < 3; ++a) FLUX_TRY(extrapolate_run(b->ex[a], depth));
        Mac m{(const float*)b->vel[0]->d, (const float*)b->vel[1]->d, (const float*)b->vel[2]->d};
        for (int a = 0; a < 3; ++a) {
            const int64_t nf = face_count(b->g.n, a);
            k_no_flux<<<blocks_for(nf, FTPB), FTPB, 0, g_stream>>>(
                g, a, m, (const float*)b->sdf->d, b->weights[a] ? (const float*)b->weights[a]->d : nullptr,
                b->friction, (const float*)b->vel[a]->d, b->tmp[a]);
        }
        for (int a = 0; a < 3; ++a) {
            const int64_t nf = face_count(b->g.n, a);
            k_copy_floats<<<blocks_for(nf, FTPB), FTPB, 0, g_stream>>>(b->tmp[a], (float*)b->vel[a]->d, (int)nf);
        }
        const int64_t cells = (int64_t)b->g.n[0] * b->g.n[1] * b->g.n[2];
        k_blocked_boundary<<<blocks_for(cells, FTPB), FTPB, 0, g_stream>>>(
            g, (const float*)b->sdf->d, (float*)b->vel[0]->d, (float*)b->vel[1]->d, (float*)b->vel[2]->d);
    }
    for (int a = 0; a < 3; ++a) {
        const int64_t nf = face_count(b->g.n, a);
        k_walls_and_clamp<<<blocks_for(nf, FTPB), FTPB, 0, g_stream>>>(g, a, b->flags, b->inflows,
                                                                      (float*)b->vel[a]->d);
    }
    FLUX_LAUNCH_CHECK();
    return FLUX_OK;
}

int flux_boundary3_add_inflow(flux_boundary3* b, const flux_field* sdf, float vx, float vy, float vz, int mode) {
    FLUX_REQUIRE(b && sdf, "setInflows: NULL argument");
    FLUX_REQUIRE(b->inflows.count < kMaxInflows3, "setInflows: at most %d inflows", kMaxInflows3);
    FLUX_REQUIRE(mode == 0 || mode == 1, "setInflows: mode must be 'set' (0) or 'add' (1)");
    FLUX_TRY(check_field(sdf, FLUX_F32, (int64_t)b->g.n[0] * b->g.n[1] * b->g.n[2], "inflow sdf"));
    const int q = b->inflows.count++;
    b->inflows.sdf[q] = (const float*)sdf->d;
    b->inflows.vel[q][0] = vx;
    b->inflows.vel[q][1] = vy;
    b->inflows.vel[q][2] = vz;
    b->inflows.mode[q] = mode;
    return FLUX_OK;
}

int flux_boundary3_destroy(flux_boundary3* b) {
    if (!b) return FLUX_OK;
    if (g_init) (void)hipStreamSynchronize(g_stream);
    for (int a = 0; a < 3; ++a) {
        flux_extrapolator3_destroy(b->ex[a]);
        if (b->marker_field[a]) flux_field_destroy(b->marker_field[a]);
        dfree(b->tmp[a]);
    }
    delete b;
    return FLUX_OK;
}

}  // extern "C"
