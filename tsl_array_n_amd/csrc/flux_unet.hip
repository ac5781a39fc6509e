// Learned initial guess: a 3-D conv U-Net trained and run by this library's
// own kernels (no reference counterpart: the reference has no learned part).
//
// Network (4 levels, channels 16/32/64/64, 3x3x3 kernels, zero padding,
// ReLU, 2x2x2 average pooling, nearest upsampling, skip concat):
//   enc0a 3->16, enc0b 16->16 | pool | enc1a 16->32, enc1b 32->32 | pool |
//   enc2a 32->64, enc2b 64->64 | pool | mid 64->64 | up, cat s2 | dec2 128->64 |
//   up, cat s1 | dec1a 96->32, dec1b 32->32 | up, cat s0 | dec0a 48->16,
//   dec0b 16->16 | out 16->1 (linear)
// Activations are FP32 [C][nz][ny][nx]; master weights FP32 [co][ci][27]
// (tap = dz*9 + dy*3 + dx); optimiser Adam.  Upsampling and concat are
// folded into the conv input loader (channel ci < c1 reads the half-size
// tensor at (z/2, y/2, x/2)).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "flux_internal.h"
#include "flux_solver.h"

namespace flux {

static constexpr int UTPB = 256;

struct ConvIn {
    const float* a;   // first c1 channels (half resolution when up != 0)
    const float* b;   // next c2 channels (full resolution)
    int c1, c2;
    int up;
    const float* mask_y;  // backward-data: multiply input by (mask_y > 0), same layout as a
};

__device__ __forceinline__ float conv_in_load(const ConvIn& in, int ci, int z, int y, int x, int nz, int ny, int nx) {
    if (z < 0 || y < 0 || x < 0 || z >= nz || y >= ny || x >= nx) return 0.0f;
    if (ci < in.c1) {
        if (in.up) {
            int hz = nz >> 1, hy = ny >> 1, hx = nx >> 1;
            int64_t q = (int64_t)ci * hz * hy * hx + ((int64_t)(z >> 1) * hy + (y >> 1)) * hx + (x >> 1);
            return in.a[q];
        }
        int64_t q = (int64_t)ci * nz * ny * nx + ((int64_t)z * ny + y) * nx + x;
        float v = in.a[q];
        if (in.mask_y) v = in.mask_y[q] > 0.0f ? v : 0.0f;
        return v;
    }
    int64_t q = (int64_t)(ci - in.c1) * nz * ny * nx + ((int64_t)z * ny + y) * nx + x;
    return in.b[q];
}

// Direct 3x3x3 convolution, one thread per output voxel, all COUT channels
// in registers; weights w[co][ci][27] are uniform across the block.
template <int COUT>
__global__ void __launch_bounds__(UTPB) k_conv3(ConvIn in, const float* __restrict__ w, const float* __restrict__ bias,
                                                float* __restrict__ out, int cout, int nz, int ny, int nx, int relu) {
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t nv = (int64_t)nz * ny * nx;
    if (v >= nv) return;
    int x = (int)(v % nx);
    int y = (int)((v / nx) % ny);
    int z = (int)(v / ((int64_t)nx * ny));
    const int cin = in.c1 + in.c2;
    float acc[COUT];
#pragma unroll
    for (int co = 0; co < COUT; ++co) acc[co] = (co < cout && bias) ? bias[co] : 0.0f;
    for (int ci = 0; ci < cin; ++ci) {
        float t[27];
#pragma unroll
        for (int dz = 0; dz < 3; ++dz)
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int dx = 0; dx < 3; ++dx)
                    t[dz * 9 + dy * 3 + dx] = conv_in_load(in, ci, z + dz - 1, y + dy - 1, x + dx - 1, nz, ny, nx);
        const float* wc = w + (size_t)ci * 27;
#pragma unroll
        for (int co = 0; co < COUT; ++co) {
            if (co < cout) {
                const float* wk = wc + (size_t)co * cin * 27;
                float s = acc[co];
#pragma unroll
                for (int k = 0; k < 27; ++k) s += wk[k] * t[k];
                acc[co] = s;
            }
        }
    }
#pragma unroll
    for (int co = 0; co < COUT; ++co)
        if (co < cout) out[(int64_t)co * nv + v] = relu ? fmaxf(acc[co], 0.0f) : acc[co];
}

// weight gradient: dW[co][ci][k] = sum_v g[co][v] * in[ci][v + off(k)], with
// g masked by (y > 0) for ReLU layers; one block per (co, ci), fixed-order
// block reduction (deterministic).
__global__ void __launch_bounds__(UTPB) k_conv3_wgrad(ConvIn in, const float* __restrict__ g,
                                                      const float* __restrict__ ymask, float* __restrict__ dw,
                                                      float* __restrict__ db, int cout, int nz, int ny, int nx) {
    const int co = blockIdx.x % cout;
    const int ci = blockIdx.x / cout;
    const int cin = in.c1 + in.c2;
    const int64_t nv = (int64_t)nz * ny * nx;
    float acc[27];
#pragma unroll
    for (int k = 0; k < 27; ++k) acc[k] = 0.0f;
    float accb = 0.0f;
    for (int64_t v = threadIdx.x; v < nv; v += blockDim.x) {
        float gv = g[(int64_t)co * nv + v];
        if (ymask && !(ymask[(int64_t)co * nv + v] > 0.0f)) gv = 0.0f;
        if (gv == 0.0f) continue;
        accb += gv;
        int x = (int)(v % nx);
        int y = (int)((v / nx) % ny);
        int z = (int)(v / ((int64_t)nx * ny));
#pragma unroll
        for (int dz = 0; dz < 3; ++dz)
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int dx = 0; dx < 3; ++dx)
                    acc[dz * 9 + dy * 3 + dx] += gv * conv_in_load(in, ci, z + dz - 1, y + dy - 1, x + dx - 1, nz, ny, nx);
    }
    __shared__ float red[UTPB];
    for (int k = 0; k <= 27; ++k) {
        red[threadIdx.x] = (k < 27) ? acc[k] : accb;
        __syncthreads();
        for (int s = UTPB / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            if (k < 27) dw[((size_t)co * cin + ci) * 27 + k] = red[0];
            else if (ci == 0) db[co] = red[0];
        }
        __syncthreads();
    }
}

// Wt[ci][co][26-k] = W[co][ci][k]: the data gradient is the convolution of the
// masked output gradient with the flipped, transposed kernel.
__global__ void k_flip_transpose(const float* __restrict__ w, float* __restrict__ wt, int cout, int cin) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= cout * cin * 27) return;
    int k = q % 27;
    int ci = (q / 27) % cin;
    int co = q / (27 * cin);
    wt[((size_t)ci * cout + co) * 27 + (26 - k)] = w[q];
}

__global__ void k_pool(const float* __restrict__ in, float* __restrict__ out, int c, int nz, int ny, int nx) {
    int hz = nz >> 1, hy = ny >> 1, hx = nx >> 1;
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t n = (int64_t)c * hz * hy * hx;
    if (q >= n) return;
    int x = (int)(q % hx), y = (int)((q / hx) % hy), z = (int)((q / ((int64_t)hx * hy)) % hz);
    int ch = (int)(q / ((int64_t)hx * hy * hz));
    const float* src = in + (int64_t)ch * nz * ny * nx;
    float s = 0.0f;
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) s += src[((int64_t)(2 * z + dz) * ny + (2 * y + dy)) * nx + (2 * x + dx)];
    out[q] = s * 0.125f;
}

// dst(full) += pool_backward(g(half)) = g/8 broadcast
__global__ void k_pool_back_add(const float* __restrict__ g, float* __restrict__ dst, int c, int nz, int ny, int nx) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t n = (int64_t)c * nz * ny * nx;
    if (q >= n) return;
    int x = (int)(q % nx), y = (int)((q / nx) % ny), z = (int)((q / ((int64_t)nx * ny)) % nz);
    int ch = (int)(q / ((int64_t)nx * ny * nz));
    int hz = nz >> 1, hy = ny >> 1, hx = nx >> 1;
    dst[q] += 0.125f * g[(int64_t)ch * hz * hy * hx + ((int64_t)(z >> 1) * hy + (y >> 1)) * hx + (x >> 1)];
}

// half-res dst = up_backward(g(full)): sum of the 8 children
__global__ void k_up_back(const float* __restrict__ g, float* __restrict__ dst, int c, int nz, int ny, int nx) {
    int hz = nz >> 1, hy = ny >> 1, hx = nx >> 1;
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t n = (int64_t)c * hz * hy * hx;
    if (q >= n) return;
    int x = (int)(q % hx), y = (int)((q / hx) % hy), z = (int)((q / ((int64_t)hx * hy)) % hz);
    int ch = (int)(q / ((int64_t)hx * hy * hz));
    const float* src = g + (int64_t)ch * nz * ny * nx;
    float s = 0.0f;
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) s += src[((int64_t)(2 * z + dz) * ny + (2 * y + dy)) * nx + (2 * x + dx)];
    dst[q] = s;
}

__global__ void k_ucopy(const float* __restrict__ a, float* __restrict__ b, int64_t n) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) b[q] = a[q];
}

// MSE: loss partials and gradient 2 (y - t) / n
__global__ void k_mse(const float* __restrict__ y, const float* __restrict__ t, float* __restrict__ g,
                      double* __restrict__ part, int64_t n) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double d2 = 0.0;
    if (q < n) {
        float d = y[q] - t[q];
        g[q] = 2.0f * d / (float)n;
        d2 = (double)d * d;
    }
    __shared__ double red[UTPB];
    red[threadIdx.x] = d2;
    __syncthreads();
    for (int s = UTPB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       int64_t n, float lr, float b1, float b2, float eps, float c1, float c2) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    float gq = g[q];
    float mq = b1 * m[q] + (1.0f - b1) * gq;
    float vq = b2 * v[q] + (1.0f - b2) * gq * gq;
    m[q] = mq;
    v[q] = vq;
    p[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
}

// input features from the assembled pressure system: ch0 = b / rms(b),
// ch1 = Dirichlet mask, ch2 = log(mean face weight of the cell) / log(ratio)
__global__ void k_features(const float* __restrict__ b, const uint8_t* __restrict__ mask, const float* __restrict__ wu,
                           const float* __restrict__ wv, const float* __restrict__ ww, int nx, int ny, int nz,
                           const double* __restrict__ bb, float inv_log_ratio, float* __restrict__ feat) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t nv = (int64_t)nx * ny * nz;
    if (c >= nv) return;
    int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / ((int64_t)nx * ny));
    float rms = (float)sqrt(fmax(*bb / (double)nv, 1e-300));
    feat[c] = b[c] / rms;
    feat[nv + c] = (mask && mask[c]) ? 1.0f : 0.0f;
    float wsum = 1.0f;
    if (wu) {
        int fu = i + (nx + 1) * (j + ny * k), fv = i + nx * (j + (ny + 1) * k);
        wsum = (wu[fu] + wu[fu + 1] + wv[fv] + wv[fv + nx] + ww[c] + ww[c + (int64_t)nx * ny]) / 6.0f;
    }
    feat[2 * nv + c] = logf(fmaxf(wsum, 1e-6f)) * inv_log_ratio;
}

__global__ void k_sumsq(const float* __restrict__ a, int64_t n, double* __restrict__ part) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double s = 0.0;
    if (q < n) s = (double)a[q] * a[q];
    __shared__ double red[UTPB];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = UTPB / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// sums partials in a fixed order into out[0]
__global__ void k_reduce_fixed(const double* __restrict__ part, int n, double* __restrict__ out) {
    __shared__ double red[UTPB];
    double s = 0.0;
    for (int q = threadIdx.x; q < n; q += blockDim.x) s += part[q];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = UTPB / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// guess = out * rms(b)
__global__ void k_scale_out(const float* __restrict__ y, const double* __restrict__ bb, int64_t nv, float* __restrict__ p) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nv) return;
    float rms = (float)sqrt(fmax(*bb / (double)nv, 1e-300));
    p[c] = y[c] * rms;
}

__global__ void k_guess_into(const float* __restrict__ y, const double* __restrict__ bb, const uint8_t* __restrict__ mask,
                             int64_t nv, float* __restrict__ p) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nv) return;
    if (mask && mask[c]) return;
    float rms = (float)sqrt(fmax(*bb / (double)nv, 1e-300));
    p[c] = y[c] * rms;
}

// ------------------------------------------------------------------ layers
struct LayerSpec {
    const char* name;
    int cin, cout, relu, level;
};

static constexpr LayerSpec kLayers[] = {
    {"enc0a", 3, 16, 1, 0},  {"enc0b", 16, 16, 1, 0}, {"enc1a", 16, 32, 1, 1}, {"enc1b", 32, 32, 1, 1},
    {"enc2a", 32, 64, 1, 2}, {"enc2b", 64, 64, 1, 2}, {"mid", 64, 64, 1, 3},   {"dec2", 128, 64, 1, 2},
    {"dec1a", 96, 32, 1, 1}, {"dec1b", 32, 32, 1, 1}, {"dec0a", 48, 16, 1, 0}, {"dec0b", 16, 16, 1, 0},
    {"out", 16, 1, 0, 0},
};
static constexpr int kNumLayers = 13;
enum { L_E0A, L_E0B, L_E1A, L_E1B, L_E2A, L_E2B, L_MID, L_D2, L_D1A, L_D1B, L_D0A, L_D0B, L_OUT };

// the BF16 convolution kernels are instantiated for output tiles of 1, 2 or
// 4 column blocks of 16 channels: every layer's padded width must be one
constexpr bool mfma_widths_supported() {
    for (int l = 0; l < kNumLayers; ++l) {
        const int ntn = (kLayers[l].cout + 15) / 16;
        if (ntn != 1 && ntn != 2 && ntn != 4) return false;
    }
    return true;
}
static_assert(mfma_widths_supported(), "a layer's output channels pad to a width the MFMA kernels do not cover");

int64_t unet_param_count() {
    int64_t n = 0;
    for (int l = 0; l < kNumLayers; ++l) n += (int64_t)kLayers[l].cin * kLayers[l].cout * 27 + kLayers[l].cout;
    return n;
}

// ------------------------------------------------ BF16 inference on MFMA
// Implicit GEMM per layer: M = voxels, N = output channels (padded to 16),
// K = 27 taps x input channels (padded to 8), ordered tap-major in groups of
// 8 channels: group G = tap * CB + cb holds channels cb*8 .. cb*8+7 of tap
// (dz*9 + dy*3 + dx).  One v_mfma_f32_16x16x32_bf16 consumes 4 groups.
// Operand layout of that instruction (lane l, r = l % 16, q = l / 16):
//   A[row r][k 8q+j], B[k 8q+j][col r], D[row 4q+i][col r], j < 8, i < 4.
// Activations are channels-last BF16 ([voxel][C], 16-byte channel groups);
// the accumulation is FP32; bias, ReLU and the rounding to BF16 are the
// epilogue.  The last layer (one channel) is stored as FP32.
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

// round to nearest even (activations are finite)
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint32_t h) { return __uint_as_float(h << 16); }

struct MfmaIn {
    const uint16_t* a;  // c1 channels per voxel, half resolution when up != 0
    const uint16_t* b;  // c2 channels per voxel, full resolution
    int c1, c2, up;
};

// weights packed as wq[G][npad][8]: B fragment of group G, column co = 8 bf16
__global__ void k_pack_weights(const float* __restrict__ w, uint16_t* __restrict__ wq, int cout, int cin, int cb_per_tap,
                               int groups, int npad) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t total = (int64_t)groups * npad * 8;
    if (t >= total) return;
    int j = (int)(t % 8);
    int co = (int)((t / 8) % npad);
    int g = (int)(t / (8 * (int64_t)npad));
    int tap = g / cb_per_tap, ci = (g % cb_per_tap) * 8 + j;
    float v = (tap < 27 && co < cout && ci < cin) ? w[((int64_t)co * cin + ci) * 27 + tap] : 0.0f;
    wq[t] = f32_to_bf16(v);
}

// [3][nz][ny][nx] FP32 -> [voxel][8] BF16 (channels 3..7 zero)
__global__ void k_input_to_cl(const float* __restrict__ x, uint16_t* __restrict__ out, int64_t nv) {
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    uint4 o;
    o.x = f32_to_bf16(x[v]) | ((uint32_t)f32_to_bf16(x[nv + v]) << 16);
    o.y = f32_to_bf16(x[2 * nv + v]);
    o.z = 0;
    o.w = 0;
    reinterpret_cast<uint4*>(out)[v] = o;
}

// 2x2x2 average pooling of channels-last BF16, one thread per (coarse voxel,
// group of 8 channels); the mean of 8 BF16 values is exact in FP32
__global__ void k_pool_cl(const uint16_t* __restrict__ in, uint16_t* __restrict__ out, int c, int nz, int ny, int nx) {
    const int hz = nz >> 1, hy = ny >> 1, hx = nx >> 1, cg = c / 8;
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t total = (int64_t)hz * hy * hx * cg;
    if (t >= total) return;
    int g = (int)(t % cg);
    int64_t cv = t / cg;
    int x = (int)(cv % hx), y = (int)((cv / hx) % hy), z = (int)(cv / ((int64_t)hx * hy));
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int d = 0; d < 8; ++d) {
        int64_t fv = ((int64_t)(2 * z + (d >> 2)) * ny + (2 * y + ((d >> 1) & 1))) * nx + (2 * x + (d & 1));
        uint4 r = reinterpret_cast<const uint4*>(in + fv * c)[g];
        uint32_t w4[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            s[2 * k] += bf16_to_f32(w4[k] & 0xFFFFu);
            s[2 * k + 1] += bf16_to_f32(w4[k] >> 16);
        }
    }
    uint4 o;
    uint32_t* ow = &o.x;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        ow[k] = f32_to_bf16(s[2 * k] * 0.125f) | ((uint32_t)f32_to_bf16(s[2 * k + 1] * 0.125f) << 16);
    reinterpret_cast<uint4*>(out + cv * c)[g] = o;
}

// one wavefront computes MT x NT tiles of 16 voxels x 16 channels; the block
// holds 4 independent wavefronts.  The fragments of the next K step are
// loaded before the MFMAs of the current one.
__device__ __forceinline__ uint32_t bit_of(bool b) { return b ? 1u : 0u; }

// Split K (gridDim.y > 1, `part` non-null): split blockIdx.y runs K steps
// [y * steps, (y + 1) * steps) and stores its raw FP32 sums to
// part[(split * nv + v) * cout + co]; k_split_finish adds the splits in a
// fixed order and applies bias, ReLU and the rounding.
template <int MT, int NT>
__global__ void __launch_bounds__(256) k_conv3_mfma(MfmaIn in, const uint16_t* __restrict__ wq,
                                                    const float* __restrict__ bias, void* __restrict__ out, int cout,
                                                    int npad, int groups, int nz, int ny, int nx, int relu,
                                                    int out_f32, float* __restrict__ part, int steps) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int n_groups_n = npad / (16 * NT);
    const int64_t mgrp = wave / n_groups_n;
    const int ngrp = (int)(wave % n_groups_n);
    const int64_t nv = (int64_t)nz * ny * nx;
    const int64_t vbase = mgrp * 16 * MT;
    if (vbase >= nv) return;  // wavefront-uniform
    const int r = lane & 15, q = lane >> 4;
    const int cb_per_tap = (in.c1 + in.c2) >> 3, cb1 = in.c1 >> 3;
    const int hz = nz >> 1, hy = ny >> 1, hx = nx >> 1;
    (void)hz;
    // per tile row: flat voxel index, coordinates and the 27-bit mask of the
    // taps whose neighbour lies inside the grid (0 past the last voxel)
    int vx[MT], vy[MT], vz[MT], vidx[MT];
    uint32_t vmask[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m) {
        // voxel counts stay below 2^31 (checked on the host): 32-bit division
        const uint32_t v = (uint32_t)vbase + 16u * m + r;
        const uint32_t row = v / (uint32_t)nx;
        vidx[m] = (int)v;
        vx[m] = (int)(v - row * (uint32_t)nx);
        vz[m] = (int)(row / (uint32_t)ny);
        vy[m] = (int)(row - (uint32_t)vz[m] * (uint32_t)ny);
        // tap t = dz*9 + dy*3 + dx is valid when each axis' neighbour is:
        // x repeats its 3-bit pattern 9 times, y fills 3-bit groups, z 9-bit ones
        const uint32_t mx = (vx[m] > 0 ? 1u : 0u) | 2u | (vx[m] < nx - 1 ? 4u : 0u);
        const uint32_t rx = mx * 0x1249249u;
        const uint32_t ry = ((bit_of(vy[m] > 0) * 7u) | (7u << 3) | (bit_of(vy[m] < ny - 1) * (7u << 6))) * 0x40201u;
        const uint32_t rz = (bit_of(vz[m] > 0) * 0x1FFu) | (0x1FFu << 9) | (bit_of(vz[m] < nz - 1) * (0x1FFu << 18));
        vmask[m] = (int64_t)v < nv ? (rx & ry & rz) : 0u;
    }
    const int c_begin = (int)blockIdx.y * steps;
    const int c_end = min(groups >> 2, c_begin + steps);
    int tap = (4 * c_begin + q) / cb_per_tap, cb = (4 * c_begin + q) % cb_per_tap;  // this lane's group 4c + q
    const uint16_t* wlane = wq + ((int64_t)q * npad + ngrp * NT * 16 + r) * 8;
    const int64_t wstep = (int64_t)4 * npad * 8;

    auto load_a = [&](int tp, int cbk, bf16x8* a) {
        const int dz = tp / 9, dy = (tp / 3) % 3, dx = tp % 3;
        const bool from_a = cbk < cb1;
        const bool up = in.up && from_a;
        const uint16_t* base = from_a ? in.a : in.b;
        const int64_t cs = from_a ? in.c1 : in.c2;
        const int coff = (from_a ? cbk : cbk - cb1) * 8;
        const int delta = ((dz - 1) * ny + (dy - 1)) * nx + (dx - 1);
        const uint32_t bit = tp < 27 ? 1u << tp : 0u;
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            uint4 raw = make_uint4(0, 0, 0, 0);
            if (vmask[m] & bit) {
                int64_t qv;
                if (up)
                    qv = ((int64_t)((vz[m] + dz - 1) >> 1) * hy + ((vy[m] + dy - 1) >> 1)) * hx + ((vx[m] + dx - 1) >> 1);
                else
                    qv = (int64_t)(vidx[m] + delta);
                raw = *reinterpret_cast<const uint4*>(base + qv * cs + coff);
            }
            a[m] = __builtin_bit_cast(bf16x8, raw);
        }
    };
    auto load_b = [&](const uint16_t* wp, bf16x8* b) {
#pragma unroll
        for (int n = 0; n < NT; ++n) b[n] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(wp + n * 128));
    };

    f32x4 acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int n = 0; n < NT; ++n) acc[m][n] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    bf16x8 a[MT], b[NT], a2[MT], b2[NT];
    load_a(tap, cb, a);
    load_b(wlane + c_begin * wstep, b);
    for (int c = c_begin; c < c_end; ++c) {
        const bool more = c + 1 < c_end;
        if (more) {
            cb += 4;
            while (cb >= cb_per_tap) {
                cb -= cb_per_tap;
                ++tap;
            }
            load_a(tap, cb, a2);
            load_b(wlane + (c + 1) * wstep, b2);
        }
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int n = 0; n < NT; ++n) acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], b[n], acc[m][n], 0, 0, 0);
        if (more) {
#pragma unroll
            for (int m = 0; m < MT; ++m) a[m] = a2[m];
#pragma unroll
            for (int n = 0; n < NT; ++n) b[n] = b2[n];
        }
    }
#pragma unroll
    for (int n = 0; n < NT; ++n) {
        const int co = (ngrp * NT + n) * 16 + r;
        if (co >= cout) continue;
        const float bco = bias[co];
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int64_t v = vbase + 16 * m + 4 * q + i;
                if (v >= nv) continue;
                if (part) {
                    part[((int64_t)blockIdx.y * nv + v) * cout + co] = acc[m][n][i];
                    continue;
                }
                float s = acc[m][n][i] + bco;
                if (relu) s = fmaxf(s, 0.0f);
                if (out_f32) reinterpret_cast<float*>(out)[v * cout + co] = s;
                else reinterpret_cast<uint16_t*>(out)[v * cout + co] = f32_to_bf16(s);
            }
    }
}

// sum of the K splits in split order, then bias, ReLU, rounding
__global__ void k_split_finish(const float* __restrict__ part, int splits, const float* __restrict__ bias,
                               void* __restrict__ out, int64_t nv, int cout, int relu, int out_f32) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = nv * cout;
    if (t >= n) return;
    float acc = part[t];
    for (int k = 1; k < splits; ++k) acc += part[(int64_t)k * n + t];
    float s = acc + bias[t % cout];
    if (relu) s = fmaxf(s, 0.0f);
    if (out_f32) reinterpret_cast<float*>(out)[t] = s;
    else reinterpret_cast<uint16_t*>(out)[t] = f32_to_bf16(s);
}

// Halo-tiled form for levels made of whole 16 x 4 x 4 bricks: the block
// copies its brick's input with a one-voxel halo (18 x 6 x 6 voxels, all
// channels, zero outside the grid; upsampling and concat resolved while
// copying) to LDS once, then each wavefront computes one z-slice of the
// brick (tile row m = y offset, lane row = x offset) with every N tile,
// reading A fragments from LDS.  Same K order and fragments as
// k_conv3_mfma, so the results are bit-identical.
static constexpr int kHaloVox = 18 * 6 * 6;

template <int NT>
__global__ void __launch_bounds__(256) k_conv3_mfma_halo(MfmaIn in, const uint16_t* __restrict__ wq,
                                                         const float* __restrict__ bias, void* __restrict__ out,
                                                         int cout, int npad, int groups, int nz, int ny, int nx,
                                                         int relu, int out_f32) {
    extern __shared__ uint4 halo[];  // [halo voxel][channel group]
    const int cg = (in.c1 + in.c2) >> 3, cg1 = in.c1 >> 3;
    const int bx = nx >> 4, by = ny >> 2;
    const int brick = blockIdx.x;
    const int x0 = (brick % bx) * 16, y0 = ((brick / bx) % by) * 4, z0 = (brick / (bx * by)) * 4;
    const int hz = nz >> 1, hy = ny >> 1, hx = nx >> 1;
    (void)hz;
    const int total = kHaloVox * cg;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
        const int g = t % cg, hv = t / cg;
        const int x = x0 - 1 + hv % 18, y = y0 - 1 + (hv / 18) % 6, z = z0 - 1 + hv / 108;
        uint4 val = make_uint4(0, 0, 0, 0);
        if (x >= 0 && y >= 0 && z >= 0 && x < nx && y < ny && z < nz) {
            if (g < cg1) {
                const int64_t qv = in.up ? (((int64_t)(z >> 1) * hy + (y >> 1)) * hx + (x >> 1))
                                         : (((int64_t)z * ny + y) * nx + x);
                val = *reinterpret_cast<const uint4*>(in.a + qv * in.c1 + g * 8);
            } else {
                val = *reinterpret_cast<const uint4*>(in.b + (((int64_t)z * ny + y) * nx + x) * in.c2 + (g - cg1) * 8);
            }
        }
        halo[t] = val;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r = lane & 15, q = lane >> 4;
    int tap = q / cg, cb = q % cg;
    const uint16_t* wlane = wq + ((int64_t)q * npad + r) * 8;
    const int64_t wstep = (int64_t)4 * npad * 8;
    auto load_a = [&](int tp, int cbk, bf16x8* a) {
        const bool ok = tp < 27;
        const int tt = ok ? tp : 0;
        const int dz = tt / 9, dy = (tt / 3) % 3, dx = tt % 3;
        const int base = ((w + dz) * 6 + dy) * 18 + r + dx;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const uint4 raw = halo[(base + m * 18) * cg + cbk];
            a[m] = __builtin_bit_cast(bf16x8, ok ? raw : make_uint4(0, 0, 0, 0));
        }
    };
    auto load_b = [&](const uint16_t* wp, bf16x8* b) {
#pragma unroll
        for (int n = 0; n < NT; ++n) b[n] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(wp + n * 128));
    };
    f32x4 acc[4][NT];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int n = 0; n < NT; ++n) acc[m][n] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    bf16x8 a[4], b[NT], a2[4], b2[NT];
    load_a(tap, cb, a);
    load_b(wlane, b);
    const int nchunks = groups >> 2;
    for (int c = 0; c < nchunks; ++c) {
        const bool more = c + 1 < nchunks;
        if (more) {
            cb += 4;
            while (cb >= cg) {
                cb -= cg;
                ++tap;
            }
            load_a(tap, cb, a2);
            load_b(wlane + (c + 1) * wstep, b2);
        }
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int n = 0; n < NT; ++n) acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], b[n], acc[m][n], 0, 0, 0);
        if (more) {
#pragma unroll
            for (int m = 0; m < 4; ++m) a[m] = a2[m];
#pragma unroll
            for (int n = 0; n < NT; ++n) b[n] = b2[n];
        }
    }
#pragma unroll
    for (int n = 0; n < NT; ++n) {
        const int co = n * 16 + r;
        if (co >= cout) continue;
        const float bco = bias[co];
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int64_t v = ((int64_t)(z0 + w) * ny + (y0 + m)) * nx + (x0 + 4 * q + i);
                float s2 = acc[m][n][i] + bco;
                if (relu) s2 = fmaxf(s2, 0.0f);
                if (out_f32) reinterpret_cast<float*>(out)[v * cout + co] = s2;
                else reinterpret_cast<uint16_t*>(out)[v * cout + co] = f32_to_bf16(s2);
            }
    }
}

}  // namespace flux

using namespace flux;

// Energy-norm loss: the guess g = y on the free cells
__global__ void k_mask_out(const float* __restrict__ y, const uint8_t* __restrict__ mask, float* __restrict__ g,
                           int64_t n) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) g[q] = (mask && mask[q]) ? 0.0f : y[q];
}

// partials of a.b (part[0..gridDim)) and c.d (part[gridDim..2 gridDim))
__global__ void k_dot2(const float* __restrict__ a, const float* __restrict__ b, const float* __restrict__ c,
                       const float* __restrict__ d, int64_t n, double* __restrict__ part) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double s0 = 0.0, s1 = 0.0;
    if (q < n) {
        s0 = (double)a[q] * b[q];
        s1 = (double)c[q] * d[q];
    }
    __shared__ double red[2][UTPB];
    red[0][threadIdx.x] = s0;
    red[1][threadIdx.x] = s1;
    __syncthreads();
    for (int st = UTPB / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            red[0][threadIdx.x] += red[0][threadIdx.x + st];
            red[1][threadIdx.x] += red[1][threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = red[0][0];
        part[gridDim.x + blockIdx.x] = red[1][0];
    }
}

// dL/dy = c1 r + c2 (A g) on the free cells, 0 on the Dirichlet cells
__global__ void k_energy_grad(const float* __restrict__ r, const float* __restrict__ ag,
                              const uint8_t* __restrict__ mask, float c1, float c2, float* __restrict__ gy,
                              int64_t n) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) gy[q] = (mask && mask[q]) ? 0.0f : c1 * r[q] + c2 * ag[q];
}

// ------------------------------------------------ BF16 training on MFMA
// The backward pass of a BF16 network (flux_unet_set_precision): the
// output gradient of a layer, masked by its ReLU, goes to channels-last
// BF16 once (k_grad_to_cl); the data gradient is the implicit-GEMM
// convolution of that tensor with the layer's flipped, transposed kernel on
// the matrix cores (k_conv3_mfma with its own packed weights), and the
// weight gradient an MFMA GEMM over the voxels (k_wgrad_mfma).  Master
// weights, their gradients and Adam stay FP32.

// gy (FP32 [C][voxel]) masked by y > 0 (y channels-last BF16, or none) ->
// g16 (channels-last BF16, cpad >= c channels, the rest 0)
__global__ void k_grad_to_cl(const float* __restrict__ gy, const uint16_t* __restrict__ y16, int c, int cpad,
                             int64_t nv, uint16_t* __restrict__ g16) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nv * cpad) return;
    const int ch = (int)(t % cpad);
    const int64_t v = t / cpad;
    float g = 0.0f;
    if (ch < c) {
        g = gy[(int64_t)ch * nv + v];
        // y > 0: a BF16 value with the sign bit clear and not zero
        if (y16) {
            const uint32_t h = y16[v * c + ch];
            if ((h & 0x8000u) || (h & 0x7FFFu) == 0u) g = 0.0f;
        }
    }
    g16[t] = f32_to_bf16(g);
}

// channels-last FP32 [voxel][c] -> [c][voxel]
__global__ void k_cl_to_nc(const float* __restrict__ src, float* __restrict__ dst, int c, int64_t nv) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nv * c) return;
    const int64_t v = t / c;
    const int ch = (int)(t % c);
    dst[(int64_t)ch * nv + v] = src[t];
}

// Weight gradient dW[co][ci][tap] = sum_v g[v][co] x[v + off(tap)][ci] as
// MFMA GEMMs: M = output channels (rows co0 .. co0 + 15), N = input
// channels (columns ci0 .. ci0 + 15), K = voxels in steps of 32.  A
// wavefront owns one tap, one (co, ci) tile and a run of `kv` voxels (whole
// K steps).  Lane l (r = l % 16, q = l / 16) holds A[r][8q + j] =
// g[v0 + 8q + j][co0 + r] and B[8q + j][r] = x[v0 + 8q + j + off][ci0 + r],
// each of its 8 voxels with its own coordinates (any grid shape); zero
// padding outside the grid, upsampling and concat as in k_conv3_mfma.  Each
// wavefront stores its FP32 partial sums; k_wgrad_finish adds the runs in a
// fixed order.
__global__ void __launch_bounds__(256) k_wgrad_mfma(const uint16_t* __restrict__ g, int cpad, MfmaIn in, int co_tiles,
                                                    int ci_tiles, int64_t kv, int runs, int nz, int ny, int nx,
                                                    float* __restrict__ part) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int64_t tiles = (int64_t)co_tiles * ci_tiles;
    if (wave >= tiles * 27 * runs) return;  // wavefront-uniform
    const int tile = (int)(wave % tiles);
    const int tp = (int)((wave / tiles) % 27);
    const int run = (int)(wave / (tiles * 27));
    const int co0 = (tile / ci_tiles) * 16, ci0 = (tile % ci_tiles) * 16;
    const int r = lane & 15, q = lane >> 4;
    const int64_t nv = (int64_t)nz * ny * nx;
    const int64_t vbeg = (int64_t)run * kv, vend = min(nv, vbeg + kv);
    const int cin = in.c1 + in.c2;
    const int hy = ny >> 1, hx = nx >> 1;
    const int dz = tp / 9, dy = (tp / 3) % 3, dx = tp % 3;
    const int co = co0 + r, ci = ci0 + r;
    const bool co_ok = co < cpad, ci_ok = ci < cin;
    const bool from_a = ci < in.c1;
    const bool up = in.up && from_a;
    const uint16_t* base = from_a ? in.a : in.b;
    const int64_t cs = from_a ? in.c1 : in.c2;
    const int coff = from_a ? ci : ci - in.c1;
    f32x4 acc = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int64_t v0 = vbeg; v0 < vend; v0 += 32) {
        // this lane's 8 voxels vb .. vb + 7, walked with their coordinates
        const int64_t vb = v0 + 8 * q;
        const int64_t row = vb / nx;
        int xi = (int)(vb - row * nx), yi = (int)(row % ny), zi = (int)(row / ny);
        uint16_t av[8], bv[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const bool vin = vb + j < vend;
            av[j] = (vin && co_ok) ? g[(vb + j) * cpad + co] : (uint16_t)0;
            const int xx = xi + dx - 1, yy = yi + dy - 1, zz = zi + dz - 1;
            uint16_t val = 0;
            if (vin && ci_ok && xx >= 0 && xx < nx && yy >= 0 && yy < ny && zz >= 0 && zz < nz) {
                const int64_t qv = up ? ((int64_t)(zz >> 1) * hy + (yy >> 1)) * hx + (xx >> 1)
                                      : ((int64_t)zz * ny + yy) * nx + xx;
                val = base[qv * cs + coff];
            }
            bv[j] = val;
            if (++xi == nx) {
                xi = 0;
                if (++yi == ny) {
                    yi = 0;
                    ++zi;
                }
            }
        }
        bf16x8 a, b;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            a[j] = __builtin_bit_cast(__bf16, av[j]);
            b[j] = __builtin_bit_cast(__bf16, bv[j]);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    // D[4q + i][r]: row co0 + 4q + i, column ci0 + r
    const int np = ci_tiles * 16, mp = co_tiles * 16;
#pragma unroll
    for (int i = 0; i < 4; ++i) part[(((int64_t)run * 27 + tp) * mp + co0 + 4 * q + i) * np + ci0 + r] = acc[i];
}

// dW[co][ci][tap] (co < cout, ci < cin) = the runs' partial sums added in run
// order; db[co] = sum_v g[v][co] in a fixed order (one block per channel)
__global__ void k_wgrad_finish(const float* __restrict__ part, int runs, int co_tiles, int ci_tiles, int cout,
                               int cin, float* __restrict__ dw) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)cout * cin * 27) return;
    const int tp = (int)(t % 27);
    const int ci = (int)((t / 27) % cin);
    const int co = (int)(t / (27 * (int64_t)cin));
    const int np = ci_tiles * 16, mp = co_tiles * 16;
    float s = 0.0f;
    for (int k = 0; k < runs; ++k) s += part[(((int64_t)k * 27 + tp) * mp + co) * np + ci];
    dw[t] = s;
}

__global__ void __launch_bounds__(UTPB) k_bias_grad(const uint16_t* __restrict__ g, int cpad, int64_t nv,
                                                    float* __restrict__ db) {
    const int co = blockIdx.x;
    float s = 0.0f;
    for (int64_t v = threadIdx.x; v < nv; v += blockDim.x) s += bf16_to_f32(g[v * cpad + co]);
    __shared__ float red[UTPB];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = UTPB / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[co] = red[0];
}

struct flux_unet {
    int n[3];  // nx, ny, nz of the finest level
    int64_t nparams = 0;
    float* params = nullptr;  // [layer: w (co*ci*27), b (co)] ...
    float* grads = nullptr;
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    float* wt = nullptr;      // flipped/transposed weights scratch (largest layer)
    int64_t off_w[kNumLayers], off_b[kNumLayers];
    float* y[kNumLayers] = {};    // layer outputs
    float* pooled[3] = {};        // pool(s0), pool(s1), pool(s2)
    float* gy[kNumLayers] = {};   // gradient w.r.t. each layer output
    float* gin_up[kNumLayers] = {};  // gradient of the upsampled part (full res)
    float* gpool[3] = {};         // gradients of pooled tensors
    double* part = nullptr;
    double* scal = nullptr;       // [0] loss, [1] b.b
    int64_t nvox[4];
    int steps = 0;
    // BF16 inference (flux_unet_set_precision)
    int precision = FLUX_UNET_F32;
    uint16_t* wq = nullptr;        // packed weights of every layer
    int64_t off_wq[kNumLayers];
    int groups[kNumLayers], npad[kNumLayers];
    uint16_t* y16[kNumLayers] = {};  // channels-last layer outputs (the last layer writes y[L_OUT])
    uint16_t* pooled16[3] = {};
    uint16_t* x16 = nullptr;       // [voxel][8]
    int64_t version = 1, packed_version = 0;
    int halo = 1;  // halo-tiled convolutions: 0 never, 1 where measured faster, 2 wherever they fit
    int split = 1;                 // split K on small levels (FLUX_UNET_SPLIT=0 disables)
    int tile_mt = 0, tile_nt = 0;  // forced flat-kernel tile (FLUX_UNET_TILE=MxN), 0 = chosen per layer
    float* split_part = nullptr;   // FP32 partial sums of the K splits, sized for the layers that split
    // BF16 training (the backward pass on MFMA)
    uint16_t* wqt = nullptr;       // packed flipped, transposed weights of every layer (data gradients)
    int64_t off_wqt[kNumLayers];
    int groups_t[kNumLayers], npad_t[kNumLayers];
    int64_t packed_t_version = 0;
    uint16_t* g16 = nullptr;       // a layer's masked output gradient, channels-last BF16
    float* dx_cl = nullptr;        // a layer's data gradient, channels-last FP32
    float* wpart = nullptr;        // weight-gradient partial sums
    int64_t wpart_floats = 0;
    float* zero_bias = nullptr;    // 128 zeros (the data gradient has no bias)
};

namespace {

int64_t lvox(const flux_unet* u, int level) {
    return (int64_t)(u->n[0] >> level) * (u->n[1] >> level) * (u->n[2] >> level);
}

template <int C>
void launch_conv(const ConvIn& in, const float* w, const float* b, float* out, int cout, int nz, int ny, int nx,
                 int relu) {
    int64_t nv = (int64_t)nz * ny * nx;
    k_conv3<C><<<blocks_for(nv, UTPB), UTPB, 0, g_stream>>>(in, w, b, out, cout, nz, ny, nx, relu);
}

// output channels beyond 64 (the data gradient of dec2 has 128) run as
// further launches over channel groups of 64
void conv(const ConvIn& in, const float* w, const float* b, float* out, int cout, int level, const flux_unet* u,
          int relu) {
    int nx = u->n[0] >> level, ny = u->n[1] >> level, nz = u->n[2] >> level;
    const int cin = in.c1 + in.c2;
    const int64_t nv = (int64_t)nz * ny * nx;
    for (int co0 = 0; co0 < cout; co0 += 64) {
        int cc = std::min(64, cout - co0);
        const float* wg = w + (size_t)co0 * cin * 27;
        const float* bg = b ? b + co0 : nullptr;
        float* og = out + (int64_t)co0 * nv;
        if (cc <= 1) launch_conv<1>(in, wg, bg, og, cc, nz, ny, nx, relu);
        else if (cc <= 16) launch_conv<16>(in, wg, bg, og, cc, nz, ny, nx, relu);
        else if (cc <= 32) launch_conv<32>(in, wg, bg, og, cc, nz, ny, nx, relu);
        else launch_conv<64>(in, wg, bg, og, cc, nz, ny, nx, relu);
    }
}

ConvIn cin_single(const float* a, int c) { return ConvIn{a, nullptr, c, 0, 0, nullptr}; }
ConvIn cin_cat(const float* up_src, int c1, const float* skip, int c2) { return ConvIn{up_src, skip, c1, c2, 1, nullptr}; }

// the input of layer l in the forward pass
ConvIn layer_input(const flux_unet* u, int l, const float* x) {
    switch (l) {
        case L_E0A: return cin_single(x, 3);
        case L_E0B: return cin_single(u->y[L_E0A], 16);
        case L_E1A: return cin_single(u->pooled[0], 16);
        case L_E1B: return cin_single(u->y[L_E1A], 32);
        case L_E2A: return cin_single(u->pooled[1], 32);
        case L_E2B: return cin_single(u->y[L_E2A], 64);
        case L_MID: return cin_single(u->pooled[2], 64);
        case L_D2: return cin_cat(u->y[L_MID], 64, u->y[L_E2B], 64);
        case L_D1A: return cin_cat(u->y[L_D2], 64, u->y[L_E1B], 32);
        case L_D1B: return cin_single(u->y[L_D1A], 32);
        case L_D0A: return cin_cat(u->y[L_D1B], 32, u->y[L_E0B], 16);
        case L_D0B: return cin_single(u->y[L_D0A], 16);
        default: return cin_single(u->y[L_D0B], 16);
    }
}

void pool_into(const float* in, float* out, int c, int level, const flux_unet* u) {
    int nx = u->n[0] >> level, ny = u->n[1] >> level, nz = u->n[2] >> level;
    int64_t n = (int64_t)c * (nz >> 1) * (ny >> 1) * (nx >> 1);
    k_pool<<<blocks_for(n, UTPB), UTPB, 0, g_stream>>>(in, out, c, nz, ny, nx);
}

int forward(flux_unet* u, const float* x) {
    for (int l = 0; l < kNumLayers; ++l) {
        const LayerSpec& s = kLayers[l];
        ConvIn in = layer_input(u, l, x);
        conv(in, u->params + u->off_w[l], u->params + u->off_b[l], u->y[l], s.cout, s.level, u, s.relu);
        if (l == L_E0B) pool_into(u->y[L_E0B], u->pooled[0], 16, 0, u);
        if (l == L_E1B) pool_into(u->y[L_E1B], u->pooled[1], 32, 1, u);
        if (l == L_E2B) pool_into(u->y[L_E2B], u->pooled[2], 64, 2, u);
    }
    FLUX_LAUNCH_CHECK();
    return FLUX_OK;
}

// ---- BF16 inference path
static constexpr int64_t kSplitPartBytes = (int64_t)32 << 20;
MfmaIn mfma_input(const flux_unet* u, int l) {
    switch (l) {
        case L_E0A: return MfmaIn{u->x16, nullptr, 8, 0, 0};
        case L_E0B: return MfmaIn{u->y16[L_E0A], nullptr, 16, 0, 0};
        case L_E1A: return MfmaIn{u->pooled16[0], nullptr, 16, 0, 0};
        case L_E1B: return MfmaIn{u->y16[L_E1A], nullptr, 32, 0, 0};
        case L_E2A: return MfmaIn{u->pooled16[1], nullptr, 32, 0, 0};
        case L_E2B: return MfmaIn{u->y16[L_E2A], nullptr, 64, 0, 0};
        case L_MID: return MfmaIn{u->pooled16[2], nullptr, 64, 0, 0};
        case L_D2: return MfmaIn{u->y16[L_MID], u->y16[L_E2B], 64, 64, 1};
        case L_D1A: return MfmaIn{u->y16[L_D2], u->y16[L_E1B], 64, 32, 1};
        case L_D1B: return MfmaIn{u->y16[L_D1A], nullptr, 32, 0, 0};
        case L_D0A: return MfmaIn{u->y16[L_D1B], u->y16[L_E0B], 32, 16, 1};
        case L_D0B: return MfmaIn{u->y16[L_D0A], nullptr, 16, 0, 0};
        default: return MfmaIn{u->y16[L_D0B], nullptr, 16, 0, 0};
    }
}

int mfma_channels(int l) { return l == L_E0A ? 8 : kLayers[l].cin; }

// How one layer's BF16 convolution runs: the halo-tiled kernel, or the flat
// kernel with an MT x NT tile (rows of 16 voxels x columns of 16 channels)
// and K split `splits` ways.
struct MfmaPlan {
    bool halo;
    int mt, nt, splits;
    int64_t nv, halo_bytes, bricks;
};

// the widest tile (4 x N/16) that still gives >= 4096 wavefronts; smaller
// levels trade tile reuse for parallelism
MfmaPlan plan_mfma(const flux_unet* u, int l) {
    const LayerSpec& s = kLayers[l];
    const int nx = u->n[0] >> s.level, ny = u->n[1] >> s.level, nz = u->n[2] >> s.level;
    MfmaPlan p{};
    p.nv = (int64_t)nz * ny * nx;
    const int ntn = u->npad[l] / 16;
    p.halo_bytes = (int64_t)kHaloVox * (mfma_channels(l) / 8) * 16;
    p.bricks = (int64_t)(nx / 16) * (ny / 4) * (nz / 4);
    // measured: the halo form pays where the brick is small (<= 32 channels,
    // <= 41 KiB of LDS, several blocks per CU) and the level has >= 1024 bricks
    const bool fits = nx % 16 == 0 && ny % 4 == 0 && nz % 4 == 0 && p.halo_bytes <= 64 * 1024;
    const bool pays = mfma_channels(l) <= 32 && p.bricks >= 1024;
    p.halo = u->tile_mt == 0 && fits && (u->halo == 2 || (u->halo == 1 && pays));
    p.mt = 4;
    p.nt = ntn;
    auto waves = [&]() { return (p.nv + 16 * p.mt - 1) / (16 * p.mt) * (ntn / p.nt); };
    if (u->tile_mt) {
        p.mt = u->tile_mt;
        p.nt = std::min(u->tile_nt, ntn);
    } else {
        while (waves() < 4096 && (p.mt > 1 || p.nt > 1)) {
            if (p.mt > 1) p.mt >>= 1;
            else p.nt >>= 1;
        }
    }
    // levels too small to fill the GPU (< 4096 wavefronts even with 1x1
    // tiles): split K until 4096 run, each split keeping >= 8 K steps, while
    // the FP32 partial sums stay within 32 MiB (measured: splitting levels
    // that already have 4096 wavefronts only adds the finishing pass)
    p.splits = 1;
    const int chunks = u->groups[l] / 4;
    while (!p.halo && u->split && waves() * p.splits < 4096 && p.splits < 8 && chunks / (2 * p.splits) >= 8 &&
           (int64_t)2 * p.splits * p.nv * s.cout * 4 <= kSplitPartBytes)
        p.splits *= 2;
    // every split must own at least one K step
    while (p.splits > 1 && (int64_t)(p.splits - 1) * ((chunks + p.splits - 1) / p.splits) >= chunks) p.splits >>= 1;
    return p;
}

// allocate the packed weights and the channels-last activations
int bf16_setup(flux_unet* u) {
    if (u->wq) return FLUX_OK;
    // the MFMA loader keeps flat voxel indices in 32 bits
    FLUX_REQUIRE(u->nvox[0] < ((int64_t)1 << 31), "unet: BF16 inference needs fewer than 2^31 voxels");
    int64_t off = 0;
    for (int l = 0; l < kNumLayers; ++l) {
        const int cbt = mfma_channels(l) / 8;
        u->groups[l] = (27 * cbt + 3) / 4 * 4;
        u->npad[l] = (kLayers[l].cout + 15) / 16 * 16;
        u->off_wq[l] = off;
        off += (int64_t)u->groups[l] * u->npad[l] * 8;
    }
    bool ok = true;
    auto alloc = [&](uint16_t*& p, int64_t n) {
        p = (uint16_t*)dmalloc(n * 2);
        ok = ok && p;
    };
    alloc(u->wq, off);
    int64_t split_bytes = 0;
    for (int l = 0; l < kNumLayers; ++l) {
        const MfmaPlan p = plan_mfma(u, l);
        if (p.splits > 1) split_bytes = std::max<int64_t>(split_bytes, (int64_t)p.splits * p.nv * kLayers[l].cout * 4);
    }
    if (split_bytes > 0) {
        u->split_part = (float*)dmalloc((size_t)split_bytes);
        ok = ok && u->split_part;
    }
    alloc(u->x16, 8 * u->nvox[0]);
    for (int l = 0; l < L_OUT; ++l) alloc(u->y16[l], (int64_t)kLayers[l].cout * u->nvox[kLayers[l].level]);
    alloc(u->pooled16[0], 16 * u->nvox[1]);
    alloc(u->pooled16[1], 32 * u->nvox[2]);
    alloc(u->pooled16[2], 64 * u->nvox[3]);
    if (!ok) return set_error(FLUX_ERR_NOMEM, "unet: BF16 buffers: allocation failed");
    u->packed_version = 0;
    return FLUX_OK;
}

template <int MT, int NT>
void launch_mfma(const MfmaIn& in, const uint16_t* wq, const float* b, void* out, int cout, int npad, int groups,
                 int nz, int ny, int nx, int relu, int out_f32, float* part, int splits) {
    const int64_t nv = (int64_t)nz * ny * nx;
    const int64_t waves = (nv + 16 * MT - 1) / (16 * MT) * (npad / (16 * NT));
    const int chunks = groups / 4, steps = (chunks + splits - 1) / splits;
    dim3 grid((unsigned)((waves + 3) / 4), (unsigned)splits);
    k_conv3_mfma<MT, NT><<<grid, 256, 0, g_stream>>>(in, wq, b, out, cout, npad, groups, nz, ny, nx, relu, out_f32,
                                                    splits > 1 ? part : nullptr, steps);
    if (splits > 1)
        k_split_finish<<<blocks_for(nv * cout, UTPB), UTPB, 0, g_stream>>>(part, splits, b, out, nv, cout, relu,
                                                                           out_f32);
}

int conv_mfma(flux_unet* u, int l, void* out, int out_f32) {
    const LayerSpec& s = kLayers[l];
    const int nx = u->n[0] >> s.level, ny = u->n[1] >> s.level, nz = u->n[2] >> s.level;
    const MfmaPlan p = plan_mfma(u, l);
    const int ntn = u->npad[l] / 16;
    MfmaIn in = mfma_input(u, l);
    const uint16_t* wq = u->wq + u->off_wq[l];
    const float* b = u->params + u->off_b[l];
    const int cp = s.cout, np = u->npad[l], g = u->groups[l];
    if (p.halo) {
        const dim3 grid((unsigned)p.bricks);
        const size_t lds = (size_t)p.halo_bytes;
        if (ntn == 1)
            k_conv3_mfma_halo<1><<<grid, 256, lds, g_stream>>>(in, wq, b, out, cp, np, g, nz, ny, nx, s.relu, out_f32);
        else if (ntn == 2)
            k_conv3_mfma_halo<2><<<grid, 256, lds, g_stream>>>(in, wq, b, out, cp, np, g, nz, ny, nx, s.relu, out_f32);
        else if (ntn == 4)
            k_conv3_mfma_halo<4><<<grid, 256, lds, g_stream>>>(in, wq, b, out, cp, np, g, nz, ny, nx, s.relu, out_f32);
        else
            return set_error(FLUX_ERR_STATE, "unet: layer %s: no halo kernel for %d output channels", s.name, np);
        return FLUX_OK;
    }
    FLUX_REQUIRE(p.splits == 1 || u->split_part, "unet: layer %s splits K without a partial-sum buffer", s.name);
#define FLUX_MFMA_CASE(M, N)                                                                                    \
    if (p.mt == M && p.nt == N) {                                                                               \
        launch_mfma<M, N>(in, wq, b, out, cp, np, g, nz, ny, nx, s.relu, out_f32, u->split_part, p.splits);     \
        return FLUX_OK;                                                                                         \
    }
    FLUX_MFMA_CASE(4, 4) FLUX_MFMA_CASE(4, 2) FLUX_MFMA_CASE(4, 1)
    FLUX_MFMA_CASE(2, 4) FLUX_MFMA_CASE(2, 2) FLUX_MFMA_CASE(2, 1)
    FLUX_MFMA_CASE(1, 4) FLUX_MFMA_CASE(1, 2) FLUX_MFMA_CASE(1, 1)
#undef FLUX_MFMA_CASE
    return set_error(FLUX_ERR_STATE, "unet: layer %s: no MFMA kernel for a %dx%d tile over %d channels", s.name, p.mt,
                     p.nt, np);
}

int forward_bf16(flux_unet* u, const float* x) {
    FLUX_TRY(bf16_setup(u));
    if (u->packed_version != u->version) {
        for (int l = 0; l < kNumLayers; ++l) {
            const int64_t total = (int64_t)u->groups[l] * u->npad[l] * 8;
            k_pack_weights<<<blocks_for(total, UTPB), UTPB, 0, g_stream>>>(
                u->params + u->off_w[l], u->wq + u->off_wq[l], kLayers[l].cout, kLayers[l].cin,
                mfma_channels(l) / 8, u->groups[l], u->npad[l]);
        }
        u->packed_version = u->version;
    }
    k_input_to_cl<<<blocks_for(u->nvox[0], UTPB), UTPB, 0, g_stream>>>(x, u->x16, u->nvox[0]);
    auto pool = [&](int src, int dst, int c, int level) {
        const int nx = u->n[0] >> level, ny = u->n[1] >> level, nz = u->n[2] >> level;
        const int64_t total = (int64_t)(nz >> 1) * (ny >> 1) * (nx >> 1) * (c / 8);
        k_pool_cl<<<blocks_for(total, UTPB), UTPB, 0, g_stream>>>(u->y16[src], u->pooled16[dst], c, nz, ny, nx);
    };
    for (int l = 0; l < kNumLayers; ++l) {
        if (l == L_OUT) FLUX_TRY(conv_mfma(u, l, u->y[L_OUT], 1));
        else FLUX_TRY(conv_mfma(u, l, u->y16[l], 0));
        if (l == L_E0B) pool(L_E0B, 0, 16, 0);
        if (l == L_E1B) pool(L_E1B, 1, 32, 1);
        if (l == L_E2B) pool(L_E2B, 2, 64, 2);
    }
    FLUX_LAUNCH_CHECK();
    return FLUX_OK;
}

int infer(flux_unet* u, const float* x) { return u->precision == FLUX_UNET_BF16 ? forward_bf16(u, x) : forward(u, x); }

// ---- BF16 training: the backward pass on MFMA
int cpad_of(int l) { return (kLayers[l].cout + 7) / 8 * 8; }

// How a layer's weight gradient splits the voxels into runs: about 8192
// wavefronts (27 per tile and run), whole K steps per run, partial sums
// within a 64 MiB budget
struct WgradPlan {
    int co_tiles, ci_tiles, runs;
    int64_t kv, nv;
};
WgradPlan plan_wgrad(const flux_unet* u, int l) {
    const LayerSpec& s = kLayers[l];
    WgradPlan p;
    p.nv = u->nvox[s.level];
    p.co_tiles = (s.cout + 15) / 16;
    p.ci_tiles = (mfma_channels(l) + 15) / 16;
    const int64_t tiles = (int64_t)p.co_tiles * p.ci_tiles;
    const int64_t per_run = (int64_t)27 * p.co_tiles * 16 * p.ci_tiles * 16;
    const int64_t steps = (p.nv + 31) / 32;
    int64_t runs = std::max<int64_t>(1, 8192 / (27 * tiles));
    runs = std::min<int64_t>(runs, std::max<int64_t>(1, ((int64_t)16 << 20) / per_run));
    runs = std::min<int64_t>(runs, steps);
    p.kv = (steps + runs - 1) / runs * 32;
    p.runs = (int)((p.nv + p.kv - 1) / p.kv);
    return p;
}

int bf16_train_setup(flux_unet* u) {
    FLUX_TRY(bf16_setup(u));
    if (u->wqt) return FLUX_OK;
    FLUX_REQUIRE(u->grads, "unet: BF16 training needs a training network");
    int64_t off = 0, g16_n = 0, dx_n = 0, wp_n = 0;
    for (int l = 0; l < kNumLayers; ++l) {
        const LayerSpec& s = kLayers[l];
        const int64_t nv = u->nvox[s.level];
        u->groups_t[l] = (27 * (cpad_of(l) / 8) + 3) / 4 * 4;
        u->npad_t[l] = (s.cin + 15) / 16 * 16;
        u->off_wqt[l] = off;
        off += (int64_t)u->groups_t[l] * u->npad_t[l] * 8;
        g16_n = std::max<int64_t>(g16_n, (int64_t)cpad_of(l) * nv);
        if (l != L_E0A) dx_n = std::max<int64_t>(dx_n, (int64_t)s.cin * nv);
        const WgradPlan p = plan_wgrad(u, l);
        wp_n = std::max<int64_t>(wp_n, (int64_t)p.runs * 27 * p.co_tiles * 16 * p.ci_tiles * 16);
    }
    u->wqt = (uint16_t*)dmalloc(off * 2);
    u->g16 = (uint16_t*)dmalloc(g16_n * 2);
    u->dx_cl = (float*)dmalloc(dx_n * 4);
    u->wpart = (float*)dmalloc(wp_n * 4);
    u->wpart_floats = wp_n;
    u->zero_bias = (float*)dmalloc(128 * 4);
    if (!u->wqt || !u->g16 || !u->dx_cl || !u->wpart || !u->zero_bias)
        return set_error(FLUX_ERR_NOMEM, "unet: BF16 training buffers: allocation failed");
    FLUX_HIP(hipMemsetAsync(u->zero_bias, 0, 128 * 4, g_stream));
    u->packed_t_version = 0;
    return FLUX_OK;
}

// the flipped, transposed kernels of every layer packed for the data
// gradients (after each change of the master weights)
void pack_transposed(flux_unet* u) {
    if (u->packed_t_version == u->version) return;
    for (int l = 0; l < kNumLayers; ++l) {
        if (l == L_E0A) continue;
        const LayerSpec& s = kLayers[l];
        const int cnt = s.cout * s.cin * 27;
        k_flip_transpose<<<blocks_for(cnt, UTPB), UTPB, 0, g_stream>>>(u->params + u->off_w[l], u->wt, s.cout, s.cin);
        const int64_t total = (int64_t)u->groups_t[l] * u->npad_t[l] * 8;
        k_pack_weights<<<blocks_for(total, UTPB), UTPB, 0, g_stream>>>(u->wt, u->wqt + u->off_wqt[l], s.cin, s.cout,
                                                                       cpad_of(l) / 8, u->groups_t[l], u->npad_t[l]);
    }
    u->packed_t_version = u->version;
}

// layer l's masked output gradient into g16 (channels-last BF16)
void grad_to_cl(flux_unet* u, int l) {
    const LayerSpec& s = kLayers[l];
    const int64_t nv = u->nvox[s.level];
    const int cp = cpad_of(l);
    const uint16_t* y = (s.relu && l != L_OUT) ? u->y16[l] : nullptr;
    k_grad_to_cl<<<blocks_for(nv * cp, UTPB), UTPB, 0, g_stream>>>(u->gy[l], y, s.cout, cp, nv, u->g16);
}

void weight_grad_bf16(flux_unet* u, int l) {
    const LayerSpec& s = kLayers[l];
    const int nx = u->n[0] >> s.level, ny = u->n[1] >> s.level, nz = u->n[2] >> s.level;
    const WgradPlan p = plan_wgrad(u, l);
    const int64_t waves = (int64_t)p.co_tiles * p.ci_tiles * 27 * p.runs;
    k_wgrad_mfma<<<(unsigned)((waves + 3) / 4), 256, 0, g_stream>>>(u->g16, cpad_of(l), mfma_input(u, l), p.co_tiles,
                                                                     p.ci_tiles, p.kv, p.runs, nz, ny, nx, u->wpart);
    const int64_t nw = (int64_t)s.cout * s.cin * 27;
    k_wgrad_finish<<<blocks_for(nw, UTPB), UTPB, 0, g_stream>>>(u->wpart, p.runs, p.co_tiles, p.ci_tiles, s.cout,
                                                                  s.cin, u->grads + u->off_w[l]);
    k_bias_grad<<<s.cout, UTPB, 0, g_stream>>>(u->g16, cpad_of(l), p.nv, u->grads + u->off_b[l]);
}

// the data gradient of layer l into dst ([cin][voxel] FP32, [up | skip])
void data_grad_bf16(flux_unet* u, int l, float* dst) {
    const LayerSpec& s = kLayers[l];
    const int nx = u->n[0] >> s.level, ny = u->n[1] >> s.level, nz = u->n[2] >> s.level;
    const int64_t nv = u->nvox[s.level];
    const MfmaIn in{u->g16, nullptr, cpad_of(l), 0, 0};
    const uint16_t* wq = u->wqt + u->off_wqt[l];
    const int np = u->npad_t[l], g = u->groups_t[l];
    if (np % 64 == 0)
        launch_mfma<1, 4>(in, wq, u->zero_bias, u->dx_cl, s.cin, np, g, nz, ny, nx, 0, 1, nullptr, 1);
    else if (np % 32 == 0)
        launch_mfma<1, 2>(in, wq, u->zero_bias, u->dx_cl, s.cin, np, g, nz, ny, nx, 0, 1, nullptr, 1);
    else
        launch_mfma<1, 1>(in, wq, u->zero_bias, u->dx_cl, s.cin, np, g, nz, ny, nx, 0, 1, nullptr, 1);
    k_cl_to_nc<<<blocks_for(nv * s.cin, UTPB), UTPB, 0, g_stream>>>(u->dx_cl, dst, s.cin, nv);
}

// gradient of layer l's input from gy[l]: returns into dst (full-res, cin
// channels laid out [up part | skip part]) -- computed as a conv of the masked
// gradient with the flipped transposed kernel
void data_grad(flux_unet* u, int l, float* dst) {
    const LayerSpec& s = kLayers[l];
    int cnt = s.cout * s.cin * 27;
    k_flip_transpose<<<blocks_for(cnt, UTPB), UTPB, 0, g_stream>>>(u->params + u->off_w[l], u->wt, s.cout, s.cin);
    ConvIn gin{u->gy[l], nullptr, s.cout, 0, 0, s.relu ? u->y[l] : nullptr};
    conv(gin, u->wt, nullptr, dst, s.cin, s.level, u, 0);
}

void weight_grad(flux_unet* u, int l, const float* x) {
    const LayerSpec& s = kLayers[l];
    int nx = u->n[0] >> s.level, ny = u->n[1] >> s.level, nz = u->n[2] >> s.level;
    ConvIn in = layer_input(u, l, x);
    k_conv3_wgrad<<<s.cout * s.cin, UTPB, 0, g_stream>>>(in, u->gy[l], s.relu ? u->y[l] : nullptr,
                                                          u->grads + u->off_w[l], u->grads + u->off_b[l], s.cout, nz, ny,
                                                          nx);
}

int backward(flux_unet* u, const float* x) {
    auto copy_part = [&](const float* src, float* dst, int64_t n) {
        k_ucopy<<<blocks_for(n, UTPB), UTPB, 0, g_stream>>>(src, dst, n);
    };
    // BF16 networks: each layer's gradients on MFMA from its masked output
    // gradient in channels-last BF16 (weight_grad converts it, data_grad
    // follows and reuses it)
    const bool bf = u->precision == FLUX_UNET_BF16;
    if (bf) pack_transposed(u);
    auto weight_grad = [&](flux_unet* uu, int l, const float* xx) {
        if (bf) {
            grad_to_cl(uu, l);
            weight_grad_bf16(uu, l);
        } else {
            ::weight_grad(uu, l, xx);
        }
    };
    auto data_grad = [&](flux_unet* uu, int l, float* dst) {
        if (bf) data_grad_bf16(uu, l, dst);
        else ::data_grad(uu, l, dst);
    };
    // out
    weight_grad(u, L_OUT, x);
    data_grad(u, L_OUT, u->gy[L_D0B]);
    // dec0b, dec0a (input = [up(d1b) 32 | s0 16])
    weight_grad(u, L_D0B, x);
    data_grad(u, L_D0B, u->gy[L_D0A]);
    weight_grad(u, L_D0A, x);
    data_grad(u, L_D0A, u->gin_up[L_D0A]);  // 48 channels at level 0
    int64_t v0 = lvox(u, 0), v1 = lvox(u, 1), v2 = lvox(u, 2), v3 = lvox(u, 3);
    // grad of d1b = up_back(channels 0..31); grad of s0 (enc0b out) = channels 32..47
    k_up_back<<<blocks_for(32 * v1, UTPB), UTPB, 0, g_stream>>>(u->gin_up[L_D0A], u->gy[L_D1B], 32, u->n[2], u->n[1],
                                                               u->n[0]);
    copy_part(u->gin_up[L_D0A] + 32 * v0, u->gy[L_E0B], 16 * v0);
    // dec1b, dec1a (input = [up(d2) 64 | s1 32])
    weight_grad(u, L_D1B, x);
    data_grad(u, L_D1B, u->gy[L_D1A]);
    weight_grad(u, L_D1A, x);
    data_grad(u, L_D1A, u->gin_up[L_D1A]);  // 96 channels at level 1
    k_up_back<<<blocks_for(64 * v2, UTPB), UTPB, 0, g_stream>>>(u->gin_up[L_D1A], u->gy[L_D2], 64, u->n[2] >> 1,
                                                               u->n[1] >> 1, u->n[0] >> 1);
    copy_part(u->gin_up[L_D1A] + 64 * v1, u->gy[L_E1B], 32 * v1);
    // dec2 (input = [up(mid) 64 | s2 64])
    weight_grad(u, L_D2, x);
    data_grad(u, L_D2, u->gin_up[L_D2]);  // 128 channels at level 2
    k_up_back<<<blocks_for(64 * v3, UTPB), UTPB, 0, g_stream>>>(u->gin_up[L_D2], u->gy[L_MID], 64, u->n[2] >> 2,
                                                               u->n[1] >> 2, u->n[0] >> 2);
    copy_part(u->gin_up[L_D2] + 64 * v2, u->gy[L_E2B], 64 * v2);
    // mid (input pool(s2))
    weight_grad(u, L_MID, x);
    data_grad(u, L_MID, u->gpool[2]);
    k_pool_back_add<<<blocks_for(64 * v2, UTPB), UTPB, 0, g_stream>>>(u->gpool[2], u->gy[L_E2B], 64, u->n[2] >> 2,
                                                                     u->n[1] >> 2, u->n[0] >> 2);
    // enc2b, enc2a (input pool(s1))
    weight_grad(u, L_E2B, x);
    data_grad(u, L_E2B, u->gy[L_E2A]);
    weight_grad(u, L_E2A, x);
    data_grad(u, L_E2A, u->gpool[1]);
    k_pool_back_add<<<blocks_for(32 * v1, UTPB), UTPB, 0, g_stream>>>(u->gpool[1], u->gy[L_E1B], 32, u->n[2] >> 1,
                                                                     u->n[1] >> 1, u->n[0] >> 1);
    // enc1b, enc1a (input pool(s0))
    weight_grad(u, L_E1B, x);
    data_grad(u, L_E1B, u->gy[L_E1A]);
    weight_grad(u, L_E1A, x);
    data_grad(u, L_E1A, u->gpool[0]);
    k_pool_back_add<<<blocks_for(16 * v0, UTPB), UTPB, 0, g_stream>>>(u->gpool[0], u->gy[L_E0B], 16, u->n[2], u->n[1],
                                                                     u->n[0]);
    // enc0b, enc0a
    weight_grad(u, L_E0B, x);
    data_grad(u, L_E0B, u->gy[L_E0A]);
    weight_grad(u, L_E0A, x);
    FLUX_LAUNCH_CHECK();
    return FLUX_OK;
}

}  // namespace

extern "C" {

int64_t flux_unet_param_count(void) { return unet_param_count(); }

int flux_unet_destroy(flux_unet* u) {
    if (!u) return FLUX_OK;
    if (g_init) (void)hipStreamSynchronize(g_stream);
    dfree(u->params);
    dfree(u->grads);
    dfree(u->adam_m);
    dfree(u->adam_v);
    dfree(u->wt);
    for (int l = 0; l < kNumLayers; ++l) {
        dfree(u->y[l]);
        dfree(u->gy[l]);
        dfree(u->gin_up[l]);
    }
    for (int q = 0; q < 3; ++q) {
        dfree(u->pooled[q]);
        dfree(u->gpool[q]);
    }
    dfree(u->wq);
    dfree(u->split_part);
    dfree(u->x16);
    for (int l = 0; l < kNumLayers; ++l) dfree(u->y16[l]);
    for (int q = 0; q < 3; ++q) dfree(u->pooled16[q]);
    dfree(u->wqt);
    dfree(u->g16);
    dfree(u->dx_cl);
    dfree(u->wpart);
    dfree(u->zero_bias);
    dfree(u->part);
    dfree(u->scal);
    delete u;
    return FLUX_OK;
}

int flux_unet_create(int nx, int ny, int nz, int training, flux_unet** out) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(out, "unet: NULL argument");
    FLUX_REQUIRE(nx > 0 && ny > 0 && nz > 0 && nx % 8 == 0 && ny % 8 == 0 && nz % 8 == 0,
                 "unet: grid [%d, %d, %d] must be divisible by 8", nx, ny, nz);
    flux_unet* u = new flux_unet();
    {
        // FLUX_UNET_HALO=0 keeps every layer on the flat-tile kernel, =2 uses
        // the halo-tiled one wherever it fits (the same arithmetic; for
        // comparison)
        const char* e = getenv("FLUX_UNET_HALO");
        u->halo = (e && (e[0] == '0' || e[0] == '2')) ? e[0] - '0' : 1;
        const char* sp = getenv("FLUX_UNET_SPLIT");
        u->split = !(sp && sp[0] == '0');
        // FLUX_UNET_TILE=MxN (M, N in 1, 2, 4) forces the flat kernel's tile
        // on every layer (N capped at the layer's width); each output sums
        // its K terms in the same order whatever the tile, so any tile gives
        // the same bits (split K off): a check of the tile indexing
        const char* t = getenv("FLUX_UNET_TILE");
        int tm = 0, tn = 0;
        if (t && sscanf(t, "%dx%d", &tm, &tn) == 2) {
            auto valid = [](int v) { return v == 1 || v == 2 || v == 4; };
            if (!valid(tm) || !valid(tn)) {
                delete u;
                return set_error(FLUX_ERR_ARG, "unet: FLUX_UNET_TILE=%s: tile sides must be 1, 2 or 4", t);
            }
            u->tile_mt = tm;
            u->tile_nt = tn;
        }
    }
    u->n[0] = nx;
    u->n[1] = ny;
    u->n[2] = nz;
    for (int l = 0; l < 4; ++l) u->nvox[l] = lvox(u, l);
    int64_t off = 0, maxw = 0;
    for (int l = 0; l < kNumLayers; ++l) {
        u->off_w[l] = off;
        off += (int64_t)kLayers[l].cin * kLayers[l].cout * 27;
        u->off_b[l] = off;
        off += kLayers[l].cout;
        maxw = std::max<int64_t>(maxw, (int64_t)kLayers[l].cin * kLayers[l].cout * 27);
    }
    u->nparams = off;
    bool ok = true;
    auto alloc = [&](float*& p, int64_t n) {
        p = (float*)dmalloc(n * 4);
        ok = ok && p;
    };
    alloc(u->params, off);
    for (int l = 0; l < kNumLayers; ++l) alloc(u->y[l], (int64_t)kLayers[l].cout * u->nvox[kLayers[l].level]);
    alloc(u->pooled[0], 16 * u->nvox[1]);
    alloc(u->pooled[1], 32 * u->nvox[2]);
    alloc(u->pooled[2], 64 * u->nvox[3]);
    int64_t nb = blocks_for(16 * u->nvox[0] + 1, UTPB) + 16;
    u->part = (double*)dmalloc(nb * 8);
    u->scal = (double*)dmalloc(64);
    ok = ok && u->part && u->scal;
    if (training) {
        alloc(u->grads, off);
        alloc(u->adam_m, off);
        alloc(u->adam_v, off);
        alloc(u->wt, maxw);
        for (int l = 0; l < kNumLayers; ++l) alloc(u->gy[l], (int64_t)kLayers[l].cout * u->nvox[kLayers[l].level]);
        alloc(u->gin_up[L_D0A], 48 * u->nvox[0]);
        alloc(u->gin_up[L_D1A], 96 * u->nvox[1]);
        alloc(u->gin_up[L_D2], 128 * u->nvox[2]);
        alloc(u->gpool[0], 16 * u->nvox[1]);
        alloc(u->gpool[1], 32 * u->nvox[2]);
        alloc(u->gpool[2], 64 * u->nvox[3]);
    }
    if (!ok) {
        flux_unet_destroy(u);
        return set_error(FLUX_ERR_NOMEM, "unet: allocation failed");
    }
    if (training) {
        fill_device(u->adam_m, FLUX_F32, off, 0.0);
        fill_device(u->adam_v, FLUX_F32, off, 0.0);
    }
    FLUX_HIP(hipStreamSynchronize(g_stream));
    *out = u;
    return FLUX_OK;
}

int flux_unet_set_params(flux_unet* u, const float* host, int64_t count) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(u && host && count == u->nparams, "unet: expected %lld parameters", (long long)(u ? u->nparams : 0));
    FLUX_HIP(hipMemcpy(u->params, host, count * 4, hipMemcpyHostToDevice));
    u->version += 1;
    return FLUX_OK;
}

int flux_unet_get_params(flux_unet* u, float* host, int64_t count, int which) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(u && host && count == u->nparams, "unet: expected %lld parameters", (long long)(u ? u->nparams : 0));
    const float* src = which == 1 ? u->grads : u->params;
    FLUX_REQUIRE(src, "unet: gradients exist only for a training network");
    FLUX_HIP(hipStreamSynchronize(g_stream));
    FLUX_HIP(hipMemcpy(host, src, count * 4, hipMemcpyDeviceToHost));
    return FLUX_OK;
}

int flux_unet_set_precision(flux_unet* u, int precision) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(u, "unet: NULL");
    FLUX_REQUIRE(precision == FLUX_UNET_F32 || precision == FLUX_UNET_BF16, "unet: unknown precision %d", precision);
    if (precision == FLUX_UNET_BF16) FLUX_TRY(bf16_setup(u));
    u->precision = precision;
    return FLUX_OK;
}

// forward of a 3-channel input [3][nz][ny][nx]; out [nz][ny][nx]
int flux_unet_forward(flux_unet* u, const flux_field* input, flux_field* out) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(u, "unet: NULL");
    FLUX_TRY(check_field(input, FLUX_F32, 3 * u->nvox[0], "unet: input"));
    FLUX_TRY(check_field(out, FLUX_F32, u->nvox[0], "unet: output"));
    FLUX_TRY(infer(u, (const float*)input->d));
    k_ucopy<<<blocks_for(u->nvox[0], UTPB), UTPB, 0, g_stream>>>(u->y[L_OUT], (float*)out->d, u->nvox[0]);
    FLUX_LAUNCH_CHECK();
    FLUX_HIP(hipStreamSynchronize(g_stream));
    return FLUX_OK;
}

// forward + MSE against target + backward (+ Adam when lr > 0); *loss = MSE
int flux_unet_train_step(flux_unet* u, const flux_field* input, const flux_field* target, float lr, float beta1,
                         float beta2, float eps, double* loss) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(u && u->grads, "unet: not a training network");
    FLUX_TRY(check_field(input, FLUX_F32, 3 * u->nvox[0], "unet: input"));
    FLUX_TRY(check_field(target, FLUX_F32, u->nvox[0], "unet: target"));
    const float* x = (const float*)input->d;
    // BF16 networks train on the matrix cores (forward and backward)
    if (u->precision == FLUX_UNET_BF16) FLUX_TRY(bf16_train_setup(u));
    FLUX_TRY(infer(u, x));
    int64_t nv = u->nvox[0];
    unsigned nb = blocks_for(nv, UTPB);
    k_mse<<<nb, UTPB, 0, g_stream>>>(u->y[L_OUT], (const float*)target->d, u->gy[L_OUT], u->part, nv);
    k_reduce_fixed<<<1, UTPB, 0, g_stream>>>(u->part, (int)nb, u->scal);
    FLUX_TRY(backward(u, x));
    if (lr > 0.0f) {
        u->steps += 1;
        float c1 = 1.0f - powf(beta1, (float)u->steps), c2 = 1.0f - powf(beta2, (float)u->steps);
        k_adam<<<blocks_for(u->nparams, UTPB), UTPB, 0, g_stream>>>(u->params, u->grads, u->adam_m, u->adam_v,
                                                                     u->nparams, lr, beta1, beta2, eps, c1, c2);
        u->version += 1;
    }
    FLUX_LAUNCH_CHECK();
    double l = 0;
    FLUX_HIP(hipMemcpyAsync(&l, u->scal, 8, hipMemcpyDeviceToHost, g_stream));
    FLUX_HIP(hipStreamSynchronize(g_stream));
    if (loss) *loss = l / (double)nv;
    return FLUX_OK;
}

// forward + energy-norm loss + backward (+ Adam when lr > 0).  With the
// guess g = y on the free cells, A the system's operator (negative definite
// on the free cells), r the residual of the start, B = -A and e* = x*.(-r)
// the energy of the exact correction x*, the loss is the relative energy
// error left after the optimal step along g:
//   L = 1 - (g.r)^2 / ((g.B g) e*),   dL/dg = -2 (g.r) r / (q e*) - 2 (g.r)^2 A g / (q^2 e*)
// with q = g.B g = -g.A g (so a larger step along g never helps: the loss
// is scale-free, as the solver rescales g).  *loss = L.
int flux_unet_train_step_energy(flux_unet* u, const flux_field* input, flux_operator* op, const flux_field* r,
                                const flux_field* mask, double e_star, float lr, float beta1, float beta2, float eps,
                                double* loss) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(u && u->grads && op && r, "unet: energy training needs a training network, an operator and r");
    FLUX_REQUIRE(e_star > 0.0, "unet: e* must be positive, got %g", e_star);
    FLUX_TRY(check_field(input, FLUX_F32, 3 * u->nvox[0], "unet: input"));
    const int64_t nv = u->nvox[0];
    FLUX_TRY(check_field(r, FLUX_F32, nv, "unet: residual"));
    if (mask) FLUX_TRY(check_field(mask, FLUX_U8, nv, "unet: mask"));
    const uint8_t* m = mask ? (const uint8_t*)mask->d : nullptr;
    const float* x = (const float*)input->d;
    // BF16 networks train on the matrix cores (forward and backward)
    if (u->precision == FLUX_UNET_BF16) FLUX_TRY(bf16_train_setup(u));
    FLUX_TRY(infer(u, x));
    const unsigned nb = blocks_for(nv, UTPB);
    // g into gy[L_OUT] (scratch until the gradient overwrites it), A g into
    // gin_up[L_D0A] (48 channels at level 0; backward writes it only after
    // the gradient below has consumed A g)
    float* g = u->gy[L_OUT];
    float* ag = u->gin_up[L_D0A];
    k_mask_out<<<nb, UTPB, 0, g_stream>>>(u->y[L_OUT], m, g, nv);
    flux_field gf, af;
    gf.d = g;
    af.d = ag;
    for (flux_field* f : {&gf, &af}) {
        f->dtype = FLUX_F32;
        f->ndim = 1;
        f->shape[0] = nv;
        f->count = nv;
        f->bytes = (size_t)nv * 4;
    }
    FLUX_TRY(flux_operator_apply(op, &gf, &af));
    k_dot2<<<nb, UTPB, 0, g_stream>>>(g, (const float*)r->d, g, ag, nv, u->part);
    std::vector<double> h(2 * (size_t)nb);
    FLUX_HIP(hipMemcpyAsync(h.data(), u->part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, g_stream));
    FLUX_HIP(hipStreamSynchronize(g_stream));
    double gr = 0, gag = 0;
    for (unsigned q = 0; q < nb; ++q) {
        gr += h[q];
        gag += h[nb + q];
    }
    const double qv = -gag;
    double l = 1.0;
    float c1 = 0.0f, c2 = 0.0f;
    if (qv > 0 && std::isfinite(qv) && std::isfinite(gr)) {
        l = 1.0 - gr * gr / (qv * e_star);
        c1 = (float)(-2.0 * gr / (qv * e_star));
        c2 = (float)(-2.0 * gr * gr / (qv * qv * e_star));
    }
    k_energy_grad<<<nb, UTPB, 0, g_stream>>>((const float*)r->d, ag, m, c1, c2, u->gy[L_OUT], nv);
    FLUX_TRY(backward(u, x));
    if (lr > 0.0f) {
        u->steps += 1;
        float b1c = 1.0f - powf(beta1, (float)u->steps), b2c = 1.0f - powf(beta2, (float)u->steps);
        k_adam<<<blocks_for(u->nparams, UTPB), UTPB, 0, g_stream>>>(u->params, u->grads, u->adam_m, u->adam_v,
                                                                     u->nparams, lr, beta1, beta2, eps, b1c, b2c);
        u->version += 1;
    }
    FLUX_LAUNCH_CHECK();
    FLUX_HIP(hipStreamSynchronize(g_stream));
    if (loss) *loss = l;
    return FLUX_OK;
}

// the network's input channels from an assembled pressure system
int flux_unet_features(flux_unet* u, const flux_field* b, const flux_field* mask, const flux_field* const* weights,
                       float density_ratio, flux_field* input) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(u, "unet: NULL");
    int64_t nv = u->nvox[0];
    FLUX_TRY(check_field(b, FLUX_F32, nv, "unet: b"));
    FLUX_TRY(check_field(input, FLUX_F32, 3 * nv, "unet: input"));
    if (mask) FLUX_TRY(check_field(mask, FLUX_U8, nv, "unet: mask"));
    unsigned nb = blocks_for(nv, UTPB);
    k_sumsq<<<nb, UTPB, 0, g_stream>>>((const float*)b->d, nv, u->part);
    k_reduce_fixed<<<1, UTPB, 0, g_stream>>>(u->part, (int)nb, u->scal + 1);
    float ilr = density_ratio > 1.0f ? 1.0f / logf(density_ratio) : 1.0f;
    k_features<<<nb, UTPB, 0, g_stream>>>((const float*)b->d, mask ? (const uint8_t*)mask->d : nullptr,
                                          weights ? (const float*)weights[0]->d : nullptr,
                                          weights ? (const float*)weights[1]->d : nullptr,
                                          weights ? (const float*)weights[2]->d : nullptr, u->n[0], u->n[1], u->n[2],
                                          u->scal + 1, ilr, (float*)input->d);
    FLUX_LAUNCH_CHECK();
    FLUX_HIP(hipStreamSynchronize(g_stream));
    return FLUX_OK;
}

// guess = network output * rms(b) of the last flux_unet_features call
int flux_unet_guess(flux_unet* u, const flux_field* input, flux_field* guess) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(u, "unet: NULL");
    int64_t nv = u->nvox[0];
    FLUX_TRY(check_field(input, FLUX_F32, 3 * nv, "unet: input"));
    FLUX_TRY(check_field(guess, FLUX_F32, nv, "unet: guess"));
    FLUX_TRY(infer(u, (const float*)input->d));
    k_scale_out<<<blocks_for(nv, UTPB), UTPB, 0, g_stream>>>(u->y[L_OUT], u->scal + 1, nv, (float*)guess->d);
    FLUX_LAUNCH_CHECK();
    FLUX_HIP(hipStreamSynchronize(g_stream));
    return FLUX_OK;
}

int flux_unet_guess_into(flux_unet* u, const flux_field* input, const flux_field* mask, flux_field* pressure) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(u, "unet: NULL");
    int64_t nv = u->nvox[0];
    FLUX_TRY(check_field(input, FLUX_F32, 3 * nv, "unet: input"));
    FLUX_TRY(check_field(pressure, FLUX_F32, nv, "unet: pressure"));
    if (mask) FLUX_TRY(check_field(mask, FLUX_U8, nv, "unet: mask"));
    FLUX_TRY(infer(u, (const float*)input->d));
    k_guess_into<<<blocks_for(nv, UTPB), UTPB, 0, g_stream>>>(u->y[L_OUT], u->scal + 1,
                                                               mask ? (const uint8_t*)mask->d : nullptr, nv,
                                                               (float*)pressure->d);
    FLUX_LAUNCH_CHECK();
    FLUX_HIP(hipStreamSynchronize(g_stream));
    return FLUX_OK;
}

}  // extern "C"
