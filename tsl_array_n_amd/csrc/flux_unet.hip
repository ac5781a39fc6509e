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

static const LayerSpec kLayers[] = {
    {"enc0a", 3, 16, 1, 0},  {"enc0b", 16, 16, 1, 0}, {"enc1a", 16, 32, 1, 1}, {"enc1b", 32, 32, 1, 1},
    {"enc2a", 32, 64, 1, 2}, {"enc2b", 64, 64, 1, 2}, {"mid", 64, 64, 1, 3},   {"dec2", 128, 64, 1, 2},
    {"dec1a", 96, 32, 1, 1}, {"dec1b", 32, 32, 1, 1}, {"dec0a", 48, 16, 1, 0}, {"dec0b", 16, 16, 1, 0},
    {"out", 16, 1, 0, 0},
};
static constexpr int kNumLayers = 13;
enum { L_E0A, L_E0B, L_E1A, L_E1B, L_E2A, L_E2B, L_MID, L_D2, L_D1A, L_D1B, L_D0A, L_D0B, L_OUT };

int64_t unet_param_count() {
    int64_t n = 0;
    for (int l = 0; l < kNumLayers; ++l) n += (int64_t)kLayers[l].cin * kLayers[l].cout * 27 + kLayers[l].cout;
    return n;
}

}  // namespace flux

using namespace flux;

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

// forward of a 3-channel input [3][nz][ny][nx]; out [nz][ny][nx]
int flux_unet_forward(flux_unet* u, const flux_field* input, flux_field* out) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(u, "unet: NULL");
    FLUX_TRY(check_field(input, FLUX_F32, 3 * u->nvox[0], "unet: input"));
    FLUX_TRY(check_field(out, FLUX_F32, u->nvox[0], "unet: output"));
    FLUX_TRY(forward(u, (const float*)input->d));
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
    FLUX_TRY(forward(u, x));
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
    }
    FLUX_LAUNCH_CHECK();
    double l = 0;
    FLUX_HIP(hipMemcpyAsync(&l, u->scal, 8, hipMemcpyDeviceToHost, g_stream));
    FLUX_HIP(hipStreamSynchronize(g_stream));
    if (loss) *loss = l / (double)nv;
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
    FLUX_TRY(forward(u, (const float*)input->d));
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
    FLUX_TRY(forward(u, (const float*)input->d));
    k_guess_into<<<blocks_for(nv, UTPB), UTPB, 0, g_stream>>>(u->y[L_OUT], u->scal + 1,
                                                               mask ? (const uint8_t*)mask->d : nullptr, nv,
                                                               (float*)pressure->d);
    FLUX_LAUNCH_CHECK();
    FLUX_HIP(hipStreamSynchronize(g_stream));
    return FLUX_OK;
}

}  // extern "C"
