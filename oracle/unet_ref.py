"""CPU oracle for the learned initial guess (TEST INFRASTRUCTURE ONLY).

The reference has no learned component; this restates, with numpy in float64,
the network the HIP kernels implement, so that their forward outputs and
gradients can be checked:

  3-D conv U-Net, 4 levels, channels 16/32/64/64, 3x3x3 kernels, stride 1,
  zero padding 1, ReLU after every conv except the last, 2x2x2 average pooling
  down, nearest-neighbour upsampling up, skip connections by channel concat
  (upsampled first, skip second).  ~0.69M parameters.

Layout: activations [C, nz, ny, nx] (x fastest, the fields' layout), weights
[C_out, C_in, 3, 3, 3] indexed [co, ci, dz, dy, dx], bias [C_out].
"""
import numpy as np

# (name, c_in, c_out, relu)
LAYERS = [
    ("enc0a", 3, 16, True), ("enc0b", 16, 16, True),
    ("enc1a", 16, 32, True), ("enc1b", 32, 32, True),
    ("enc2a", 32, 64, True), ("enc2b", 64, 64, True),
    ("mid", 64, 64, True),
    ("dec2", 128, 64, True),
    ("dec1a", 96, 32, True), ("dec1b", 32, 32, True),
    ("dec0a", 48, 16, True), ("dec0b", 16, 16, True),
    ("out", 16, 1, False),
]


def param_count():
    return sum(ci * co * 27 + co for _, ci, co, _ in LAYERS)


def init_params(seed=0):
    """He-uniform weights, zero bias (the HIP initialiser draws the same
    numbers from the same mulberry32 stream; see tsl_array_n_amd.learned)."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, ci, co, _ in LAYERS:
        bound = np.sqrt(6.0 / (ci * 27))
        p[name + ".w"] = rng.uniform(-bound, bound, (co, ci, 3, 3, 3))
        p[name + ".b"] = np.zeros(co)
    return p


def conv3d(x, w, b):
    """x [ci, z, y, x] -> [co, z, y, x], zero padding 1."""
    ci, nz, ny, nx = x.shape
    xp = np.zeros((ci, nz + 2, ny + 2, nx + 2))
    xp[:, 1:-1, 1:-1, 1:-1] = x
    out = np.zeros((w.shape[0], nz, ny, nx))
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                patch = xp[:, dz:dz + nz, dy:dy + ny, dx:dx + nx]
                out += np.einsum("oc,czyx->ozyx", w[:, :, dz, dy, dx], patch)
    return out + b[:, None, None, None]


def conv3d_backward(x, w, gout):
    """gradients (dx, dw, db) of conv3d at x for upstream gout."""
    ci, nz, ny, nx = x.shape
    xp = np.zeros((ci, nz + 2, ny + 2, nx + 2))
    xp[:, 1:-1, 1:-1, 1:-1] = x
    dxp = np.zeros_like(xp)
    dw = np.zeros_like(w)
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                patch = xp[:, dz:dz + nz, dy:dy + ny, dx:dx + nx]
                dw[:, :, dz, dy, dx] = np.einsum("ozyx,czyx->oc", gout, patch)
                dxp[:, dz:dz + nz, dy:dy + ny, dx:dx + nx] += np.einsum("oc,ozyx->czyx", w[:, :, dz, dy, dx], gout)
    return dxp[:, 1:-1, 1:-1, 1:-1], dw, gout.sum(axis=(1, 2, 3))


def pool(x):
    c, nz, ny, nx = x.shape
    return x.reshape(c, nz // 2, 2, ny // 2, 2, nx // 2, 2).mean(axis=(2, 4, 6))


def pool_backward(g):
    return np.repeat(np.repeat(np.repeat(g, 2, 1), 2, 2), 2, 3) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 1), 2, 2), 2, 3)


def up_backward(g):
    c, nz, ny, nx = g.shape
    return g.reshape(c, nz // 2, 2, ny // 2, 2, nx // 2, 2).sum(axis=(2, 4, 6))


def forward(p, x):
    """returns (prediction [1, z, y, x], cache for backward)"""
    cache = {}

    def layer(name, inp):
        relu = dict((n, r) for n, _, _, r in LAYERS)[name]
        z = conv3d(inp, p[name + ".w"], p[name + ".b"])
        cache[name] = (inp, z)
        return np.maximum(z, 0) if relu else z

    s0 = layer("enc0b", layer("enc0a", x))
    s1 = layer("enc1b", layer("enc1a", pool(s0)))
    s2 = layer("enc2b", layer("enc2a", pool(s1)))
    m = layer("mid", pool(s2))
    d2 = layer("dec2", np.concatenate([up(m), s2], 0))
    d1 = layer("dec1b", layer("dec1a", np.concatenate([up(d2), s1], 0)))
    d0 = layer("dec0b", layer("dec0a", np.concatenate([up(d1), s0], 0)))
    y = layer("out", d0)
    return y, cache


def backward(p, cache, gy):
    """gradients of every parameter for upstream gradient gy of the output."""
    g = {}
    relu = dict((n, r) for n, _, _, r in LAYERS)

    def layer_back(name, gout):
        inp, z = cache[name]
        if relu[name]:
            gout = gout * (z > 0)
        dx, dw, db = conv3d_backward(inp, p[name + ".w"], gout)
        g[name + ".w"] = dw
        g[name + ".b"] = db
        return dx

    gd0 = layer_back("out", gy)
    gcat0 = layer_back("dec0a", layer_back("dec0b", gd0))
    gs0 = gcat0[32:]
    gcat1 = layer_back("dec1a", layer_back("dec1b", up_backward(gcat0[:32])))
    gs1 = gcat1[64:]
    gcat2 = layer_back("dec2", up_backward(gcat1[:64]))
    gs2 = gcat2[64:]
    gm = layer_back("mid", up_backward(gcat2[:64]))
    gs2 = gs2 + pool_backward(gm)
    gp1 = layer_back("enc2a", layer_back("enc2b", gs2))
    gs1 = gs1 + pool_backward(gp1)
    gp0 = layer_back("enc1a", layer_back("enc1b", gs1))
    gs0 = gs0 + pool_backward(gp0)
    gx = layer_back("enc0a", layer_back("enc0b", gs0))
    return g, gx


def mse_loss(y, t):
    d = y - t
    return float((d * d).mean()), 2.0 * d / d.size


def to_bf16(a):
    """round to the nearest BF16 value, ties to even (finite inputs), returned
    as float64: the upper 16 bits of the float32 after adding 0x7fff plus
    the lowest kept bit."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def forward_bf16(p, x):
    """the BF16 inference form: input and weights rounded to BF16, exact
    products summed at high precision, bias added and ReLU applied before
    each layer's output is rounded to BF16; pooling averages BF16 values and
    rounds the mean; the last layer's output stays unrounded.  Returns the
    prediction [1, z, y, x]."""
    q = {k: (to_bf16(v) if k.endswith(".w") else np.asarray(v, np.float32).astype(np.float64)) for k, v in p.items()}
    relu = dict((n, r) for n, _, _, r in LAYERS)

    def layer(name, inp):
        z = conv3d(inp, q[name + ".w"], q[name + ".b"])
        z = np.maximum(z, 0) if relu[name] else z
        return to_bf16(z) if name != "out" else z

    def bpool(t):
        return to_bf16(pool(t))

    s0 = layer("enc0b", layer("enc0a", to_bf16(x)))
    s1 = layer("enc1b", layer("enc1a", bpool(s0)))
    s2 = layer("enc2b", layer("enc2a", bpool(s1)))
    m = layer("mid", bpool(s2))
    d2 = layer("dec2", np.concatenate([up(m), s2], 0))
    d1 = layer("dec1b", layer("dec1a", np.concatenate([up(d2), s1], 0)))
    d0 = layer("dec0b", layer("dec0a", np.concatenate([up(d1), s0], 0)))
    return layer("out", d0)
