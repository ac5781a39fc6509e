"""The learned guess's HIP kernels against the CPU oracle (GPU tests):
forward output and every parameter gradient of the U-Net, FP32 kernels vs
the float64 oracle, relative L2 <= 1e-4."""
import numpy as np
import pytest

from oracle import unet_ref as U

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def L(flux):
    from tsl_array_n_amd import learned
    return learned


def test_param_layout(L):
    assert L.param_count() == U.param_count()


def test_forward_and_gradients_match_oracle(flux, L):
    n = (16, 8, 8)
    rng = np.random.default_rng(0)
    p = U.init_params(1)
    for k in p:
        if k.endswith(".b"):
            p[k] = rng.standard_normal(p[k].shape) * 0.1
    x = rng.standard_normal((3, n[2], n[1], n[0]))
    t = rng.standard_normal((1, n[2], n[1], n[0]))
    net = L.UNet(n, training=True, params={k: v.astype(np.float32) for k, v in p.items()})
    y = net.forward(x.astype(np.float32).ravel()).to_array()
    yr, cache = U.forward(p, x)
    assert rel(y, yr) <= 1e-4
    target = flux.array_n("float", (int(np.prod(n)),), t.astype(np.float32).ravel())
    loss = net.train_step(target, lr=0.0)
    lr_, gy = U.mse_loss(yr, t)
    g, _ = U.backward(p, cache, gy)
    assert abs(loss - lr_) <= 1e-4 * lr_
    grads = L.unflatten(net.get(which=1))
    for name in g:
        assert rel(grads[name], g[name]) <= 1e-4, name


def test_training_reduces_loss(flux, L):
    n = (16, 16, 16)
    rng = np.random.default_rng(2)
    x = rng.standard_normal((3, 16, 16, 16)).astype(np.float32)
    t = np.tanh(x[0] + 0.5 * x[1]).astype(np.float32)
    net = L.UNet(n, training=True)
    target = flux.array_n("float", (4096,), t.ravel())
    first = net.train_step(target, x=x.ravel(), lr=1e-3)
    for _ in range(60):
        last = net.train_step(target, lr=1e-3)
    assert last < 0.5 * first
