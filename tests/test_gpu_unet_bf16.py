"""BF16 inference of the learned guess on the matrix cores (GPU tests).

The implicit-GEMM convolutions (v_mfma_f32_16x16x32_bf16, BF16 weights and
channels-last BF16 activations, FP32 accumulation) against the oracle's BF16
restatement (unet_ref.forward_bf16: the same roundings, exact sums) at
relative L2 <= 5e-3 -- the sums differ in order only, which flips an
occasional BF16 rounding by one unit -- and against the float64 forward at
the BF16 bar, relative L2 <= 2e-2.  Timing of the FP32 and BF16 forwards is
printed at 64^3 and 128^3."""
import numpy as np
import pytest

from oracle import unet_ref as U

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def random_params(seed):
    rng = np.random.default_rng(seed)
    p = U.init_params(seed)
    for k in p:
        if k.endswith(".b"):
            p[k] = rng.standard_normal(p[k].shape) * 0.1
    return p


@pytest.mark.parametrize("n", [(16, 8, 8), (24, 16, 8), (8, 8, 32)])
def test_bf16_forward_matches_oracle(flux, n):
    from tsl_array_n_amd import learned as L
    p = random_params(3)
    x = np.random.default_rng(4).standard_normal((3, n[2], n[1], n[0]))
    net = L.UNet(n, params={k: v.astype(np.float32) for k, v in p.items()}, precision="bf16")
    y = net.forward(x.astype(np.float32).ravel()).to_array()
    yb = U.forward_bf16(p, x.astype(np.float32))
    y64, _ = U.forward(p, x.astype(np.float32))
    r_b, r_64 = rel(y, yb), rel(y, y64)
    print(n, "vs bf16 oracle %.2e  vs float64 %.2e  oracle bf16 vs float64 %.2e" % (r_b, r_64, rel(yb, y64)))
    assert r_b <= 5e-3 and r_64 <= 2e-2


def test_bf16_weights_follow_parameter_updates(flux):
    """the packed BF16 weights are rebuilt after set_params and after an Adam
    step, and switching precision on one network gives both forms"""
    from tsl_array_n_amd import learned as L
    n = (16, 16, 8)
    x = np.random.default_rng(5).standard_normal((3, n[2], n[1], n[0])).astype(np.float32)
    net = L.UNet(n, training=True, params={k: v.astype(np.float32) for k, v in random_params(1).items()},
                 precision="bf16")
    y1 = net.forward(x.ravel()).to_array()
    p2 = random_params(2)
    net.set_params({k: v.astype(np.float32) for k, v in p2.items()})
    y2 = net.forward().to_array()
    assert rel(y2, U.forward_bf16(p2, x)) <= 5e-3 and rel(y1, y2) > 0.1
    target = flux.array_n("float", (int(np.prod(n)),), np.tanh(x[0]).ravel())
    net.train_step(target, lr=1e-2)
    after = L.unflatten(net.get())
    y3 = net.forward().to_array()
    assert rel(y3, U.forward_bf16({k: v.astype(np.float64) for k, v in after.items()}, x)) <= 5e-3
    net.set_precision("f32")
    y32 = net.forward().to_array()
    assert rel(y32, U.forward({k: v.astype(np.float64) for k, v in after.items()}, x)[0]) <= 1e-4
    assert rel(y3, y32) <= 2e-2


@pytest.mark.parametrize("size", [64, 128])
def test_bf16_inference_time(flux, size):
    """FP32 and BF16 forward of the same network: outputs agree to the BF16
    bar and the time of each is printed"""
    F = flux
    from tsl_array_n_amd import learned as L
    n = (size,) * 3
    p = random_params(7)
    x = np.random.default_rng(8).standard_normal((3, size, size, size)).astype(np.float32)
    net = L.UNet(n, params={k: v.astype(np.float32) for k, v in p.items()})
    ys, ms = {}, {}
    for prec in ("f32", "bf16"):
        net.set_precision(prec)
        ys[prec] = net.forward(x.ravel()).to_array()
        reps = 3 if prec == "f32" else 20
        F.timer_start()
        for _ in range(reps):
            net.forward()
        ms[prec] = F.timer_stop() / reps
    flops = 2 * 27 * sum(ci * co * (size >> lv) ** 3 for (ci, co), lv in zip(
        [(3, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (64, 64), (128, 64), (96, 32), (32, 32), (48, 16),
         (16, 16), (16, 1)], [0, 0, 1, 1, 2, 2, 3, 2, 1, 1, 0, 0, 0]))
    print("%d^3 forward: FP32 %.3f ms (%.1f TFLOP/s)  BF16 %.3f ms (%.1f TFLOP/s)  rel %.2e" % (
        size, ms["f32"], flops / ms["f32"] * 1e-9, ms["bf16"], flops / ms["bf16"] * 1e-9, rel(ys["bf16"], ys["f32"])))
    assert rel(ys["bf16"], ys["f32"]) <= 2e-2
    assert ms["bf16"] < ms["f32"]


@pytest.mark.parametrize("n,mode", [((32, 16, 8), "2"), ((64, 32, 32), "2"), ((128, 64, 32), "1")])
def test_halo_tiled_convolutions_are_bit_identical(flux, n, mode):
    """the halo-tiled kernel (levels of whole 16x4x4 bricks, input brick in
    LDS) does the flat kernel's arithmetic in the same order: forced
    wherever it fits (mode 2) and by the default rule (mode 1)"""
    import os
    from tsl_array_n_amd import learned as L
    p = {k: v.astype(np.float32) for k, v in random_params(9).items()}
    x = np.random.default_rng(10).standard_normal(3 * int(np.prod(n))).astype(np.float32)
    outs = []
    for flag in ("0", mode):
        # split K (a different summation order) off on both sides
        saved = {k: os.environ.get(k) for k in ("FLUX_UNET_HALO", "FLUX_UNET_SPLIT")}
        os.environ["FLUX_UNET_HALO"] = flag
        os.environ["FLUX_UNET_SPLIT"] = "0"
        try:
            net = L.UNet(n, params=p, precision="bf16")
        finally:
            for k, v in saved.items():
                if v is None:
                    del os.environ[k]
                else:
                    os.environ[k] = v
        outs.append(net.forward(x).to_array())
    assert np.all(np.isfinite(outs[0])) and np.any(outs[0] != 0)
    np.testing.assert_array_equal(outs[1], outs[0])


def test_split_k_matches_unsplit(flux):
    """small levels split K and add the FP32 partial sums in a fixed order:
    the same products summed in another order, so the output agrees with
    the unsplit kernel to BF16 rounding flips (relative L2 <= 2e-3) and is
    reproducible bit for bit"""
    import os
    from tsl_array_n_amd import learned as L
    n = (32, 32, 16)
    p = {k: v.astype(np.float32) for k, v in random_params(11).items()}
    x = np.random.default_rng(12).standard_normal(3 * int(np.prod(n))).astype(np.float32)
    outs = []
    for flag in ("0", "1", "1"):
        old = os.environ.get("FLUX_UNET_SPLIT")
        os.environ["FLUX_UNET_SPLIT"] = flag
        try:
            net = L.UNet(n, params=p, precision="bf16")
        finally:
            if old is None:
                del os.environ["FLUX_UNET_SPLIT"]
            else:
                os.environ["FLUX_UNET_SPLIT"] = old
        outs.append(net.forward(x).to_array())
    print("split vs unsplit rel %.2e" % rel(outs[1], outs[0]))
    assert rel(outs[1], outs[0]) <= 2e-3
    np.testing.assert_array_equal(outs[2], outs[1])


@pytest.mark.parametrize("n", [(64, 32, 32), (40, 24, 16)])
def test_forced_tiles_are_bit_identical(flux, n, monkeypatch):
    """the flat kernel's wide tiles (MT rows of 16 voxels x NT columns of 16
    channels) sum each output's K terms in the order of the 1x1 tile, so
    every forced tile gives the 1x1 bits, ragged voxel tails (40x24x16:
    30 voxels at the coarsest level) and capped widths included; split K
    and the halo kernel are off on every side"""
    from tsl_array_n_amd import learned as L
    p = {k: v.astype(np.float32) for k, v in random_params(13).items()}
    x = np.random.default_rng(14).standard_normal(3 * int(np.prod(n))).astype(np.float32)
    monkeypatch.setenv("FLUX_UNET_SPLIT", "0")
    monkeypatch.setenv("FLUX_UNET_HALO", "0")
    outs = {}
    for tile in ("1x1", "4x4", "4x2", "2x1", "2x4", "1x2"):
        monkeypatch.setenv("FLUX_UNET_TILE", tile)
        net = L.UNet(n, params=p, precision="bf16")
        outs[tile] = net.forward(x).to_array()
        del net
    assert np.all(np.isfinite(outs["1x1"])) and np.any(outs["1x1"] != 0)
    for tile, y in outs.items():
        np.testing.assert_array_equal(y, outs["1x1"], err_msg=tile)


def test_bad_forced_tile_is_rejected(flux, monkeypatch):
    from tsl_array_n_amd import learned as L
    monkeypatch.setenv("FLUX_UNET_TILE", "3x1")
    with pytest.raises(Exception, match="tile sides"):
        L.UNet((16, 16, 16), precision="bf16")
