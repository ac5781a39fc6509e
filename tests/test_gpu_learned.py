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


def test_learned_guess_is_accepted_at_the_same_tolerance(flux, L):
    """Trained on MGPCG-produced pairs; on held-out scenes the learned start
    reaches the same FP64 true-residual tolerance as the cold start, and the
    iteration counts of both are reported."""
    F = flux
    from tsl_array_n_amd import scenes
    n = 32
    pairs = L.training_pairs(F, n, 9, seed=5, levels=4)
    net, losses = L.train(F, n, pairs, 150, lr=2e-3)
    assert losses[-1] < losses[0]
    results = []
    for seed in (101, 202):
        sc = scenes.random_scene(n, seed)
        shape = sc["shape"]
        beta = [F.array_n("float", F.face_shape(shape, a), sc["beta"][a]) for a in range(3)]
        act = F.array_n("u8", shape, sc["active"])
        tgt = F.array_n("float", shape, sc["target"])
        out = []
        for mode in ("cold", "scaled"):
            vel = [F.array_n("float", F.face_shape(shape, a), sc["vel"][a]) for a in range(3)]
            ps = F.create_grid_pressure_solver3(resolution=shape, grid_spacing=sc["spacing"], dirichlet=(act, tgt),
                                                face_weights=beta, multigrid={"levels": 4}, tolerance=1e-6)
            if mode == "cold":
                d = ps.project(vel)
            else:
                d = L.LearnedProjector(ps, net, sc["ratio"]).project(vel)
            assert d["converged"] and d["residual"] <= 1e-6 * max(d["bNorm"], 1.0)
            out.append(d["iterations"])
        results.append(out)
    print("iterations cold / scaled guess:", results)


def test_scaled_start_along_exact_and_random_directions(flux):
    """pressure += alpha g with alpha = g.r / g.Ag: along the converged
    solution alpha is 1 and the solve needs (almost) no iterations; along
    random noise alpha is ~0 and the solve costs what the cold start costs."""
    F = flux
    from tsl_array_n_amd import scenes
    n = 32
    sc = scenes.random_scene(n, 303)
    shape = sc["shape"]
    beta = [F.array_n("float", F.face_shape(shape, a), sc["beta"][a]) for a in range(3)]
    act = F.array_n("u8", shape, sc["active"])
    tgt = F.array_n("float", shape, sc["target"])

    def solver():
        return F.create_grid_pressure_solver3(resolution=shape, grid_spacing=sc["spacing"], dirichlet=(act, tgt),
                                              face_weights=beta, multigrid={"levels": 4}, tolerance=1e-6)

    def vel():
        return [F.array_n("float", F.face_shape(shape, a), sc["vel"][a]) for a in range(3)]

    cold = solver()
    d0 = cold.project(vel())
    assert d0["converged"]
    exact = F.array_n("float", shape, cold.pressure.to_array())
    s1 = solver()
    v1 = vel()
    s1.assemble(v1)
    alpha = s1.scaled_guess(exact)
    assert abs(alpha - 1.0) < 1e-3, alpha
    d1 = s1.solve(v1)
    assert d1["converged"] and d1["iterations"] <= 3, d1
    noise = F.array_n("float", shape, np.random.default_rng(9).standard_normal(int(np.prod(shape))).astype(np.float32))
    s2 = solver()
    v2 = vel()
    s2.assemble(v2)
    alpha = s2.scaled_guess(noise)
    d2 = s2.solve(v2)
    assert d2["converged"] and d2["residual"] <= 1e-6 * d2["bNorm"]
    assert d2["iterations"] <= d0["iterations"] + 2, (d0, d2, alpha)
    print("cold", d0["iterations"], "exact", d1["iterations"], "noise", d2["iterations"], "alpha", alpha)


def test_energy_loss_gradients_match_oracle(flux, L):
    """flux_unet_train_step_energy: the loss and every parameter gradient
    against the float64 oracle -- dL/dy = -2 (g.r) r / (q e*) - 2 (g.r)^2 A g
    / (q^2 e*) on the free cells (g the output there, q = -g.A g) computed
    in numpy with the oracle's operator, then the oracle's backward."""
    from oracle import pyoracle as O
    F = flux
    n = (16, 8, 8)
    N = int(np.prod(n))
    rng = np.random.default_rng(4)
    p = U.init_params(3)
    for k in p:
        if k.endswith(".b"):
            p[k] = rng.standard_normal(p[k].shape) * 0.1
    x = rng.standard_normal((3, n[2], n[1], n[0]))
    mask = (rng.random(N) < 0.1).astype(np.uint8)
    w = [rng.uniform(0.5, 2.0, int(np.prod(F.face_shape(n, a)))).astype(np.float32) for a in range(3)]
    h = (0.5, 0.5, 0.5)
    r = np.where(mask == 0, rng.standard_normal(N), 0.0).astype(np.float32)
    e_star = 3.7
    net = L.UNet(n, training=True, params={k: v.astype(np.float32) for k, v in p.items()})
    mf = F.array_n("u8", n, mask)
    wf = [F.array_n("float", F.face_shape(n, a), w[a]) for a in range(3)]
    pair = {"x": F.array_n("float", (3 * N,), x.astype(np.float32).ravel()), "r": F.array_n("float", n, r),
            "mask": mf, "op": F.create_laplacian_operator(n, h, mf, wf), "e_star": e_star}
    loss = net.train_step_energy(pair, lr=0.0)
    yr, cache = U.forward(p, x)
    free = mask == 0
    g = np.where(free, yr.ravel(), 0.0)
    ag = O.apply(n, h, g, mask, [a.astype(np.float64) for a in w], suffix="f64")
    gr = float(g @ r)
    q = -float(g @ ag)
    want = 1.0 - gr * gr / (q * e_star)
    assert abs(loss - want) <= 1e-4 * max(abs(want), 1e-3), (loss, want)
    gy = np.where(free, -2 * gr / (q * e_star) * r - 2 * gr * gr / (q * q * e_star) * ag, 0.0)
    grads, _ = U.backward(p, cache, gy.reshape(1, n[2], n[1], n[0]))
    got = L.unflatten(net.get(which=1))
    for name in grads:
        assert rel(got[name], grads[name]) <= 1e-3, name


def test_energy_training_and_learned_coarse_vector(flux, L):
    """Trained on MGPCG-produced pairs with the energy-norm loss, the network
    supplies a coarse vector of the preconditioner (coarse space {free
    cells, guess}); on held-out scenes every solve reaches the same FP64
    true-residual tolerance, and the iterations and times of the cold
    start, of the free-cell correction alone and with the learned vector
    are recorded."""
    import time
    F = flux
    from tsl_array_n_amd import scenes
    n = 32
    pairs = L.energy_pairs(F, n, 9, seed=5, levels=4)
    assert len(pairs) >= 6
    net, losses = L.train_energy(F, n, pairs, 300, lr=1e-3)
    print("energy loss first %.3f last %.3f" % (np.mean(losses[:9]), np.mean(losses[-9:])))
    assert np.mean(losses[-9:]) < np.mean(losses[:9])
    rows = []
    for q, seed in enumerate((101, 202, 303)):
        sc = scenes.dam_break_system(n, box_max=(0.35, 0.55, 1.0)) if q == 2 else scenes.random_scene(n, seed)
        ratio = 1000.0 if q == 2 else sc["ratio"]
        shape = sc["shape"]
        beta = [F.array_n("float", F.face_shape(shape, a), sc["beta"][a]) for a in range(3)]
        act = F.array_n("u8", shape, sc["active"])
        tgt = F.array_n("float", shape, sc["target"])
        row = {}
        for mode in ("cold", "free", "learned"):
            vel = [F.array_n("float", F.face_shape(shape, a), sc["vel"][a]) for a in range(3)]
            ps = F.create_grid_pressure_solver3(resolution=shape, grid_spacing=sc["spacing"], dirichlet=(act, tgt),
                                                face_weights=beta, multigrid={"levels": 4}, tolerance=1e-6)
            if mode == "free":
                ps.set_coarse_space(True, [])
            proj = L.CoarseProjector(ps, net, ratio) if mode == "learned" else ps
            d = proj.project(vel)  # warm-up of the same system
            ps.pressure.fill(0.0)
            F.sync()
            t0 = time.perf_counter()
            d = proj.project(vel)
            F.sync()
            ms = (time.perf_counter() - t0) * 1e3
            # the reference's relative stop: residual <= tol * max(|b|, 1)
            assert d["converged"] and d["residual"] <= 1e-6 * max(d["bNorm"], 1.0), (mode, d)
            row[mode] = (d["iterations"], round(ms, 2))
        rows.append(row)
    print("iterations, ms (cold / free-cell correction / + learned vector):", rows)
