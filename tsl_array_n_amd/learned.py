"""Learned initial guess: the U-Net of csrc/flux_unet.hip (host side).

Input channels (built on the GPU from an assembled system): b / rms(b), the
Dirichlet mask, log(mean face weight) / log(density ratio).  Target: the
solution's shape, the pressure on the free cells scaled to unit RMS.  The
solver adds the network output along its energy-optimal step
(pressure += alpha g, alpha = g.(b - A x) / g.(A g)), which sets the
amplitude, then runs its usual PCG, so a learned solve is accepted only at
the same FP64 true-residual tolerance.
"""
import ctypes as C

import numpy as np

from . import FluxError, _check, _lib, array_n

_lib.flux_unet_param_count.restype = C.c_int64

LAYERS = [("enc0a", 3, 16), ("enc0b", 16, 16), ("enc1a", 16, 32), ("enc1b", 32, 32), ("enc2a", 32, 64),
          ("enc2b", 64, 64), ("mid", 64, 64), ("dec2", 128, 64), ("dec1a", 96, 32), ("dec1b", 32, 32),
          ("dec0a", 48, 16), ("dec0b", 16, 16), ("out", 16, 1)]


def param_count():
    return int(_lib.flux_unet_param_count())


def flatten(params):
    """dict name.w [co, ci, 3, 3, 3] / name.b [co] -> flat FP32 vector"""
    return np.concatenate([np.concatenate([np.asarray(params[n + ".w"], np.float32).ravel(),
                                           np.asarray(params[n + ".b"], np.float32).ravel()]) for n, _, _ in LAYERS])


def unflatten(vec):
    out, o = {}, 0
    for n, ci, co in LAYERS:
        k = ci * co * 27
        out[n + ".w"] = vec[o:o + k].reshape(co, ci, 3, 3, 3)
        o += k
        out[n + ".b"] = vec[o:o + co]
        o += co
    return out


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    p = {}
    for n, ci, co in LAYERS:
        bound = np.sqrt(6.0 / (ci * 27))
        p[n + ".w"] = rng.uniform(-bound, bound, (co, ci, 3, 3, 3)).astype(np.float32)
        p[n + ".b"] = np.zeros(co, np.float32)
    # a small output layer: the guess starts near zero
    p["out.w"] *= 0.1
    return p


class UNet:
    def __init__(self, n, training=False, params=None, precision="f32"):
        """precision "f32" (direct FP32 convolutions) or "bf16" (implicit-GEMM
        convolutions on the matrix cores, BF16 activations and gradients with
        FP32 accumulation) for forward/guess and for training, whose master
        weights, gradients and Adam stay FP32"""
        self.n = tuple(n)
        h = C.c_void_p()
        _check(_lib.flux_unet_create(int(n[0]), int(n[1]), int(n[2]), int(bool(training)), C.byref(h)), "unet")
        self.h = h.value
        self.set_precision(precision)
        self.count = param_count()
        self.set_params(init_params() if params is None else params)
        cells = int(np.prod(n))
        self.input = array_n("float", (3 * cells,))
        self.output = array_n("float", (cells,))

    def set_precision(self, precision):
        code = {"f32": 0, "bf16": 1}.get(precision)
        if code is None:
            raise FluxError("unet: precision must be 'f32' or 'bf16', got %r" % (precision,))
        _check(_lib.flux_unet_set_precision(C.c_void_p(self.h), code))
        self.precision = precision

    def set_params(self, params):
        vec = flatten(params) if isinstance(params, dict) else np.ascontiguousarray(params, np.float32)
        _check(_lib.flux_unet_set_params(C.c_void_p(self.h), vec.ctypes.data_as(C.c_void_p), C.c_int64(vec.size)))

    def get(self, which=0):
        out = np.empty(self.count, np.float32)
        _check(_lib.flux_unet_get_params(C.c_void_p(self.h), out.ctypes.data_as(C.c_void_p), C.c_int64(self.count),
                                         int(which)))
        return out

    def forward(self, x=None):
        if x is not None:
            self.input.from_array(x)
        _check(_lib.flux_unet_forward(C.c_void_p(self.h), C.c_void_p(self.input.h), C.c_void_p(self.output.h)))
        return self.output

    def train_step(self, target, x=None, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        if x is not None:
            self.input.from_array(x)
        loss = C.c_double()
        _check(_lib.flux_unet_train_step(C.c_void_p(self.h), C.c_void_p(self.input.h), C.c_void_p(target.h),
                                         C.c_float(lr), C.c_float(beta1), C.c_float(beta2), C.c_float(eps),
                                         C.byref(loss)))
        return loss.value

    def train_step_energy(self, pair, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        """one Adam step on the energy-norm loss of an energy_pairs() pair:
        L = 1 - (g.r)^2 / ((g.B g) e*) with g the output on the free cells,
        B = -A (flux_unet_train_step_energy); returns L"""
        self.input.copy_from(pair["x"])
        loss = C.c_double()
        _check(_lib.flux_unet_train_step_energy(C.c_void_p(self.h), C.c_void_p(self.input.h),
                                                C.c_void_p(pair["op"].h), C.c_void_p(pair["r"].h),
                                                C.c_void_p(pair["mask"].h if pair["mask"] else None),
                                                C.c_double(pair["e_star"]), C.c_float(lr), C.c_float(beta1),
                                                C.c_float(beta2), C.c_float(eps), C.byref(loss)), "train_step_energy")
        return loss.value

    def features(self, b, mask, weights, density_ratio):
        w = (C.c_void_p * 3)(*[f.h for f in weights]) if weights is not None else None
        _check(_lib.flux_unet_features(C.c_void_p(self.h), C.c_void_p(b.h), C.c_void_p(mask.h if mask else None), w,
                                       C.c_float(density_ratio), C.c_void_p(self.input.h)))
        return self.input

    def guess(self, out_field):
        _check(_lib.flux_unet_guess(C.c_void_p(self.h), C.c_void_p(self.input.h), C.c_void_p(out_field.h)))
        return out_field

    def __del__(self):
        try:
            from . import _initialised
            if _initialised and self.h:
                _lib.flux_unet_destroy(C.c_void_p(self.h))
        except Exception:
            pass


class LearnedProjector:
    """project() with the network's guess as the PCG start:
    assemble -> features -> guess -> solve.  The solve is the plain solver's,
    so it is accepted only at the same FP64 true-residual tolerance; a bad
    guess costs iterations, never correctness.

    scaled=True (default) adds the guess to the assembled start along its
    energy-optimal step (pressure += alpha * guess, alpha = g.r / g.Ag), so a
    guess pointing the wrong way is scaled toward zero; scaled=False writes
    the raw guess into the free cells."""

    def __init__(self, solver, net, density_ratio, scaled=True):
        self.solver = solver
        self.net = net
        self.density_ratio = float(density_ratio)
        self.scaled = scaled
        self.alpha = None
        self._g = array_n("float", solver.resolution) if scaled else None

    def project(self, vin, vout=None):
        s = self.solver
        s.assemble(vin)
        self.net.features(s.b, s.dirichlet_mask, s.weights, self.density_ratio)
        if self.scaled:
            self.net.guess(self._g)
            self.alpha = s.scaled_guess(self._g)
        else:
            _check(_lib.flux_unet_guess_into(C.c_void_p(self.net.h), C.c_void_p(self.net.input.h),
                                             C.c_void_p(s.dirichlet_mask.h if s.dirichlet_mask else None),
                                             C.c_void_p(s.pressure.h)), "guess")
        return s.solve(vin, vout)


class CoarseProjector:
    """project() with the network's guess as a coarse vector of the
    preconditioner: assemble -> features -> guess g -> coarse space {free
    cells, g} (flux_pressure3_set_coarse_space) -> the plain solve, accepted
    at the same FP64 true-residual tolerance.  With start=True the guess is
    also the PCG start along its energy-optimal step."""

    def __init__(self, solver, net, density_ratio, start=False):
        self.solver = solver
        self.net = net
        self.density_ratio = float(density_ratio)
        self.start = start
        self.g = array_n("float", solver.resolution)
        solver.set_coarse_space(True, [self.g])

    def project(self, vin, vout=None):
        s = self.solver
        s.assemble(vin)
        self.net.features(s.b, s.dirichlet_mask, s.weights, self.density_ratio)
        self.net.guess(self.g)
        if self.start:
            s.scaled_guess(self.g)
        return s.solve(vin, vout)


def scene_for(n, q, rng):
    """training / evaluation scenes: dam-break variants (every third) and
    random two-phase scenes"""
    from . import scenes
    if q % 3 == 2:
        return scenes.dam_break_system(n, box_max=(rng.uniform(0.2, 0.7), rng.uniform(0.3, 0.8), 1.0)), 1000.0
    sc = scenes.random_scene(n, int(rng.integers(0, 2 ** 31)))
    return sc, sc["ratio"]


def energy_pairs(F, n, count, seed=0, levels=None):
    """Training pairs for the energy-norm loss, produced by the MGPCG solver
    itself: the input features, the start's residual r (b on the free
    cells), the Dirichlet mask, the system's operator and e* = x*.(-r) of
    the converged solution x*; everything stays on the GPU."""
    levels = levels or max(2, int(np.log2(n)) - 2)
    rng = np.random.default_rng(seed)
    pairs = []
    net = UNet((n, n, n), training=False)
    q = 0
    while len(pairs) < count and q < 4 * count:
        sc, ratio = scene_for(n, q, rng)
        q += 1
        shape = sc["shape"]
        vel = [array_n("float", F.face_shape(shape, a), sc["vel"][a]) for a in range(3)]
        beta = [array_n("float", F.face_shape(shape, a), sc["beta"][a]) for a in range(3)]
        act = array_n("u8", shape, sc["active"])
        tgt = array_n("float", shape, sc["target"])
        ps = F.create_grid_pressure_solver3(resolution=shape, grid_spacing=sc["spacing"], dirichlet=(act, tgt),
                                            face_weights=beta, multigrid={"levels": levels}, tolerance=1e-6)
        d = ps.project(vel)
        if not d["converged"]:
            continue
        mask = ps.dirichlet_mask.to_array() if ps.dirichlet_mask else np.zeros(n ** 3, np.uint8)
        free = mask == 0
        b = ps.b.to_array().astype(np.float64)
        x = ps.pressure.to_array().astype(np.float64)
        e_star = -float(np.dot(x[free], b[free]))
        if not e_star > 0:
            continue
        net.features(ps.b, ps.dirichlet_mask, ps.weights, ratio)
        feats = array_n("float", (3 * n ** 3,), net.input.to_array())
        mfield = array_n("u8", shape, mask)
        r = array_n("float", shape, np.where(free, b, 0.0).astype(np.float32))
        wf = [array_n("float", F.face_shape(shape, a), ps.weights[a].to_array()) for a in range(3)]
        op = F.create_laplacian_operator(shape, sc["spacing"], mfield, wf)
        pairs.append({"x": feats, "r": r, "mask": mfield, "op": op, "e_star": e_star, "keep": (wf,),
                      "ratio": ratio})
    return pairs


def train_energy(F, n, pairs, steps, lr=1e-3, seed=0, log=None):
    net = UNet((n, n, n), training=True, params=init_params(seed))
    losses = []
    for s in range(steps):
        losses.append(net.train_step_energy(pairs[s % len(pairs)], lr=lr))
        if log and (s % 100 == 0 or s == steps - 1):
            log("step %d energy loss %.4e" % (s, losses[-1]))
    return net, losses


def training_pairs(F, n, count, seed=0, levels=None):
    """(input features, normalized pressure) pairs produced by the MGPCG
    solver itself on randomized two-phase scenes (random dense blobs, density
    ratio log-uniform in [1, 1000], random divergence) and dam-break variants;
    each pair stays on the GPU."""
    from . import scenes
    levels = levels or max(2, int(np.log2(n)) - 2)
    rng = np.random.default_rng(seed)
    pairs = []
    net = UNet((n, n, n), training=False)
    for q in range(count):
        if q % 3 == 2:
            sc = scenes.dam_break_system(n, box_max=(rng.uniform(0.2, 0.7), rng.uniform(0.3, 0.8), 1.0))
            ratio = 1000.0
        else:
            sc = scenes.random_scene(n, int(rng.integers(0, 2 ** 31)))
            ratio = sc["ratio"]
        shape = sc["shape"]
        vel = [array_n("float", F.face_shape(shape, a), sc["vel"][a]) for a in range(3)]
        beta = [array_n("float", F.face_shape(shape, a), sc["beta"][a]) for a in range(3)]
        act = array_n("u8", shape, sc["active"])
        tgt = array_n("float", shape, sc["target"])
        ps = F.create_grid_pressure_solver3(resolution=shape, grid_spacing=sc["spacing"], dirichlet=(act, tgt),
                                            face_weights=beta, multigrid={"levels": levels}, tolerance=1e-6)
        d = ps.project(vel)
        if not d["converged"]:
            continue
        net.features(ps.b, ps.dirichlet_mask, ps.weights, ratio)
        x = array_n("float", (3 * n ** 3,), net.input.to_array())
        # the target is the solution's shape (unit RMS over the free cells);
        # the amplitude is set at solve time by the energy-optimal step
        free = ps.dirichlet_mask.to_array() == 0 if ps.dirichlet_mask else np.ones(n ** 3, bool)
        p = np.where(free, ps.pressure.to_array().astype(np.float64), 0.0)
        rms = np.sqrt((p * p).sum() / max(int(free.sum()), 1))
        if not rms > 0:
            continue
        t = array_n("float", (n ** 3,), (p / rms).astype(np.float32))
        pairs.append((x, t))
    return pairs


def train(F, n, pairs, steps, lr=1e-3, seed=0, log=None):
    net = UNet((n, n, n), training=True, params=init_params(seed))
    losses = []
    for s in range(steps):
        x, t = pairs[s % len(pairs)]
        net.input.copy_from(x)
        losses.append(net.train_step(t, lr=lr))
        if log and (s % 50 == 0 or s == steps - 1):
            log("step %d loss %.4e" % (s, losses[-1]))
    return net, losses


def evaluate(F, train_size=64, pairs=24, steps=600, size=256, reps=3, levels=6, storage="f16", log=print):
    """Train on MGPCG pairs at train_size with the energy-norm loss, then
    solve the size^3 two-phase dam break (the benchmark's system) with the
    network's guess as a coarse vector (BF16 inference at size^3) against the
    cold start and the free-cell correction alone; returns a table of
    iterations and ms per projection (inference included)."""
    import time
    from . import scenes
    t0 = time.perf_counter()
    data = energy_pairs(F, train_size, pairs, seed=11)
    net, losses = train_energy(F, train_size, data, steps, lr=1e-3, log=log)
    log("trained on %d pairs at %d^3, %d steps, %.1f s; energy loss %.3f -> %.3f"
        % (len(data), train_size, steps, time.perf_counter() - t0, float(np.mean(losses[:len(data)])),
           float(np.mean(losses[-len(data):]))))
    params = net.get()
    del data
    big = UNet((size,) * 3, precision="bf16", params=params)
    sc = scenes.dam_break_system(size)
    shape = sc["shape"]
    beta = [array_n("float", F.face_shape(shape, a), sc["beta"][a]) for a in range(3)]
    act = array_n("u8", shape, sc["active"])
    tgt = array_n("float", shape, sc["target"])
    vin = [array_n("float", F.face_shape(shape, a), sc["vel"][a]) for a in range(3)]
    vout = [array_n("float", F.face_shape(shape, a)) for a in range(3)]
    table = {}
    for mode in ("cold", "free", "learned", "learned+start"):
        ps = F.create_grid_pressure_solver3(resolution=shape, grid_spacing=sc["spacing"], dirichlet=(act, tgt),
                                            face_weights=beta, multigrid={"levels": levels, "storage": storage},
                                            tolerance=1e-6)
        if mode == "free":
            ps.set_coarse_space(True, [])
        proj = CoarseProjector(ps, big, 1000.0, start=mode.endswith("start")) if mode.startswith("learned") else ps
        ps.pressure.fill(0.0)
        d = proj.project(vin, vout)
        times, its = [], []
        for _ in range(reps):
            ps.pressure.fill(0.0)
            F.sync()
            t = time.perf_counter()
            d = proj.project(vin, vout)
            F.sync()
            times.append((time.perf_counter() - t) * 1e3)
            its.append(d["iterations"])
        table[mode] = {"iterations": int(np.median(its)), "ms": round(float(np.median(times)), 2),
                       "converged": d["converged"], "fp64_relative_residual": d["residual"] / d["bNorm"]}
        log("%s: %s" % (mode, table[mode]))
        del ps, proj
    return table


def evaluate_headline(F, train_size=64, pairs=48, steps=3000, size=256, reps=5, log=print):
    """The learned coarse vector against the benchmark's headline form: train
    a BF16 network on the matrix cores (energy-norm loss, pairs from MGPCG
    solves at train_size), then project the system of step 3 of the size^3
    two-phase FLIP dam break with the headline form (FP16 V-cycle, brick
    smoother with 12 sweeps, 6 levels) and the free-cell correction alone,
    with the network's guess as a second coarse vector (BF16 inference at
    size^3 included in the time; also as the start, and with the flexible
    Polak-Ribiere beta), and with the exact solution in its place
    (the bound of any vector); returns iterations and ms per projection."""
    import time
    from . import flip as FL
    t0 = time.perf_counter()
    data = energy_pairs(F, train_size, pairs, seed=11)
    net = UNet((train_size,) * 3, training=True, precision="bf16")
    losses = []
    for s_ in range(steps):
        losses.append(net.train_step_energy(data[s_ % len(data)], lr=1e-3))
        if (s_ + 1) % 500 == 0:
            log("step %d: mean energy loss %.3f" % (s_ + 1, float(np.mean(losses[-len(data):]))))
    log("trained (BF16 on MFMA) on %d pairs at %d^3, %d steps, %.1f s; energy loss %.3f -> %.3f"
        % (len(data), train_size, steps, time.perf_counter() - t0, float(np.mean(losses[:len(data)])),
           float(np.mean(losses[-len(data):]))))
    params = net.get()
    del data, net
    big = UNet((size,) * 3, precision="bf16", params=params)
    fs, _ = FL.two_phase_dam_break(size, levels=6)
    fs.capture_projection_input(True)
    for _ in range(3):
        fs.on_advance_time_step()
    sc = fs.captured_system()
    del fs
    shape = sc["shape"]
    beta = [array_n("float", F.face_shape(shape, a), sc["beta"][a]) for a in range(3)]
    act = array_n("u8", shape, sc["active"])
    tgt = array_n("float", shape, sc["target"])
    vin = [array_n("float", F.face_shape(shape, a), sc["vel"][a]) for a in range(3)]
    vout = [array_n("float", F.face_shape(shape, a)) for a in range(3)]
    mg = {"levels": 6, "storage": "f16", "smoother": "brick", "nu_down": 12, "nu_up": 12, "n_final": 12}
    table, exact = {}, None
    for mode in ("free", "learned", "learned+start", "learned+flexible", "exact"):
        ps = F.create_grid_pressure_solver3(resolution=shape, grid_spacing=sc["spacing"], dirichlet=(act, tgt),
                                            face_weights=beta, multigrid=mg, tolerance=1e-6)
        if mode == "free":
            ps.set_coarse_space(True, [])
            proj = ps
        elif mode.startswith("learned"):
            proj = CoarseProjector(ps, big, 1000.0, start=mode.endswith("start"))
            if mode.endswith("flexible"):
                ps.set_flexible(True)
        else:
            ps.set_coarse_space(True, [array_n("float", shape, exact)])
            proj = ps
        ps.pressure.fill(0.0)
        d = proj.project(vin, vout)
        if mode == "free":
            exact = ps.pressure.to_array()
        times, its = [], []
        for _ in range(reps):
            ps.pressure.fill(0.0)
            F.sync()
            t = time.perf_counter()
            d = proj.project(vin, vout)
            F.sync()
            times.append((time.perf_counter() - t) * 1e3)
            its.append(d["iterations"])
        table[mode] = {"iterations": int(np.median(its)), "ms": round(float(np.median(times)), 2),
                       "converged": d["converged"], "stopped_by": d["stoppedBy"],
                       "fp64_relative_residual": d["residual"] / d["bNorm"]}
        log("%s: %s" % (mode, table[mode]))
        del ps, proj
    return table


def main(argv=None):
    """time the network's forward: python -m tsl_array_n_amd.learned
    --size 128 --precision bf16 --reps 20 (random He-uniform weights, random
    input); prints one line with the time per forward and the FLOP rate.
    With --evaluate: train with the energy-norm loss at --train-size and
    report the learned coarse vector at --size (evaluate())."""
    import argparse
    import json
    from . import init, timer_start, timer_stop
    ap = argparse.ArgumentParser()
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--precision", default="bf16", choices=["f32", "bf16"])
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--evaluate", action="store_true")
    ap.add_argument("--train-size", type=int, default=64)
    ap.add_argument("--pairs", type=int, default=24)
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--headline", action="store_true",
                    help="with --evaluate: against the benchmark's headline form (evaluate_headline)")
    a = ap.parse_args(argv)
    init(0)
    if a.evaluate and a.headline:
        import tsl_array_n_amd as F
        table = evaluate_headline(F, a.train_size, a.pairs, a.steps, a.size)
        print(json.dumps(table), flush=True)
        return
    if a.evaluate:
        import tsl_array_n_amd as F
        table = evaluate(F, a.train_size, a.pairs, a.steps, a.size, reps=3, levels=max(2, int(np.log2(a.size)) - 2))
        print(json.dumps(table), flush=True)
        return
    n = (a.size,) * 3
    net = UNet(n, precision=a.precision)
    rng = np.random.default_rng(0)
    net.forward(rng.standard_normal(3 * a.size ** 3).astype(np.float32))
    timer_start()
    for _ in range(a.reps):
        net.forward()
    ms = timer_stop() / a.reps
    levels = {"enc0a": 0, "enc0b": 0, "enc1a": 1, "enc1b": 1, "enc2a": 2, "enc2b": 2, "mid": 3, "dec2": 2,
              "dec1a": 1, "dec1b": 1, "dec0a": 0, "dec0b": 0, "out": 0}
    flops = sum(2 * 27 * ci * co * (a.size >> levels[name]) ** 3 for name, ci, co in LAYERS)
    print("unet forward %d^3 %s: %.3f ms, %.1f TFLOP/s" % (a.size, a.precision, ms, flops / ms * 1e-9))


if __name__ == "__main__":
    main()
