"""The brick smoother (FLUX_SMOOTH_BRICK: red-black sweeps brick by brick,
two passes per smoothing, the second on bricks shifted by half a brick) on
the GPU against the oracle's orc_brick_relax V-cycle.

Tolerances as for the other smoothers: one V-cycle relative L2 <= 1e-5
(FP32 storage); PCG solves relative L2 <= 1e-5 with iteration counts within 1
of the oracle's mixed arm.  FP16/BF16 storage: the V-cycle within 1e-3 / 1e-2
of the FP32 oracle, and projections accepted at the same FP64 true residual.
The kernels compute each update as x + omega (b/d - sum (c_n/d) x_n - x) with
the coefficients formed once per pass, the oracle as the reference's
x + ((b - A x)/d) omega: the two round differently, hence tolerances, not
bit equality.
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from tsl_array_n_amd import scenes

from test_gpu_solver import rel, upload, weights_fields, random_case

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _product_bricks():
    O.set_brick(64, 8, 8, 2)
    yield


def _dam_break(F, n):
    sc = scenes.dam_break_system(n)
    sh, h = sc["shape"], sc["spacing"]
    A = O.pressure3_assemble(sh, h, sc["vel"], active=sc["active"], target=sc["target"], beta=sc["beta"],
                             suffix="f32")
    m = upload(F, "u8", sh, A["mask"])
    v = upload(F, "u8", sh, A["vent"])
    w = [upload(F, "float", F.face_shape(sh, a), A["weights"][a]) for a in range(3)]
    return sc, sh, h, A, m, v, w


@pytest.mark.parametrize("nu", [1, 2, 4])
def test_brick_vcycle_dam_break_64(flux, nu):
    """64^3, 4 levels: levels 0-2 smooth brick-wise (shifted bricks partial
    at every face), level 3 (8^3) is the coarsest."""
    F = flux
    sc, sh, h, A, m, v, w = _dam_break(F, 64)
    mg = F.create_multigrid_preconditioner(sh, h, dirichlet_mask=m, coarse_dirichlet_mask=v, face_weights=w,
                                           levels=4, nu_down=nu, nu_up=nu, n_final=nu, smoother="brick")
    ref = O.Multigrid(sh, h, levels=4, nu_down=nu, nu_up=nu, n_final=nu, mask=A["mask"], coarse_mask=A["vent"],
                      weights=A["weights"], suffix="f32", smoother="brick")
    r = np.random.default_rng(nu).standard_normal(int(np.prod(sh))).astype(np.float32)
    z = F.array_n("float", sh)
    mg(upload(F, "float", sh, r), z)
    err = rel(z.to_array(), ref.apply(r))
    print("nu", nu, "rel", err)
    assert err <= 1e-5


@pytest.mark.parametrize("shape,spacing,levels", [((36, 20, 12), (0.5, 1.0, 1.0), 2),
                                                  ((96, 24, 40), (1.0, 1.0, 1.0), 3),
                                                  ((128, 16, 8), (1.0, 0.5, 1.0), 2),
                                                  ((64, 8, 8), (1.0, 1.0, 1.0), 2)])
def test_brick_vcycle_ragged_shapes(flux, shape, spacing, levels):
    """Bricks cut by the grid's end on every axis, a level of exactly one
    brick (64x8x8 is smoothed globally: it fits in one brick), random masks
    and face weights."""
    F = flux
    mask, ws = random_case(shape, 11)
    m = upload(F, "u8", shape, mask)
    w = weights_fields(F, shape, ws)
    mg = F.create_multigrid_preconditioner(shape, spacing, dirichlet_mask=m, face_weights=w, levels=levels,
                                           nu_down=3, nu_up=3, n_final=3, smoother="brick")
    r = np.random.default_rng(12).standard_normal(int(np.prod(shape))).astype(np.float32)
    z = F.array_n("float", shape)
    mg(upload(F, "float", shape, r), z)
    ref = O.Multigrid(shape, spacing, levels=levels, nu_down=3, nu_up=3, n_final=3, mask=mask, weights=ws,
                      suffix="f32", smoother="brick").apply(r)
    err = rel(z.to_array(), ref)
    print(shape, "rel", err)
    assert err <= 1e-5


def test_brick_vcycle_captured_frame(flux, frame110):
    F, c = flux, frame110
    sh = c["shape"]
    m = upload(F, "u8", sh, c["mask"])
    v = upload(F, "u8", sh, c["vent"])
    w = weights_fields(F, sh, c["weights"])
    mg = F.create_multigrid_preconditioner(sh, c["spacing"], dirichlet_mask=m, coarse_dirichlet_mask=v,
                                           face_weights=w, levels=4, smoother="brick")
    ref = O.Multigrid(sh, c["spacing"], levels=4, mask=c["mask"], coarse_mask=c["vent"], weights=c["weights"],
                      suffix="f32", smoother="brick")
    z = F.array_n("float", sh)
    mg(upload(F, "float", sh, c["b"]), z)
    assert rel(z.to_array(), ref.apply(c["b"])) <= 1e-5


def test_brick_vcycle_is_symmetric(flux):
    """u.M v = v.M u on the free cells (FP32 arithmetic, coarsest solved
    to convergence)."""
    F = flux
    sc, sh, h, A, m, v, w = _dam_break(F, 64)
    mg = F.create_multigrid_preconditioner(sh, h, dirichlet_mask=m, coarse_dirichlet_mask=v, face_weights=w,
                                           levels=4, nu_down=3, nu_up=3, n_final=3, n_coarsest=200,
                                           smoother="brick")
    rng = np.random.default_rng(5)
    free = A["mask"] == 0
    u = (rng.standard_normal(free.size) * free).astype(np.float32)
    q = (rng.standard_normal(free.size) * free).astype(np.float32)
    mu, mq = F.array_n("float", sh), F.array_n("float", sh)
    mg(upload(F, "float", sh, u), mu)
    mg(upload(F, "float", sh, q), mq)
    lhs = float(np.dot(q.astype(np.float64), mu.to_array().astype(np.float64)))
    rhs = float(np.dot(u.astype(np.float64), mq.to_array().astype(np.float64)))
    print("u.Mq", rhs, "q.Mu", lhs)
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)
    assert lhs < 0 or rhs < 0


@pytest.mark.parametrize("frame", ["frame110", "frame112"])
def test_brick_pcg_matches_oracle(flux, frame, request):
    F = flux
    c = request.getfixturevalue(frame)
    sh = c["shape"]
    m = upload(F, "u8", sh, c["mask"])
    v = upload(F, "u8", sh, c["vent"])
    w = weights_fields(F, sh, c["weights"])
    op = F.create_laplacian_operator(sh, c["spacing"], m, w)
    mg = F.create_multigrid_preconditioner(sh, c["spacing"], dirichlet_mask=m, coarse_dirichlet_mask=v,
                                           face_weights=w, levels=4, smoother="brick", nu_down=4, nu_up=4,
                                           n_final=4)
    b = upload(F, "float", sh, c["b"])
    x = F.array_n("float", sh)
    s = F.create_pcg_solver(op, "multigrid", b, x, multigrid=mg)
    ok = s.solve(1e-6, 200, relative_tolerance=True)
    ref_mg = O.Multigrid(sh, c["spacing"], levels=4, nu_down=4, nu_up=4, n_final=4, mask=c["mask"],
                         coarse_mask=c["vent"], weights=c["weights"], suffix="f32", smoother="brick")
    xr, _, st = O.pcg_solve_mixed(sh, c["spacing"], c["b"], mask=c["mask"], weights=c["weights"], mg=ref_mg,
                                  tol=1e-6, maxiter=200)
    print("gpu", s.state, "oracle", st)
    assert ok and st["converged"]
    assert abs(s.state["iterations"] - st["iterations"]) <= 1
    assert rel(x.to_array(), xr) <= 1e-5


def _project(F, sc, storage, nu=4, levels=5):
    n_, h = sc["shape"], sc["spacing"]
    vin = [F.array_n("float", F.face_shape(n_, a), sc["vel"][a]) for a in range(3)]
    beta = [F.array_n("float", F.face_shape(n_, a), sc["beta"][a]) for a in range(3)]
    act = F.array_n("u8", n_, sc["active"])
    tgt = F.array_n("float", n_, sc["target"])
    ps = F.create_grid_pressure_solver3(resolution=n_, grid_spacing=h, dirichlet=(act, tgt), face_weights=beta,
                                        multigrid={"levels": levels, "storage": storage, "smoother": "brick",
                                                   "nu_down": nu, "nu_up": nu, "n_final": nu}, tolerance=1e-6)
    d = ps.project(vin)
    return ps, d


def test_brick_projection_dam_break_matches_oracle(flux):
    """A 64^3 dam-break projection with the brick smoother (FP32 storage)
    reaches the FP64 1e-6 true residual and matches the oracle's solve:
    iterations within 1, pressure 1e-5."""
    F = flux
    sc = scenes.dam_break_system(64)
    n_, h = sc["shape"], sc["spacing"]
    ps, d = _project(F, sc, "f32")
    A = O.pressure3_assemble(n_, h, sc["vel"], active=sc["active"], target=sc["target"], beta=sc["beta"],
                             suffix="f32")
    ref_mg = O.Multigrid(n_, h, levels=5, nu_down=4, nu_up=4, n_final=4, mask=A["mask"], coarse_mask=A["vent"],
                         weights=A["weights"], suffix="f32", smoother="brick")
    xr, _, st = O.pcg_solve_mixed(n_, h, A["b"], x0=A["pressure"], mask=A["mask"], weights=A["weights"],
                                  mg=ref_mg, tol=1e-6, maxiter=200)
    print("gpu", d, "oracle", st)
    assert d["converged"] and d["residual"] <= 1e-6 * d["bNorm"]
    assert abs(d["iterations"] - st["iterations"]) <= 1
    assert rel(ps.pressure.to_array(), xr) <= 1e-5


@pytest.mark.parametrize("storage,vtol", [("f16", 1e-3), ("bf16", 1e-2)])
def test_brick_mixed_precision(flux, storage, vtol):
    """FP16/BF16 storage with the brick smoother: one V-cycle within the
    storage precision of the FP32 oracle's (measured 1e-5 / 8e-5 at 64^3),
    and a projection accepted at the same FP64 true residual as the FP32
    form: FP16 in at most 4 more iterations (its rounding moves the count by
    a few either way: 64^3 with 4 sweeps 21 -> 25, 128^3 17 -> 17, 8 sweeps
    16 -> 19 and 12 -> 13; red-black moves alike), BF16 (8-bit mantissa)
    within 1.5x as for red-black (64^3: 21 -> 26); pressure within 1e-4."""
    F = flux
    sc, sh, h, A, m, v, w = _dam_break(F, 64)
    mg = F.create_multigrid_preconditioner(sh, h, dirichlet_mask=m, coarse_dirichlet_mask=v, face_weights=w,
                                           levels=5, nu_down=4, nu_up=4, n_final=4, smoother="brick",
                                           storage=storage)
    r = np.random.default_rng(3).standard_normal(int(np.prod(sh))).astype(np.float32)
    z = F.array_n("float", sh)
    mg(upload(F, "float", sh, r), z)
    ref = O.Multigrid(sh, h, levels=5, nu_down=4, nu_up=4, n_final=4, mask=A["mask"], coarse_mask=A["vent"],
                      weights=A["weights"], suffix="f32", smoother="brick").apply(r)
    err = rel(z.to_array(), ref)
    print(storage, "V-cycle rel vs FP32 oracle", err)
    assert err <= vtol
    p32, d32 = _project(F, sc, "f32")
    pm, dm = _project(F, sc, storage)
    print("f32", d32, storage, dm)
    assert dm["converged"] and dm["residual"] <= 1e-6 * dm["bNorm"]
    if storage == "f16":
        assert dm["iterations"] <= d32["iterations"] + 4
    else:
        assert dm["iterations"] <= int(1.5 * d32["iterations"])
    assert rel(pm.pressure.to_array(), p32.pressure.to_array()) <= 1e-4


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_brick_wrapped_shifted_pass_is_bit_identical(flux, storage, monkeypatch):
    """On a grid of whole bricks the shifted pass runs full bricks that wrap
    around the grid (the brick at the origin takes the last half brick of
    each axis with the first) instead of partial bricks at every face; the
    coefficient across the seam is 0 (a boundary face), so a 128^3 V-cycle
    (levels of 128^3 and 64^3 wrapped, 32^3 partial) agrees bit for bit with
    the partial-brick pass (FLUX_BRICK_WRAP=0), and so does a projection."""
    F = flux
    sc, sh, h, A, m, v, w = _dam_break(F, 128)
    r = np.random.default_rng(23).standard_normal(int(np.prod(sh))).astype(np.float32)
    rd = upload(F, "float", sh, r)
    mg = F.create_multigrid_preconditioner(sh, h, dirichlet_mask=m, coarse_dirichlet_mask=v, face_weights=w,
                                           levels=4, nu_down=3, nu_up=3, n_final=3, smoother="brick",
                                           storage=storage)
    outs = []
    for wrap in ("0", "1"):
        monkeypatch.setenv("FLUX_BRICK_WRAP", wrap)
        z = F.array_n("float", sh)
        mg(rd, z)
        outs.append(z.to_array())
    assert np.array_equal(outs[0], outs[1])
    res = []
    for wrap in ("0", "1"):
        monkeypatch.setenv("FLUX_BRICK_WRAP", wrap)
        ps, d = _project(F, sc, storage)
        res.append((ps.pressure.to_array(), d))
    assert res[0][1]["iterations"] == res[1][1]["iterations"]
    assert np.array_equal(res[0][0], res[1][0])


@pytest.mark.parametrize("storage,tol", [("f32", 1e-6), ("f16", 1e-3)])
def test_brick_zfolded_restriction_matches_stored_residual(flux, storage, tol, monkeypatch):
    """The brick levels' residual folded in z into the restriction (r never
    stored; the restriction sums the z-taps first, then the x-y taps) against
    the stored residual and k_restrict3 (x-y taps first): the same sums in
    another order, so a 96^3 V-cycle agrees to rounding (FP32: measured
    ~1e-7; FP16 storage rounds the coarse right-hand sides, so to its
    precision), and projections take the same iterations within 1."""
    F = flux
    sc, sh, h, A, m, v, w = _dam_break(F, 96)
    r = np.random.default_rng(29).standard_normal(int(np.prod(sh))).astype(np.float32)
    rd = upload(F, "float", sh, r)
    mg = F.create_multigrid_preconditioner(sh, h, dirichlet_mask=m, coarse_dirichlet_mask=v, face_weights=w,
                                           levels=4, nu_down=3, nu_up=3, n_final=3, smoother="brick",
                                           storage=storage)
    outs = []
    for zf in ("0", "1"):
        monkeypatch.setenv("FLUX_ZFOLD", zf)
        z = F.array_n("float", sh)
        mg(rd, z)
        outs.append(z.to_array())
    err = rel(outs[1], outs[0].astype(np.float64))
    print(storage, "folded vs stored", err)
    assert err <= tol
    its = []
    for zf in ("0", "1"):
        monkeypatch.setenv("FLUX_ZFOLD", zf)
        ps, d = _project(F, sc, storage)
        assert d["converged"] and d["residual"] <= 1e-6 * d["bNorm"]
        its.append(d["iterations"])
    assert abs(its[0] - its[1]) <= 1
