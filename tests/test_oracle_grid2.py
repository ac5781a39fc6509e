"""The 2-D Eulerian-step oracle pinned against the reference's own known
answers and closed forms (CPU only).

Polygon SDF: polygon_sdf.test.js:8-40 (square and union values) and a
circle 48-gon against the analytic distance; computeAdaptiveSubSteps:
cfl.test.js:11-78; bilinear and monotone cubic sampling reproduce linear
data; face advection leaves a shear flow unchanged; constrainVelocity's
walls, inflows, clamp and no-flux projection at a grid-aligned box; the
outflow condition's relaxation formula and its fixed point; the pressure
system's divergence and Dirichlet terms; a projection leaves a
divergence-free field; the uniform stream is a fixed point of a whole step.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as O
from tsl_array_n_amd import scenes

SQUARE = [(0, 0), (4, 0), (4, 4), (0, 4)]


def test_polygon_distance_known_answers():
    # polygon_sdf.test.js:8-40
    assert O.polygons_distance2([SQUARE], 2, 2) == pytest.approx(-2, abs=1e-6)
    assert O.polygons_distance2([SQUARE], 2, -1) == pytest.approx(1, abs=1e-6)
    assert O.polygons_distance2([SQUARE], 5, 2) == pytest.approx(1, abs=1e-6)
    sq_b = [(10, 0), (14, 0), (14, 4), (10, 4)]
    assert O.polygons_distance2([SQUARE, sq_b], 2, 2) == pytest.approx(-2, abs=1e-6)
    assert O.polygons_distance2([SQUARE, sq_b], 7, 2) == pytest.approx(3, abs=1e-6)
    # corner region: Euclidean distance to the vertex
    assert O.polygons_distance2([SQUARE], 7, 8) == pytest.approx(5.0, abs=1e-12)


def test_polygon_sdf_of_a_circle_48gon():
    n, h, o = (40, 30), (1.0, 1.0), (0.0, 0.0)
    ring = scenes.circle_polygon(17.3, 14.6, 8.0, 48)
    sdf = O.polygon_sdf2([ring], n, h, o)
    # inscribed 48-gon: |sdf - (|x - c| - r)| <= r (1 - cos(pi/48)) near the boundary
    sag = 8.0 * (1 - math.cos(math.pi / 48))
    for j in range(n[1]):
        for i in range(n[0]):
            x, y = i + 0.5, j + 0.5
            exact = math.hypot(x - 17.3, y - 14.6) - 8.0
            got = float(sdf[i + n[0] * j])
            assert abs(got - exact) <= sag + 1e-5 or abs(exact) > 8.0 * math.sin(math.pi / 48) + 1.0
            assert (got < 0) == (exact < -sag) or abs(exact) <= sag
    inv = O.polygon_sdf2([ring], n, h, o, invert=True)
    np.testing.assert_array_equal(inv, -sdf)


def test_adaptive_substeps_known_answers():
    # cfl.test.js:11-78
    k, dt = O.adaptive_substeps(0.1, 1, 1 / 30)
    assert k == 1 and dt == pytest.approx(1 / 30)
    k, dt = O.adaptive_substeps(0.0, 1, 1 / 30)
    assert k == 1
    k, dt = O.adaptive_substeps(100, 1, 1 / 30, courant=1)
    assert k == 4 and dt == pytest.approx(1 / 120)
    k, dt = O.adaptive_substeps(50, 0.5, 1 / 24, courant=2)
    assert k * dt == pytest.approx(1 / 24)
    assert O.adaptive_substeps(1e6, 1, 1 / 30, courant=1, max_sub=32)[0] == 32
    assert O.adaptive_substeps(5, 1, 1 / 30)[0] == 1
    assert O.adaptive_substeps(5, 1, 1 / 30, courant=0.1)[0] > 1


def test_max_abs_is_the_largest_magnitude():
    a = np.array([0.5, -3.25, 2.0], np.float32)
    b = np.array([1.0, -0.0], np.float32)
    assert O.max_abs([a, b]) == 3.25
    assert O.max_abs([np.zeros(4, np.float32)]) == 0.0


def test_sampling_reproduces_linear_data():
    s = (9, 7)
    o, h = (0.25, -0.5), (0.5, 0.75)
    ii, jj = np.meshgrid(np.arange(s[0]), np.arange(s[1]), indexing="xy")
    x = o[0] + h[0] * ii
    y = o[1] + h[1] * jj
    d = (0.3 + 1.5 * x - 0.75 * y).astype(np.float32).ravel()
    rng = np.random.default_rng(1)
    for _ in range(200):
        px = rng.uniform(o[0] + h[0], o[0] + h[0] * (s[0] - 2))
        py = rng.uniform(o[1] + h[1], o[1] + h[1] * (s[1] - 2))
        exact = 0.3 + 1.5 * px - 0.75 * py
        assert abs(O.sample_bilinear2(d, s, o, h, px, py) - exact) < 2e-5
        assert abs(O.sample_cubic2(d, s, o, h, px, py) - exact) < 2e-5
    # at a node both return the stored value exactly
    assert O.sample_bilinear2(d, s, o, h, o[0] + 3 * h[0], o[1] + 2 * h[1]) == d[3 + s[0] * 2]
    assert O.sample_cubic2(d, s, o, h, o[0] + 3 * h[0], o[1] + 2 * h[1]) == d[3 + s[0] * 2]


def test_advection_leaves_a_shear_flow_unchanged():
    n, h, o = (16, 12), (1.0, 1.0), (0.0, 0.0)
    su = O.face_shape2(n, 0)
    u = np.array([[0.2 + 0.1 * j for i in range(su[0])] for j in range(su[1])], np.float32).ravel()
    v = np.zeros(n[0] * (n[1] + 1), np.float32)
    uo, vo = O.advect_face2(u, v, n, h, o, 0.7)
    np.testing.assert_array_equal(uo, u)
    np.testing.assert_array_equal(vo, v)


def test_constrain_walls_inflow_and_clamp():
    n, h, o = (10, 8), (1.0, 1.0), (0.0, 0.0)
    rng = np.random.default_rng(2)
    u = rng.uniform(-2, 2, (n[0] + 1) * n[1]).astype(np.float32)
    v = rng.uniform(-2, 2, n[0] * (n[1] + 1)).astype(np.float32)
    u[5] = 250.0
    v[7] = np.nan
    strip = O.polygon_sdf2([scenes.wall_strip_polygon(2.0, -1000.0, n[1])], n, h, o)
    uo, vo = O.constrain_velocity2([u, v], n, h, o, closed_flags=4 | 8, inflows=[(strip, (1.5, 0.25), "set")])
    U = uo.reshape(n[1], n[0] + 1)
    V = vo.reshape(n[1] + 1, n[0])
    # inflow strip x < 2: U faces at x = 0, 1 and V faces at x = 0.5, 1.5
    assert np.all(U[:, :2] == np.float32(1.5)) and np.all(V[1:-1, :2] == np.float32(0.25))
    # closed bottom / top zero the V boundary rows (then the inflow sets x < 2)
    assert np.all(V[0, 2:] == 0) and np.all(V[-1, 2:] == 0)
    assert U.max() <= 100 and np.isfinite(vo).all()
    # untouched faces keep their values
    keep = np.ones_like(U, bool)
    keep[:, :2] = False
    ur = u.reshape(n[1], n[0] + 1)
    np.testing.assert_array_equal(U[keep], np.clip(ur[keep], -100, 100))
    # 'add' adds to the existing value
    uo2, _ = O.constrain_velocity2([u, v], n, h, o, closed_flags=0, inflows=[(strip, (1.5, 0.0), "add")])
    np.testing.assert_array_equal(uo2.reshape(n[1], n[0] + 1)[:, :2], (ur[:, :2] + np.float32(1.5)).astype(np.float32))


def test_constrain_no_flux_at_a_grid_aligned_box():
    """a box collider [8, 12] x [4, 10]: every face between a solid and a
    fluid cell carries zero velocity; inside the box next to its vertical
    sides, where the SDF gradient is (+-1, 0), the normal component U is 0."""
    n, h, o = (20, 14), (1.0, 1.0), (0.0, 0.0)
    box = [(8, 4), (12, 4), (12, 10), (8, 10)]
    sdf = O.polygon_sdf2([box], n, h, o)
    rng = np.random.default_rng(4)
    u = rng.uniform(-1, 1, (n[0] + 1) * n[1]).astype(np.float32)
    v = rng.uniform(-1, 1, n[0] * (n[1] + 1)).astype(np.float32)
    uo, vo = O.constrain_velocity2([u, v], n, h, o, closed_flags=15, sdf=sdf)
    U = uo.reshape(n[1], n[0] + 1)
    V = vo.reshape(n[1] + 1, n[0])
    S = sdf.reshape(n[1], n[0]) < 0
    for j in range(n[1]):
        for i in range(n[0]):
            if not S[j, i]:
                continue
            if i > 0 and not S[j, i - 1]:
                assert U[j, i] == 0
            if i < n[0] - 1 and not S[j, i + 1]:
                assert U[j, i + 1] == 0
            if j > 0 and not S[j - 1, i]:
                assert V[j, i] == 0
            if j < n[1] - 1 and not S[j + 1, i]:
                assert V[j + 1, i] == 0
    # U faces at x = 9 (inside, one cell from the left side), rows well inside
    for j in range(6, 8):
        assert U[j, 9] == 0
    assert np.all(U[:, 0] == 0) and np.all(U[:, -1] == 0) and np.all(V[0] == 0) and np.all(V[-1] == 0)


def test_outflow_relaxation_and_fixed_point():
    n, h, o = (12, 6), (1.0, 1.0), (0.0, 0.0)
    slab = O.polygon_sdf2([scenes.wall_strip_polygon(n[0] - 2.0, n[0] + 1000.0, n[1])], n, h, o)
    su = O.face_shape2(n, 0)
    u = np.full(su[0] * su[1], 0.8, np.float32)
    v = np.zeros(n[0] * (n[1] + 1), np.float32)
    uo, vo = O.outflow_bc2([u, v], n, h, o, 0.5, slab)
    np.testing.assert_array_equal(uo, u)
    np.testing.assert_array_equal(vo, v)
    # a step: a inside the slab (x > 10), b upstream; rows uniform so samples are exact
    a, b, dt = np.float32(0.4), np.float32(1.2), np.float32(0.5)
    U = np.full((su[1], su[0]), b, np.float32)
    U[:, 11:] = a
    uo, _ = O.outflow_bc2([U.ravel(), v], n, h, o, dt, slab)
    Uo = uo.reshape(su[1], su[0])
    # face x = 11: bulk u = a, outward normal (1, 0), upstream sample at x = 10 reads b
    courant = np.float32(min(max((a * dt) / np.float32(1.0), 0.0), 1.0))
    want = np.float32(a + (b - a) * courant)
    assert np.all(Uo[:, 11] == want)
    assert np.all(Uo[:, :11] == b)


def test_pressure_system_divergence_and_dirichlet_terms():
    n, h = (6, 5), (0.5, 0.25)
    su = O.face_shape2(n, 0)
    xs = np.arange(su[0]) * h[0]
    u = np.tile(xs, su[1]).astype(np.float32)            # du/dx = 1
    v = np.zeros(n[0] * (n[1] + 1), np.float32)
    A = O.pressure2_assemble(n, h, [u, v])
    np.testing.assert_allclose(A["b"], 1.0, rtol=0, atol=1e-6)
    active = np.zeros(n[0] * n[1], np.uint8)
    target = np.zeros(n[0] * n[1], np.float32)
    active[n[0] - 1::n[0]] = 1                            # right column pinned
    target[n[0] - 1::n[0]] = 0.125
    A = O.pressure2_assemble(n, h, [u, v], active=active, target=target)
    b = A["b"].reshape(n[1], n[0])
    assert np.all(b[:, -1] == -0.125) and np.all(A["pressure"].reshape(n[1], n[0])[:, -1] == 0.125)
    np.testing.assert_allclose(b[:, -2], 1.0 - 0.125 / h[0] ** 2, atol=1e-6)
    np.testing.assert_allclose(b[:, :-2], 1.0, atol=1e-6)


def test_projection_leaves_a_divergence_free_field():
    n, h = (32, 16), (1.0, 1.0)
    rng = np.random.default_rng(5)
    u = rng.uniform(-1, 1, (n[0] + 1) * n[1]).astype(np.float32)
    v = rng.uniform(-1, 1, n[0] * (n[1] + 1)).astype(np.float32)
    # closed box: zero normal velocity on the walls
    U = u.reshape(n[1], n[0] + 1)
    U[:, 0] = U[:, -1] = 0
    V = v.reshape(n[1] + 1, n[0])
    V[0] = V[-1] = 0
    A = O.pressure2_assemble(n, h, [u, v])
    A["b"] -= A["b"].mean()                                # the compatible (zero-mean) system
    mg = O.Multigrid(n, h, levels=3, suffix="f32")
    x, _, d = O.pcg_solve_mixed(n, h, A["b"], mg=mg, tol=1e-6, maxiter=200)
    assert d["converged"]
    uo, vo = O.pressure2_correct(n, h, x, [u, v])
    div = O.pressure2_assemble(n, h, [uo, vo])["b"]
    assert np.abs(div - div.mean()).max() < 1e-4 * np.abs(A["b"]).max()


def test_uniform_stream_is_a_fixed_point_of_the_step():
    n, h, o = (24, 12), (1.0, 1.0), (0.0, 0.0)
    inflow = O.polygon_sdf2([scenes.wall_strip_polygon(2.0, -1000.0, n[1])], n, h, o)
    outflow = O.polygon_sdf2([scenes.wall_strip_polygon(n[0] - 2.0, n[0] + 1000.0, n[1])], n, h, o)
    su = O.face_shape2(n, 0)
    st = {"u": np.ones(su[0] * su[1], np.float32), "v": np.zeros(n[0] * (n[1] + 1), np.float32),
          "pressure": np.zeros(n[0] * n[1], np.float32), "p_accepted": np.zeros(n[0] * n[1], np.float32)}
    for _ in range(2):
        d = O.solver2_step(st, n, h, o, 0.25, inflows=[(inflow, (1.0, 0.0), "set")], outflow_sdf=outflow,
                           flags=12, levels=2)
        assert d["converged"] and not d["rejected"]
    np.testing.assert_array_equal(st["u"], 1.0)
    np.testing.assert_array_equal(st["v"], 0.0)
    np.testing.assert_array_equal(st["pressure"], 0.0)
