"""The FLIP-stage oracle pinned against closed forms (CPU only).

fractionInsideCpu against exact polygon clipping for linear level sets;
primitive distances against their textbook forms and rotation invariance;
computeFaceWeights on grid-aligned and half-space geometry; the Jacobi
extrapolation's propagation front; fluid marking on known cells; and
constrainVelocity's walls, clamp and no-flux projection against a planar
wall, where the projected velocity has a closed form.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as O


def clip_area(phi):
    """area of {phi < 0} in the unit square for phi(x, y) = a + b x + c y
    (exact: Sutherland-Hodgman clip of the square by a half-plane)"""
    poly = [(0.0, 0.0), (1.0, 0.0), (1.0, 1.0), (0.0, 1.0)]
    out = []
    for q in range(4):
        p0, p1 = poly[q], poly[(q + 1) % 4]
        f0, f1 = phi(*p0), phi(*p1)
        if f0 < 0:
            out.append(p0)
        if (f0 < 0) != (f1 < 0):
            t = f0 / (f0 - f1)
            out.append((p0[0] + t * (p1[0] - p0[0]), p0[1] + t * (p1[1] - p0[1])))
    if len(out) < 3:
        return 0.0
    return 0.5 * abs(sum(out[q][0] * out[(q + 1) % len(out)][1] - out[(q + 1) % len(out)][0] * out[q][1]
                         for q in range(len(out))))


def test_fraction_inside_is_exact_for_linear_level_sets():
    rng = np.random.default_rng(3)
    for _ in range(400):
        a, b, c = rng.uniform(-1, 1, 3)
        phi = lambda x, y: a + b * x + c * y  # noqa: E731
        got = O.fraction_inside(phi(0, 0), phi(1, 0), phi(0, 1), phi(1, 1))
        assert got == pytest.approx(clip_area(phi), abs=1e-12)
    assert O.fraction_inside(-1, -1, -1, -1) == 1.0
    assert O.fraction_inside(1, 2, 3, 4) == 0.0


def test_primitive_distances():
    s = [O_sphere((1.0, 2.0, 3.0), 0.5)]
    assert O.primitives_distance3(s, 1.0, 2.0, 5.0) == pytest.approx(1.5)
    b = [{"type": "box", "center": (0.0, 0.0, 0.0), "halfExtents": (1.0, 2.0, 3.0)}]
    assert O.primitives_distance3(b, 2.0, 3.0, 0.0) == pytest.approx(math.hypot(1.0, 1.0))   # outside an edge
    assert O.primitives_distance3(b, 0.5, 0.0, 0.0) == pytest.approx(-0.5)                   # inside: nearest face
    assert O.primitives_distance3(b + s, 1.0, 2.0, 3.0) == pytest.approx(-0.5)               # union = min
    # a box rotated about z by angle: distance at R(angle) p equals the unrotated box's at p
    rb = [{"type": "box", "center": (0.0, 0.0, 0.0), "halfExtents": (1.0, 2.0, 3.0),
           "rotation": {"axis": (0.0, 0.0, 2.0), "angle": 0.7}}]
    rng = np.random.default_rng(5)
    for _ in range(50):
        p = rng.uniform(-4, 4, 3)
        ca, sa = math.cos(0.7), math.sin(0.7)
        q = (ca * p[0] - sa * p[1], sa * p[0] + ca * p[1], p[2])
        assert O.primitives_distance3(rb, *q) == pytest.approx(O.primitives_distance3(b, *p), abs=1e-12)


def O_sphere(center, radius):
    return {"type": "sphere", "center": center, "radius": radius}


def test_face_weights_grid_aligned_box_and_half_space():
    n, h, o = (8, 8, 8), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)
    # box exactly covering cells [2,5] on every axis: faces on or inside it close
    box = [{"type": "box", "center": (4.0, 4.0, 4.0), "halfExtents": (2.0, 2.0, 2.0)}]
    wu, wv, ww = O.face_weights3(box, n, h, o)
    U = wu.reshape(8, 8, 9)  # [k][j][i]
    assert U[3, 3, 2] == 0.0 and U[3, 3, 6] == 0.0 and U[3, 3, 4] == 0.0   # on the surface / inside
    assert U[3, 3, 1] == 1.0 and U[3, 3, 7] == 1.0 and U[0, 0, 4] == 1.0   # clear of it
    # an oblique plane x + 0.5 y < 4.3 (the +x face of a huge rotated box):
    # every partial face carries the exact open fraction of its square
    nrm = np.array([1.0, 0.5, 0.0]) / math.hypot(1.0, 0.5)
    big = 1000.0
    ang = math.atan2(nrm[1], nrm[0])
    p0 = np.array([4.3, 0.0, 4.0])  # on the plane
    ctr = p0 - nrm * big
    plane_box = [{"type": "box", "center": tuple(ctr), "halfExtents": (big, big, big),
                  "rotation": {"axis": (0.0, 0.0, 1.0), "angle": ang}}]
    wu, wv, ww = O.face_weights3(plane_box, n, h, o)
    U = wu.reshape(8, 8, 9)
    for (k, j, i) in [(4, 2, 3), (4, 3, 3), (4, 5, 2), (1, 6, 1)]:
        x = float(i)
        phi = lambda yy, zz: (x + 0.5 * (j + yy) - 4.3) / math.hypot(1, 0.5)  # noqa: E731
        frac = clip_area(lambda s, t: phi(s, t))
        want = min(1.0, max(0.0, 1.0 - frac))
        if 0 < want < 0.01:
            want = 0.01
        assert U[k, j, i] == pytest.approx(want, abs=2e-7)


def test_face_weight_floor():
    # a sphere covering all but a sliver of one face's square: open but < 1%
    n, h, o = (4, 4, 4), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)
    big = [{"type": "box", "center": (2.0 - 1000.0 + 0.995, 2.0, 2.0), "halfExtents": (1000.0, 1000.0, 1000.0)}]
    wu, wv, ww = O.face_weights3(big, n, h, o)
    V = wv.reshape(4, 5, 4)
    # the surface is the plane x = 2.995: V faces at x-centre 2.5 (square
    # spanning x in [2, 3]) are 0.5% open and get the 0.01 floor; those at
    # 1.5 are closed and those at 3.5 fully open
    assert V[1, 1, 2] == pytest.approx(0.01)
    assert V[1, 1, 1] == 0.0 and V[1, 1, 3] == 1.0


def test_extrapolation_front_and_means():
    n = (7, 1, 1)
    val = np.zeros(7, np.float32)
    valid = np.zeros(7, np.int32)
    val[3], valid[3] = 2.0, 1
    for it in range(5):
        out = O.extrapolate3(n, val, valid, it)
        for i in range(7):
            assert out[i] == (2.0 if abs(i - 3) <= it else 0.0)
    # Jacobi: a cell between two valid ones takes their mean in one step
    val = np.array([1.0, 9.0, 3.0], np.float32)
    out = O.extrapolate3((3, 1, 1), val, np.array([1, 0, 1], np.int32), 1)
    assert out.tolist() == [1.0, 2.0, 3.0]
    # all valid: identity
    rnd = np.random.default_rng(1).standard_normal(60).astype(np.float32)
    np.testing.assert_array_equal(O.extrapolate3((3, 4, 5), rnd, np.ones(60, np.int32), 4), rnd)


def test_mark_fluid_cells():
    n, h, o = (4, 3, 2), (0.5, 0.5, 0.5), (1.0, 0.0, 0.0)
    pos = np.array([[1.1, 0.1, 0.1], [2.9, 1.4, 0.9], [0.0, -5.0, 7.0]], np.float32)
    m = O.mark_fluid3(pos.ravel(), n, h, o).reshape(2, 3, 4)
    want = np.zeros((2, 3, 4), np.float32)
    want[0, 0, 0] = 1   # first particle
    want[1, 2, 3] = 1   # second
    want[1, 0, 0] = 1   # third clamps to (0, 0, 1)
    np.testing.assert_array_equal(m, want)
    adj = O.fluid_adjacent3(m.ravel(), n, 0).reshape(2, 3, 5)
    assert adj[0, 0, 0] == 1 and adj[0, 0, 1] == 1 and adj[0, 0, 2] == 0   # faces of cell (0,0,0), clamped at i=0
    assert adj[1, 2, 4] == 1 and adj[1, 2, 3] == 1 and adj[1, 2, 2] == 0


def test_constrain_walls_and_clamp():
    n, h, o = (4, 5, 6), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)
    rng = np.random.default_rng(2)
    vel = [rng.uniform(-50, 50, int(np.prod(O.face_shape(n, a)))).astype(np.float32) for a in range(3)]
    vel[0][7] = 150.0
    vel[1][3] = np.nan
    vel[2][5] = -np.inf
    out = O.constrain_velocity3(vel, n, h, o, closed_flags=1 | 8 | 32)
    U = out[0].reshape(6, 5, 5)
    V = out[1].reshape(6, 6, 4)
    W = out[2].reshape(7, 5, 4)
    assert np.all(U[:, :, 0] == 0) and not np.all(U[:, :, 4] == 0)      # left closed, right open
    assert np.all(V[:, 5, :] == 0) and not np.all(V[:, 0, :] == 0)      # up closed, down open
    assert np.all(W[6, :, :] == 0) and not np.all(W[0, :, :] == 0)      # front closed, back open
    assert out[0][7] == 100.0 and out[1][3] == 0.0 and out[2][5] == 0.0   # clamp; non-finite -> 0
    assert np.all(np.abs(np.concatenate(out)) <= 100)


def test_constrain_inflows_set_add_and_order():
    """inflows after the walls, in order, by the sign of the SDF sampled at
    each face: half-spaces phi = x - 3.25 ('set') and phi = 1.75 - y
    ('add'), linear so the trilinear samples are exact; faces at x < 3.25
    take the set velocity, faces at y > 1.75 add theirs, and where both
    hold the later inflow wins ('add' after 'set' adds to the set value,
    'set' after 'add' overwrites it)"""
    n, h, o = (6, 5, 4), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)
    rng = np.random.default_rng(4)
    vel = [rng.uniform(-5, 5, int(np.prod(O.face_shape(n, a)))).astype(np.float32) for a in range(3)]
    k, j, i = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    phi_set = ((i + 0.5) - 3.25).astype(np.float32).ravel()
    phi_add = (1.75 - (j + 0.5)).astype(np.float32).ravel()
    vs, va = (2.0, -1.0, 0.5), (0.25, 3.0, -0.75)
    for order in ("set-add", "add-set"):
        inflows = [(phi_set, vs, "set"), (phi_add, va, "add")]
        if order == "add-set":
            inflows = inflows[::-1]
        out = O.constrain_velocity3(vel, n, h, o, closed_flags=0, inflows=inflows)
        for a in range(3):
            fs = O.face_shape(n, a)
            kk, jj, ii = np.meshgrid(np.arange(fs[2]), np.arange(fs[1]), np.arange(fs[0]), indexing="ij")
            x = ii + (0.0 if a == 0 else 0.5)
            y = jj + (0.0 if a == 1 else 0.5)
            in_set = (x < 3.25).ravel()
            in_add = (y > 1.75).ravel()
            want = vel[a].copy()
            for name in order.split("-"):
                if name == "set":
                    want = np.where(in_set, np.float32(vs[a]), want)
                else:
                    want = np.where(in_add, want + np.float32(va[a]), want)
            np.testing.assert_array_equal(out[a], want.astype(np.float32), err_msg="%s axis %d" % (order, a))


@pytest.mark.parametrize("use_weights", [False, True])
def test_constrain_no_flux_against_a_planar_wall(use_weights):
    """A solid half-space x >= 8 and a uniform velocity (a, b, c): blocked faces
    near the wall get the tangential velocity (0, b, c) scaled by the friction
    factor max(1 - max(a, 0) mu / |(b, c)|, 0)."""
    n, h, o = (16, 16, 16), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)
    big = 1000.0
    wall = [{"type": "box", "center": (8.0 + big, 8.0, 8.0), "halfExtents": (big, big, big)}]
    sdf = O.collider_sdf3(wall, n, h, o)
    wts = O.face_weights3(wall, n, h, o) if use_weights else None
    a, b, c, mu = 0.5, 0.25, -0.125, 0.5
    vel = [np.full(int(np.prod(O.face_shape(n, ax))), val, np.float32) for ax, val in enumerate((a, b, c))]
    out = O.constrain_velocity3(vel, n, h, o, closed_flags=0, sdf=sdf, weights=wts, friction=mu, depth=5)
    scale = max(1.0 - max(a, 0.0) * mu / math.hypot(b, c), 0.0)
    U = out[0].reshape(16, 16, 17)
    V = out[1].reshape(16, 17, 16)
    W = out[2].reshape(17, 16, 16)
    sl = (slice(3, 13), slice(3, 13))
    # blocked faces one to three cells inside the wall
    for i in (9, 10, 11):
        assert np.all(U[sl[0], sl[1], i] == 0.0)
    for i in (8, 9, 10):
        np.testing.assert_allclose(V[3:13, 3:13, i], b * scale, rtol=1e-6)
        np.testing.assert_allclose(W[3:13, 3:13, i], c * scale, rtol=1e-6)
    # the open side keeps the uniform flow; the face on the wall is closed
    assert np.all(U[sl[0], sl[1], 4] == a) and np.all(V[3:13, 3:13, 4] == b)
    assert np.all(U[sl[0], sl[1], 8] == 0.0)


def test_resample_moves_lowest_donors_to_needing_cells():
    """A 3x3x1 slab of cells: the centre-left cell holds 9 particles (> 8, all
    donors), its neighbours hold 4 each, the centre cell is empty with 3
    neighbours holding >= 4: it receives the 4 lowest-index donors, placed at
    its centre with the grid velocity there."""
    n, h, o = (3, 3, 1), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)
    rng = np.random.default_rng(0)
    pos = []

    def put(i, j, k):
        pos.append([i + rng.uniform(0.1, 0.9), j + rng.uniform(0.1, 0.9), k + rng.uniform(0.1, 0.9)])

    for _ in range(4):
        put(1, 0, 0)
        put(1, 2, 0)
    donors_at = len(pos)
    for _ in range(9):
        put(0, 1, 0)
    for _ in range(4):
        put(2, 1, 0)
    pos = np.array(pos, np.float32)
    vel = np.zeros_like(pos)
    u = np.full(int(np.prod(O.face_shape(n, 0))), 0.5, np.float32)
    v = np.full(int(np.prod(O.face_shape(n, 1))), -0.25, np.float32)
    w = np.full(int(np.prod(O.face_shape(n, 2))), 2.0, np.float32)
    p2, v2 = O.resample3(pos.ravel(), vel.ravel(), u, v, w, n, h, o)
    p2 = p2.reshape(-1, 3)
    v2 = v2.reshape(-1, 3)
    moved = [q for q in range(len(pos)) if not np.array_equal(p2[q], pos[q])]
    assert moved == list(range(donors_at, donors_at + 4))
    for q in moved:
        assert p2[q].tolist() == [1.5, 1.5, 0.5] and v2[q].tolist() == [0.5, -0.25, 2.0]


def layer_state(n, h, height_cells, ppa=2, block_x=None):
    pos = []
    nx = n[0] if block_x is None else block_x
    for k in range(n[2]):
        for j in range(height_cells):
            for i in range(nx):
                for sk in range(ppa):
                    for sj in range(ppa):
                        for si in range(ppa):
                            pos.append([(i + (si + 0.5) / ppa) * h[0], (j + (sj + 0.5) / ppa) * h[1],
                                        (k + (sk + 0.5) / ppa) * h[2]])
    pos = np.array(pos, np.float32).ravel()
    st = {"pos": pos, "vel": np.zeros_like(pos), "pressure": np.zeros(int(np.prod(n)), np.float32),
          "p_accepted": np.zeros(int(np.prod(n)), np.float32)}
    for a, c in enumerate("uvw"):
        st[c] = np.zeros(int(np.prod(O.face_shape(n, a))), np.float32)
    return st


def test_flip_step_keeps_a_layer_at_rest():
    """Reduced pressure: a liquid layer at rest in a closed box is in
    hydrostatic balance; one step leaves velocities and particles still."""
    n, h = (8, 8, 8), (1 / 8, 1 / 8, 1 / 8)
    st = layer_state(n, h, 4)
    p0 = st["pos"].copy()
    diag = O.flip3_step(st, n, h, (0, 0, 0), 1 / 60, levels=2, damping=0.0)
    assert diag["converged"] and not diag["rejected"]
    assert np.max(np.abs(st["vel"])) < 1e-5
    assert np.max(np.abs(st["pos"] - p0)) < 1e-6


def test_flip_step_dam_break_starts_to_collapse():
    n, h = (8, 8, 8), (1 / 8, 1 / 8, 1 / 8)
    st = layer_state(n, h, 6, block_x=3)
    count = st["pos"].size // 3
    for _ in range(3):
        diag = O.flip3_step(st, n, h, (0, 0, 0), 1 / 60, levels=2)
        assert diag["converged"]
    p = st["pos"].reshape(-1, 3)
    v = st["vel"].reshape(-1, 3)
    assert p.shape[0] == count and np.all(p >= 0) and np.all(p <= 1)
    top = p[:, 1] > 0.6
    assert np.mean(v[top, 1]) < -0.05            # the top of the column falls
    front = p[:, 0] > 0.3
    assert np.mean(v[front, 0]) > 0.05           # the front runs out along +x


def smoke_state(n):
    cells = int(np.prod(n))
    st = {"dens": np.zeros(cells, np.float32), "temp": np.zeros(cells, np.float32),
          "pressure": np.zeros(cells, np.float32), "p_accepted": np.zeros(cells, np.float32)}
    for a, c in enumerate("uvw"):
        st[c] = np.zeros(int(np.prod(O.face_shape(n, a))), np.float32)
    return st


def test_smoke_uniform_buoyancy_is_a_pure_gradient():
    """A uniform temperature gives a uniform vertical force: a gradient, which
    the projection removes entirely in a closed box; the scalars only decay."""
    n, h = (12, 12, 12), (1.0, 1.0, 1.0)
    st = smoke_state(n)
    st["temp"][:] = 0.5
    st["dens"][:] = 0.25
    d = O.smoke3_step(st, n, h, (0, 0, 0), 0.1, levels=2)
    assert d["converged"]
    assert max(np.max(np.abs(st[c])) for c in "uvw") < 1e-5
    np.testing.assert_allclose(st["temp"], 0.5 * (1 - 0.001), rtol=1e-6)
    np.testing.assert_allclose(st["dens"], 0.25 * (1 - 0.001), rtol=1e-6)


def test_smoke_hot_blob_rises():
    n, h = (16, 16, 16), (1.0, 1.0, 1.0)
    st = smoke_state(n)
    st["temp"] = O.source_sphere3(st["temp"], n, h, (0, 0, 0), (8.0, 4.0, 8.0), 2.5, 1.0)
    y = np.repeat(np.arange(16), 16)[None, :].repeat(16, 0).ravel() + 0.5
    cy0 = np.sum(st["temp"] * y) / np.sum(st["temp"])
    for _ in range(4):
        assert O.smoke3_step(st, n, h, (0, 0, 0), 0.25, b=5.0, levels=2)["converged"]
    cy1 = np.sum(st["temp"] * y) / np.sum(st["temp"])
    V = st["v"].reshape(16, 17, 16)
    assert cy1 > cy0 + 0.05            # the centre of heat moves up
    assert V[8, 5, 8] > 0.1            # and the velocity above the source points up
