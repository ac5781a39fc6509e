"""Closed-form pins of the oracle's time integration and fixed-point rounding.

* backTrace (advection_solver3.js:20-62, advection_solver2.js:81-127): RK2
  midpoint sub-steps.  Trilinear (bilinear) sampling of a linear velocity
  field v(x) = M x + c is exact away from the clamped border, so one
  sub-step of size s from x0 lands at x0 - d s v(x0 - d s/2 v(x0)), i.e.
  x0 - d s (M x0 + c) + s^2/2 M (M x0 + c) -- a forward-Euler step misses the
  second-order term by s^2/2 |M v|, far above FP32 rounding here.
* The sub-step count max(ceil(|v| T / h), 1): a rigid rotation whose speed
  needs several sub-steps, against the loop restated in float64.
* Face and scalar advection (advection_solver3.js:130-360): the monotone cubic
  of a linear field is exact, so the advected value is the field at the
  closed-form trace.
* P2G fixed point (grid_flip_solver3.js:508-535): WGSL round() is
  round-half-to-even; particles placed so that value * w * 65536 and
  w * 65536 are exact half-integers must round to the even neighbour.
"""
import math

import numpy as np

from oracle import pyoracle as O


def _face_positions(n, h, origin, comp):
    o = O.mac_origin(origin, h, comp)
    fs = O.face_shape(n, comp)
    ii, jj, kk = np.meshgrid(*[np.arange(s) for s in fs], indexing="ij")
    X = np.stack([o[0] + ii * h[0], o[1] + jj * h[1], o[2] + kk * h[2]], axis=-1)
    return X.reshape(-1, 3, order="F").astype(np.float64), fs


def _linear_mac3(n, h, origin, M, c):
    """Face arrays (first index fastest) of v(x) = M x + c."""
    out = []
    for comp in range(3):
        X, _ = _face_positions(n, h, origin, comp)
        out.append((X @ M[comp] + c[comp]).astype(np.float32))
    return out


def _midpoint(x0, M, c, s, d):
    v0 = x0 @ M.T + c
    mid = x0 - d * (s * 0.5) * v0
    return x0 - d * s * (mid @ M.T + c)


def _euler(x0, M, c, s, d):
    return x0 - d * s * (x0 @ M.T + c)


N3 = (16, 16, 16)
H3 = (1.0 / 16,) * 3
M3 = np.array([[0.9, -1.3, 0.4], [1.1, 0.3, -0.8], [-0.5, 0.7, 1.2]])
C3 = np.array([0.2, -0.1, 0.15])


def test_particle_advection_is_one_midpoint_step_of_a_linear_field():
    rng = np.random.default_rng(7)
    u, v, w = _linear_mac3(N3, H3, [0, 0, 0], M3, C3)
    x0 = rng.uniform(0.35, 0.65, (400, 3))
    # |v| <= ~1.1 on this box, so |v| dt < h: one sub-step
    dt = 0.05
    assert np.max(np.linalg.norm(x0 @ M3.T + C3, axis=1)) * dt < H3[0]
    got = O.advect_particles3(x0.astype(np.float32), u, v, w, N3, H3, [0, 0, 0], dt).astype(np.float64)
    want = _midpoint(x0.astype(np.float32).astype(np.float64), M3, C3, dt, -1.0)
    euler = _euler(x0.astype(np.float32).astype(np.float64), M3, C3, dt, -1.0)
    err = np.max(np.abs(got - want))
    assert err < 2e-6, err
    # the second-order term is what the test resolves
    assert np.min(np.max(np.abs(euler - want), axis=1)) > 50 * err


def test_semi_lagrangian_trace_is_one_midpoint_step_backwards():
    rng = np.random.default_rng(8)
    u, v, w = _linear_mac3(N3, H3, [0, 0, 0], M3, C3)
    x0 = rng.uniform(0.35, 0.65, (300, 3)).astype(np.float32)
    got = O.trace3(x0, u, v, w, N3, H3, [0, 0, 0], 0.05, direction=1.0).astype(np.float64)
    want = _midpoint(x0.astype(np.float64), M3, C3, 0.05, 1.0)
    assert np.max(np.abs(got - want)) < 2e-6


def _substep_loop(x0, M, c, T, h, d, max_sub=32):
    """advection_solver3.js:20-62 without a collider, in float64."""
    pt = x0.copy()
    rem = T
    for _ in range(max_sub):
        if rem <= 1e-6:
            break
        v0 = M @ pt + c
        ns = max(math.ceil(np.linalg.norm(v0) * rem / h), 1)
        s = rem / ns
        mid = pt - d * (s * 0.5) * v0
        pt = pt - d * s * (M @ mid + c)
        rem -= s
    return pt


def test_substep_count_on_a_rigid_rotation():
    om = 20.0
    ctr = np.array([0.5, 0.5, 0.5])
    M = np.array([[0.0, -om, 0.0], [om, 0.0, 0.0], [0.0, 0.0, 0.0]])
    c = -M @ ctr
    u, v, w = _linear_mac3(N3, H3, [0, 0, 0], M, c)
    rng = np.random.default_rng(9)
    ang = rng.uniform(0, 2 * np.pi, 64)
    rad = rng.uniform(0.05, 0.12, 64)
    x0 = np.stack([0.5 + rad * np.cos(ang), 0.5 + rad * np.sin(ang), rng.uniform(0.4, 0.6, 64)], axis=1)
    x0 = x0.astype(np.float32)
    T = 0.12  # |v| T / h = 20 r * 0.12 * 16 -> 2..5 sub-steps
    got = O.advect_particles3(x0, u, v, w, N3, H3, [0, 0, 0], T).astype(np.float64)
    subs = set()
    for p in range(64):
        want = _substep_loop(x0[p].astype(np.float64), M, c, T, H3[0], -1.0)
        assert np.max(np.abs(got[p] - want)) < 5e-6, (p, got[p], want)
        subs.add(max(math.ceil(np.linalg.norm(M @ x0[p] + c) * T / H3[0]), 1))
    assert max(subs) >= 3
    # one midpoint step over T instead of the sub-steps lands elsewhere
    one = _midpoint(x0.astype(np.float64), M, c, T, -1.0)
    assert np.max(np.abs(one - got)) > 1e-3


def test_face_advection_of_a_linear_field_samples_the_closed_form_trace():
    n = (24, 24, 24)
    h = (1.0 / 24,) * 3
    M = 0.5 * M3
    u, v, w = _linear_mac3(n, h, [0, 0, 0], M, C3)
    dt = 0.06
    outs = O.advect_face3(u, v, w, n, h, [0, 0, 0], dt)
    for comp in range(3):
        X, fs = _face_positions(n, h, [0, 0, 0], comp)
        inner = np.all((X > 0.3) & (X < 0.7), axis=1)
        xs = X[inner].astype(np.float32).astype(np.float64)
        tr = _midpoint(xs, M, C3, dt, 1.0)
        want = tr @ M[comp] + C3[comp]
        got = outs[comp][inner].astype(np.float64)
        assert np.max(np.abs(got - want)) < 2e-5, comp
        eu = _euler(xs, M, C3, dt, 1.0) @ M[comp] + C3[comp]
        assert np.max(np.abs(eu - want)) > 2e-4


def test_scalar_advection_of_a_linear_field():
    n = (24, 24, 24)
    h = (1.0 / 24,) * 3
    M = 0.5 * M3
    u, v, w = _linear_mac3(n, h, [0, 0, 0], M, C3)
    ii, jj, kk = np.meshgrid(*[np.arange(s) for s in n], indexing="ij")
    X = np.stack([(ii + 0.5) * h[0], (jj + 0.5) * h[1], (kk + 0.5) * h[2]], -1).reshape(-1, 3, order="F")
    a = np.array([0.7, -1.9, 0.45])
    src = (X @ a + 0.3).astype(np.float32)
    dt = 0.06
    out = O.advect_scalar3(src, u, v, w, n, h, [0, 0, 0], dt).astype(np.float64)
    inner = np.all((X > 0.3) & (X < 0.7), axis=1)
    tr = _midpoint(X[inner].astype(np.float32).astype(np.float64), M, C3, dt, 1.0)
    assert np.max(np.abs(out[inner] - (tr @ a + 0.3))) < 2e-5


def test_face_advection_with_zero_velocity_is_the_identity():
    n = (8, 8, 8)
    h = (0.5, 0.5, 0.5)
    rng = np.random.default_rng(3)
    data = [rng.standard_normal(int(np.prod(O.face_shape(n, a)))).astype(np.float32) for a in range(3)]
    zero = [np.zeros_like(d) for d in data]
    for order in (1, 2):
        outs = O.advect_faces3_vd(zero, data, n, h, [0, 0, 0], 0.1, order=order)
        for a in range(3):
            np.testing.assert_array_equal(outs[a], data[a])


# ----------------------------------------------------------------- 2-D
def _linear_mac2(n, h, origin, M, c):
    out = []
    for comp in range(2):
        fs = [n[0] + (comp == 0), n[1] + (comp == 1)]
        o = [origin[0] + (0.0 if comp == 0 else h[0] * 0.5), origin[1] + (0.0 if comp == 1 else h[1] * 0.5)]
        ii, jj = np.meshgrid(np.arange(fs[0]), np.arange(fs[1]), indexing="ij")
        X = np.stack([o[0] + ii * h[0], o[1] + jj * h[1]], -1).reshape(-1, 2, order="F")
        out.append(((X @ M[comp]) + c[comp]).astype(np.float32))
    return out


def test_trace2_is_one_midpoint_step_of_a_linear_field():
    n = (32, 32)
    h = (1.0, 1.0)
    M = np.array([[0.08, -0.11], [0.09, -0.04]])
    c = np.array([0.3, -0.2])
    u, v = _linear_mac2(n, h, [0, 0], M, c)
    rng = np.random.default_rng(11)
    x0 = rng.uniform(10, 22, (200, 2)).astype(np.float32)
    dt = 0.4  # |v| dt < 1 cell
    assert np.max(np.linalg.norm(x0 @ M.T + c, axis=1)) * dt < 1.0
    for d in (1.0, -1.0):
        got = O.trace2(x0, u, v, n, h, [0, 0], dt, direction=d).astype(np.float64)
        want = _midpoint(x0.astype(np.float64), M, c, dt, d)
        err = np.max(np.abs(got - want))
        assert err < 5e-5, err
        assert np.min(np.max(np.abs(_euler(x0.astype(np.float64), M, c, dt, d) - want), axis=1)) > 20 * err


def test_face_advection2_of_a_linear_field():
    n = (32, 32)
    h = (1.0, 1.0)
    M = np.array([[0.08, -0.11], [0.09, -0.04]])
    c = np.array([0.3, -0.2])
    u, v = _linear_mac2(n, h, [0, 0], M, c)
    dt = 0.4
    uo, vo = O.advect_face2(u, v, n, h, [0, 0], dt)
    for comp, got in ((0, uo), (1, vo)):
        fs = [n[0] + (comp == 0), n[1] + (comp == 1)]
        o = [0.0 if comp == 0 else 0.5, 0.0 if comp == 1 else 0.5]
        ii, jj = np.meshgrid(np.arange(fs[0]), np.arange(fs[1]), indexing="ij")
        X = np.stack([o[0] + ii * h[0], o[1] + jj * h[1]], -1).reshape(-1, 2, order="F")
        inner = np.all((X > 8) & (X < 24), axis=1)
        tr = _midpoint(X[inner], M, c, dt, 1.0)
        want = tr @ M[comp] + c[comp]
        assert np.max(np.abs(got[inner] - want)) < 5e-5, comp
        assert np.max(np.abs(_euler(X[inner], M, c, dt, 1.0) @ M[comp] + c[comp] - want)) > 5e-4


# ------------------------------------------------------ P2G tie rounding
def _tie_particles():
    """Particles on the U-face lattice of a 6^3 grid (h = 1, origin 0; U data
    origin (0, 0.5, 0.5)).  Each value particle sits exactly on one face, so
    its only non-zero weight is 1 and value * 65536 is the half-integer given;
    one weight particle sits a fraction f = 2.5 / 65536 past face (1, 1, 1)
    along x, so its two weights scale to 65533.5 and 2.5."""
    halves = [0.5, 1.5, 2.5, 3.5, -0.5, -1.5, -2.5, 7.5]
    pos, vel = [], []
    for q, k in enumerate(halves):
        i, j = 1 + q % 4, 2 + q // 4
        pos.append((float(i), j + 0.5, 3.5))
        vel.append((k / 65536.0, 0.0, 0.0))
    f = 2.5 / 65536.0
    pos.append((1.0 + f, 1.5, 1.5))
    vel.append((0.0, 0.0, 0.0))
    return np.array(pos, np.float32), np.array(vel, np.float32), halves


def test_p2g_fixed_point_rounds_half_to_even():
    pos, vel, halves = _tie_particles()
    n = (6, 6, 6)
    out, valid, numer, denom = O.p2g_fixed3(pos, vel, n, (1.0, 1.0, 1.0), [0, 0, 0], 0)
    fs = O.face_shape(n, 0)

    def fidx(i, j, k):
        return i + fs[0] * (j + fs[1] * k)

    for q, k in enumerate(halves):
        i, j = 1 + q % 4, 2 + q // 4
        f = fidx(i, j, 3)
        assert numer[f] == round(k), (k, numer[f])  # Python's round() is half-to-even
        assert denom[f] == 65536
    # weight ties: 65533.5 -> 65534 and 2.5 -> 2 (half away from zero gives 3)
    assert denom[fidx(1, 1, 1)] == 65534
    assert denom[fidx(2, 1, 1)] == 2
    assert numer.sum() == sum(round(k) for k in halves)


def test_mass_weighted_p2g_rounds_half_to_even():
    """The mass-weighted scatter (grid_flip_solver3.js:493-531) with equal
    densities has mass 1 and must round exactly like the plain one."""
    pos, vel, halves = _tie_particles()
    n = (6, 6, 6)
    conc = np.linspace(0, 1, len(pos)).astype(np.float32)
    a = O.p2g_fixed3(pos, vel, n, (1.0, 1.0, 1.0), [0, 0, 0], 0)
    b = O.p2g_fixed3(pos, vel, n, (1.0, 1.0, 1.0), [0, 0, 0], 0, mass=(conc, 1000.0, 1000.0))
    np.testing.assert_array_equal(a[2], b[2])
    np.testing.assert_array_equal(a[3], b[3])
