"""The oracle's 3-D collider paths (CPU): the back-trace's stop at a static
collider (advection_solver3.js:41-52), advection that skips points inside the
collider, the particle push-out (grid_flip_solver3.js:945-964), checked on
closed forms; and an SDF that is positive everywhere reproduces the
collider-free routines bit for bit."""
import numpy as np

from oracle import pyoracle as O


def _uniform(n, vx=1.0):
    u = np.full(int(np.prod(O.face_shape(n, 0))), vx, np.float32)
    v = np.zeros(int(np.prod(O.face_shape(n, 1))), np.float32)
    w = np.zeros(int(np.prod(O.face_shape(n, 2))), np.float32)
    return u, v, w


def _plane_sdf(n, x0):
    """phi = x0 - x at the cell centres (unit cells): the collider fills
    x > x0; linear, so trilinear samples are exact inside the grid"""
    k, j, i = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    return (np.float32(x0) - (i + 0.5).astype(np.float32)).ravel().astype(np.float32)


def test_trace_stops_at_the_crossing_of_a_planar_collider():
    """u = 1, a particle at x = 4 traced forward over dt = 2 (two unit
    sub-steps): the first ends at x = 5 (phi 0.3), the second would end at
    x = 6 (phi -0.7) and stops at the SDF-weighted point 5 * 0.7 + 6 * 0.3 =
    5.3, the plane itself; without the collider the trace reaches x = 6"""
    n, h, o = (12, 6, 6), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)
    u, v, w = _uniform(n)
    sdf = _plane_sdf(n, 5.3)
    p = np.array([4.0, 3.0, 3.0], np.float32)
    free = O.trace3_c(p, u, v, w, n, h, o, None, 2.0, direction=-1.0)
    hit = O.trace3_c(p, u, v, w, n, h, o, sdf, 2.0, direction=-1.0)
    assert abs(free[0] - 6.0) <= 1e-5 and abs(hit[0] - 5.3) <= 1e-5
    assert hit[1] == 3.0 and hit[2] == 3.0
    moved = O.advect_particles3_c(p, u, v, w, n, h, o, sdf, 2.0)
    np.testing.assert_array_equal(moved, hit)


def test_positive_sdf_is_the_collider_free_path():
    n, h, o = (9, 7, 8), (0.5, 0.5, 0.5), (0.1, -0.2, 0.0)
    rng = np.random.default_rng(3)
    vel = [rng.standard_normal(int(np.prod(O.face_shape(n, a)))).astype(np.float32) for a in range(3)]
    data = [rng.standard_normal(int(np.prod(O.face_shape(n, a)))).astype(np.float32) for a in range(3)]
    src = rng.standard_normal(int(np.prod(n))).astype(np.float32)
    far = np.full(int(np.prod(n)), 5.0, np.float32)
    for order in (1, 2):
        a = O.advect_faces3_vd(vel, data, n, h, o, 0.3, order=order)
        b = O.advect_faces3_c(vel, data, n, h, o, far, 0.3, order=order)
        for q in range(3):
            np.testing.assert_array_equal(a[q], b[q])
        np.testing.assert_array_equal(O.advect_scalar3(src, *vel, n, h, o, 0.3, order=order),
                                      O.advect_scalar3_c(src, *vel, n, h, o, far, 0.3, order=order))
    pos = (rng.uniform(0.2, 3.8, (50, 3))).astype(np.float32).ravel()
    np.testing.assert_array_equal(O.advect_particles3(pos, *vel, n, h, o, 0.3),
                                  O.advect_particles3_c(pos, *vel, n, h, o, far, 0.3))


def test_points_inside_the_collider_keep_the_output():
    """faces whose positions lie in the collider keep the output's values
    (order 1 and 2); faces far from it, whose traces never reach it, get the
    collider-free result"""
    n, h, o = (16, 6, 6), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)
    u, v, w = _uniform(n, 1.0)
    sdf = _plane_sdf(n, 10.3)
    rng = np.random.default_rng(5)
    data = [rng.standard_normal(int(np.prod(O.face_shape(n, a)))).astype(np.float32) for a in range(3)]
    sentinel = [np.full_like(d, 777.0) for d in data]
    for order in (1, 2):
        got = O.advect_faces3_c([u, v, w], data, n, h, o, sdf, 0.5, order=order, out=sentinel)
        ref = O.advect_faces3_vd([u, v, w], data, n, h, o, 0.5, order=order)
        for a in range(3):
            fs = O.face_shape(n, a)
            x = np.tile(np.arange(fs[0]) + (0.0 if a == 0 else 0.5), fs[1] * fs[2])
            inside = x > 10.3
            assert np.all(got[a][inside] == 777.0)
            far = x < 7.0
            np.testing.assert_array_equal(got[a][far], ref[a][far])


def test_push_out_moves_particles_to_the_surface():
    """phi = 5.3 - x: a particle at x = 6 (phi -0.7) moves along the
    gradient (-1, 0, 0) by (0 - phi) + shift; one outside stays"""
    n, h, o = (12, 6, 6), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)
    sdf = _plane_sdf(n, 5.3)
    pos = np.array([6.0, 3.0, 3.0, 4.0, 2.5, 2.5], np.float32)
    out = O.push_out3(pos, sdf, n, h, o, 0.0, 0.0)
    assert abs(out[0] - 5.3) <= 1e-5 and out[1] == 3.0 and out[2] == 3.0
    np.testing.assert_array_equal(out[3:], pos[3:])
    out = O.push_out3(pos, sdf, n, h, o, 0.0, 0.2)
    assert abs(out[0] - 5.1) <= 1e-5
