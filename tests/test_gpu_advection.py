"""createSemiLagrangianAdvectionSolver3 on the GPU against the oracle
(advection_solver3.js:67-362): faces and cell-centred scalars advected
through a separate velocity grid, order 1 and MacCormack, bit for bit; dt
changed between calls; the checks on aliasing and order."""
import numpy as np
import pytest

from oracle import pyoracle as O
from tsl_array_n_amd import transfer as T

pytestmark = pytest.mark.gpu


def smooth_mac(n, seed, amp=1.0, shift=0.2):
    rng = np.random.default_rng(seed)
    out = []
    for a in range(3):
        s = O.face_shape(n, a)
        i, j, k = np.meshgrid(*[np.arange(x, dtype=np.float64) for x in s], indexing="ij")
        ph, f = rng.uniform(0, 6.28, 3), rng.uniform(0.15, 0.45, 3)
        v = amp * np.sin(f[0] * i + ph[0]) * np.cos(f[1] * j + ph[1]) * np.cos(f[2] * k + ph[2]) + shift
        out.append(v.ravel(order="F").astype(np.float32))
    return out


def features(shape, seed):
    rng = np.random.default_rng(seed)
    i, j, k = np.meshgrid(*[np.arange(x, dtype=np.float64) for x in shape], indexing="ij")
    v = np.sin(0.5 * i) * np.cos(0.35 * j + 0.2 * k) + (i + 0.7 * j > shape[0] * 0.6) * 2.0
    return (v + rng.uniform(-0.05, 0.05, v.shape)).ravel(order="F").astype(np.float32)


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("n,dt", [((20, 16, 12), 1.3), ((33, 17, 24), 3.1)])
def test_faces_match_oracle(flux, order, n, dt):
    F = flux
    g = T.Grid(n, (1.0, 1.0, 1.0))
    V = smooth_mac(n, 1, 1.2)
    D = [features(O.face_shape(n, a), 10 + a) for a in range(3)]
    vel = T.mac_fields(g, V)
    din = T.mac_fields(g, D)
    out = T.mac_fields(g)
    adv = T.SemiLagrangianAdvectionSolver3(g, vel, dt, order=order)
    adv.advect_face_centered3(din, out)
    want = O.advect_faces3_vd(V, D, n, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), dt, order=order)
    for a in range(3):
        np.testing.assert_array_equal(out[a].to_array(), want[a])
    # self-advection, as the grid solver runs it (velocityGrid = input)
    self_adv = T.SemiLagrangianAdvectionSolver3(g, vel, dt, order=order)
    self_adv.advect_face_centered3(vel, out)
    want = O.advect_face3(*V, n, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), dt, order=order)
    for a in range(3):
        np.testing.assert_array_equal(out[a].to_array(), want[a])
    if order == 1:  # the stand-alone order-1 entry point agrees
        out2 = T.mac_fields(g)
        T.advect_faces(g, vel, out2, dt)
        for a in range(3):
            np.testing.assert_array_equal(out2[a].to_array(), want[a])


@pytest.mark.parametrize("order", [1, 2])
def test_scalar_matches_oracle_and_dt_is_live(flux, order):
    F = flux
    n, h, o = (24, 20, 18), (0.5, 0.5, 0.5), (0.25, -1.0, 0.0)
    g = T.Grid(n, h, o)
    V = smooth_mac(n, 2, 0.8)
    src = features(n, 3)
    vel = T.mac_fields(g, V)
    a_in = F.array_n("float", n, src)
    a_out = F.array_n("float", n)
    adv = T.SemiLagrangianAdvectionSolver3(g, vel, 0.9, order=order)
    for dt in (0.9, 2.2):
        adv.set_dt(dt)
        adv.advect_scalar3(a_in, a_out)
        want = O.advect_scalar3(src, *V, n, h, o, dt, order=order)
        np.testing.assert_array_equal(a_out.to_array(), want)


def test_rejects_aliasing_and_bad_order(flux):
    import ctypes as C
    from tsl_array_n_amd import FluxError, _lib
    n = (8, 8, 8)
    g = T.Grid(n, (1.0, 1.0, 1.0))
    vel = T.mac_fields(g, smooth_mac(n, 4))
    h = C.c_void_p()
    assert _lib.flux_advection3_create(C.byref(g.c), T._mac(vel), C.c_float(0.5), 32, 3, C.byref(h)) != 0
    assert "order" in _lib.flux_last_error().decode()
    for order in (1, 2):
        adv = T.SemiLagrangianAdvectionSolver3(g, vel, 0.5, order=order)
        other = T.mac_fields(g)
        with pytest.raises(FluxError, match="alias"):
            adv.advect_face_centered3(other, vel)  # output = the bound velocity
        with pytest.raises(FluxError, match="alias"):
            adv.advect_face_centered3(other, other)
        c = flux.array_n("float", n)
        with pytest.raises(FluxError, match="alias"):
            adv.advect_scalar3(c, c)


def _ball_sdf(n, c, r):
    from tsl_array_n_amd import boundary as B
    return O.collider_sdf3([B.sphere(c, r)], n, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0))


@pytest.mark.parametrize("order", [1, 2])
def test_faces_and_scalar_through_a_collider_match_oracle(flux, order):
    """the collider's SDF bound to the solver: traces stop at the ball,
    points inside it keep the output's previous values, order 2 clamps over
    the corners outside it; bit for bit against the oracle"""
    F = flux
    n, dt = (24, 18, 16), 2.2
    g = T.Grid(n, (1.0, 1.0, 1.0))
    V = smooth_mac(n, 4, 1.5, shift=0.6)
    D = [features(O.face_shape(n, a), 20 + a) for a in range(3)]
    prev = [np.full(int(np.prod(O.face_shape(n, a))), -3.0, np.float32) for a in range(3)]
    sdf = _ball_sdf(n, (12.0, 9.0, 8.0), 4.3)
    vel = T.mac_fields(g, V)
    din = T.mac_fields(g, D)
    out = T.mac_fields(g, prev)
    col = F.array_n("float", n, sdf)
    adv = T.SemiLagrangianAdvectionSolver3(g, vel, dt, order=order, collider=col)
    adv.advect_face_centered3(din, out)
    want = O.advect_faces3_c(V, D, n, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), sdf, dt, order=order, out=prev)
    for a in range(3):
        np.testing.assert_array_equal(out[a].to_array(), want[a])
    assert np.count_nonzero(want[0] == -3.0) > 0  # some faces lie in the ball
    src = features(n, 30)
    sin = F.array_n("float", n, src)
    sout = F.array_n("float", n, np.full(int(np.prod(n)), 9.0, np.float32))
    adv.advect_scalar3(sin, sout)
    want = O.advect_scalar3_c(src, *V, n, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), sdf, dt, order=order,
                              out=np.full(int(np.prod(n)), 9.0, np.float32))
    np.testing.assert_array_equal(sout.to_array(), want)


def test_particles_through_a_collider_match_oracle(flux):
    F = flux
    n, dt = (20, 16, 16), 3.0
    g = T.Grid(n, (1.0, 1.0, 1.0))
    V = smooth_mac(n, 6, 1.0, shift=1.0)
    sdf = _ball_sdf(n, (12.0, 8.0, 8.0), 3.5)
    rng = np.random.default_rng(8)
    pos = rng.uniform(1.0, 15.0, (4000, 3)).astype(np.float32).ravel()
    P = F.array_n("vec3", 4000, pos)
    vel = T.mac_fields(g, V)
    T.advect_particles_collider(g, P, 4000, vel, F.array_n("float", n, sdf), dt)
    want = O.advect_particles3_c(pos, *V, n, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), sdf, dt)
    got = P.to_array()
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-5)
    free = O.advect_particles3(pos, *V, n, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), dt)
    assert np.count_nonzero(np.abs(free - want) > 1e-3) > 0  # some traces stopped at the ball
