"""FLIP grid stages on the GPU against the oracle (GPU tests).

Extrapolation, fluid marking and adjacency are integer/select work and must
match bit for bit; the collider face weights are computed in FP64 on both
sides (|diff| <= 1e-6 after the f32 store); constrainVelocity is f32 work in
the same operation order (bit-exact expected, checked to 1e-6 relative).
"""
import numpy as np
import pytest

import tsl_array_n_amd as F
from oracle import pyoracle as O
from tsl_array_n_amd import boundary as B
from tsl_array_n_amd.transfer import Grid

pytestmark = pytest.mark.gpu


def shapes_case():
    return [B.sphere((5.3, 6.1, 4.7), 2.6),
            B.box((11.0, 5.0, 9.0), (2.0, 3.5, 1.5), {"axis": (0.3, 1.0, 0.2), "angle": 0.6})]


@pytest.mark.parametrize("iterations", [0, 1, 4, 5])
@pytest.mark.parametrize("valid_type", ["int", "u8"])
def test_extrapolation_matches_oracle(flux, iterations, valid_type):
    n = (13, 9, 7)
    rng = np.random.default_rng(iterations)
    val = rng.standard_normal(int(np.prod(n))).astype(np.float32)
    valid = (rng.random(val.size) < 0.15).astype(np.int32)
    v = F.array_n("float", n, val)
    m = F.array_n(valid_type, n, valid.astype(np.uint8) if valid_type == "u8" else valid)
    out = F.array_n("float", n)
    ex = B.Extrapolator(v, m, out)
    ex(iterations)
    np.testing.assert_array_equal(out.to_array(), O.extrapolate3(n, val, valid, iterations))
    B.Extrapolator(v, m)(iterations)  # in place
    np.testing.assert_array_equal(v.to_array(), O.extrapolate3(n, val, valid, iterations))


def test_fluid_mask_and_adjacency_match_oracle(flux):
    n, h, o = (20, 12, 10), (0.1, 0.1, 0.1), (0.3, -0.2, 0.0)
    rng = np.random.default_rng(4)
    pos = (np.array(o) + rng.uniform(-0.1, 1.0, (5000, 3)) * np.array(n) * 0.1 * np.array([0.6, 0.9, 1.0])).astype(
        np.float32)
    g = Grid(n, h, o)
    p = F.array_n("vec3", 5000, pos.ravel())
    mask = F.array_n("float", n)
    B.mark_fluid_cells(g, p, 5000, mask)
    want = O.mark_fluid3(pos.ravel(), n, h, o)
    np.testing.assert_array_equal(mask.to_array(), want)
    for a in range(3):
        adj = F.array_n("int", F.face_shape(n, a))
        B.fluid_adjacent(g, mask, a, adj)
        np.testing.assert_array_equal(adj.to_array(), O.fluid_adjacent3(want, n, a))


@pytest.mark.parametrize("invert", [False, True])
def test_collider_geometry_matches_oracle(flux, invert):
    n, h, o = (16, 12, 14), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)
    sh = shapes_case()
    ws = [F.array_n("float", F.face_shape(n, a)) for a in range(3)]
    B.collider_face_weights(n, h, o, sh, ws, invert=invert)
    want = O.face_weights3(sh, n, h, o, invert=invert)
    for a in range(3):
        got = ws[a].to_array()
        assert np.max(np.abs(got - want[a])) <= 1e-6
        assert 0 < np.count_nonzero(got < 1) < got.size
    sdf = F.array_n("float", n)
    B.collider_sdf(n, h, o, sh, sdf, invert=invert)
    np.testing.assert_allclose(sdf.to_array(), O.collider_sdf3(sh, n, h, o, invert=invert), atol=1e-6)


@pytest.mark.parametrize("collider", ["none", "sdf", "weights"])
def test_constrain_velocity_matches_oracle(flux, collider):
    n, h, o = (16, 12, 14), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)
    rng = np.random.default_rng(9)
    vel = [rng.uniform(-3, 3, int(np.prod(F.face_shape(n, a)))).astype(np.float32) for a in range(3)]
    vel[0][11] = 250.0
    vel[1][17] = np.nan
    g = Grid(n, h, o)
    fields = [F.array_n("float", F.face_shape(n, a), vel[a]) for a in range(3)]
    sh = shapes_case()
    sdf_h = sdf = wts = wts_h = None
    if collider != "none":
        sdf_h = O.collider_sdf3(sh, n, h, o)
        sdf = F.array_n("float", n, sdf_h)
    if collider == "weights":
        wts_h = O.face_weights3(sh, n, h, o)
        wts = [F.array_n("float", F.face_shape(n, a), wts_h[a]) for a in range(3)]
    bs = B.BoundarySolver3(g, fields, sdf=sdf, face_weights=wts, friction=0.5)
    bs.closed_domain_boundary_flag = B.DIRECTION_LEFT | B.DIRECTION_DOWN | B.DIRECTION_FRONT
    bs.constrain_velocity(5)
    want = O.constrain_velocity3(vel, n, h, o, closed_flags=1 | 4 | 32, sdf=sdf_h, weights=wts_h, friction=0.5,
                                 depth=5)
    for a in range(3):
        np.testing.assert_allclose(fields[a].to_array(), want[a], rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("collider", ["none", "sdf"])
def test_constrain_velocity_with_inflows_matches_oracle(flux, collider):
    """createSDFInflow3 objects in constrainVelocity: a 'set' ball and an
    'add' box, applied after the walls and before the clamp, in order (the
    ball's set region overlaps the box's add region)"""
    n, h, o = (16, 12, 14), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)
    rng = np.random.default_rng(19)
    vel = [rng.uniform(-3, 3, int(np.prod(F.face_shape(n, a)))).astype(np.float32) for a in range(3)]
    g = Grid(n, h, o)
    fields = [F.array_n("float", F.face_shape(n, a), vel[a]) for a in range(3)]
    ball = O.collider_sdf3([B.sphere((3.5, 6.0, 7.0), 2.5)], n, h, o)
    box = O.collider_sdf3([B.box((4.5, 6.0, 7.0), (2.0, 2.0, 3.0))], n, h, o)
    inflows = [(F.array_n("float", n, ball), (1.5, -0.25, 0.5), "set"),
               (F.array_n("float", n, box), (0.75, 0.0, -2.0), "add")]
    sdf_h = sdf = None
    if collider == "sdf":
        sdf_h = O.collider_sdf3([B.sphere((12.0, 6.0, 7.0), 2.2)], n, h, o)
        sdf = F.array_n("float", n, sdf_h)
    bs = B.BoundarySolver3(g, fields, sdf=sdf, friction=0.5, inflows=inflows)
    bs.closed_domain_boundary_flag = 63
    bs.constrain_velocity(5)
    want = O.constrain_velocity3(vel, n, h, o, closed_flags=63, sdf=sdf_h, friction=0.5, depth=5,
                                 inflows=[(ball, (1.5, -0.25, 0.5), "set"), (box, (0.75, 0.0, -2.0), "add")])
    for a in range(3):
        np.testing.assert_allclose(fields[a].to_array(), want[a], rtol=1e-6, atol=1e-6)
    # the ball's faces hold its velocity (or it plus the box's where both cover them)
    assert np.count_nonzero(fields[0].to_array() == np.float32(1.5)) > 0
    with pytest.raises(Exception, match="mode"):
        B.BoundarySolver3(g, fields, inflows=[(inflows[0][0], (0, 0, 0), "replace")])


def test_outflow_dirichlet_matches_oracle_and_drives_a_projection(flux):
    """createOutflowPressureDirichlet3 over two outflow objects combined with
    a user pin (combineDirichlet3): the fields equal the oracle's; a
    projection with them has zero pressure inside the outflows, the pin's
    target at the pin, and converges"""
    n, h, o = (20, 16, 12), (0.5, 0.5, 0.5), (0.0, 0.0, 0.0)
    g = Grid(n, h, o)
    o1 = O.collider_sdf3([B.sphere((8.0, 7.5, 3.0), 1.6)], n, h, o)
    o2 = O.collider_sdf3([B.box((1.5, 4.0, 3.0), (1.0, 4.0, 3.0))], n, h, o)
    cells = int(np.prod(n))
    ua = np.zeros(cells, np.uint8)
    ut = np.zeros(cells, np.float32)
    pin = 3 + n[0] * (2 + n[1] * 4)
    ua[pin] = 1
    ut[pin] = 0.75
    act, tgt = B.outflow_dirichlet3(g, [F.array_n("float", n, o1), F.array_n("float", n, o2)],
                                    F.array_n("u8", n, ua), F.array_n("float", n, ut))
    wa, wt = O.outflow_dirichlet3([o1, o2], n, h, o, ua, ut)
    np.testing.assert_array_equal(act.to_array(), wa)
    np.testing.assert_array_equal(tgt.to_array(), wt)
    assert wa.sum() > 1 + 10 and wa[pin] == 1
    rng = np.random.default_rng(3)
    vel = [F.array_n("float", F.face_shape(n, a), rng.standard_normal(int(np.prod(F.face_shape(n, a))))
                     .astype(np.float32)) for a in range(3)]
    ps = F.create_grid_pressure_solver3(resolution=n, grid_spacing=h, dirichlet=(act, tgt),
                                        multigrid={"levels": 2}, tolerance=1e-6)
    d = ps.project(vel)
    assert d["converged"], d
    p = ps.pressure.to_array()
    np.testing.assert_allclose(p[(wa != 0) & (np.arange(cells) != pin)], 0.0, atol=1e-6)
    assert abs(p[pin] - 0.75) <= 1e-6
