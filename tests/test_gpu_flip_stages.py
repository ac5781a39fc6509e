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
