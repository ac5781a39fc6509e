"""Transfers on the GPU against the oracle (GPU tests).

Bit-exact: binning (cell index, stable permutation, cell starts), P2G in
both modes (the gather's f32 sums run in the same order as the oracle's; the
fixed-point mode's integer sums are order-free), two-phase beta.
Tolerance 1e-6 absolute (positions of order 1): G2P, particle advection,
semi-Lagrangian face advection (sqrt/ceil/floor may differ in the last bit
near ties, which moves a back-trace by a rounding error only).
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from tsl_array_n_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T(flux):
    from tsl_array_n_amd import transfer
    return transfer


def seeded(n, box_max=(0.4, 0.6, 1.0)):
    h = 1.0 / n
    pos = O.flip_box_seed3([0, 0, 0], box_max, [h, h, h], ppa=2, jitter=0.2)
    return pos, (h, h, h)


def test_binning_is_bit_exact(flux, T):
    F = flux
    n = (24, 20, 16)
    rng = np.random.default_rng(4)
    L = np.array(n, np.float64) * 0.05
    pos = (rng.uniform(0, 1, (20000, 3)) * L).astype(np.float32).ravel()
    g = T.Grid(n, (0.05, 0.05, 0.05))
    P = F.array_n("vec3", 20000, pos)
    bins = T.bin_particles(g, P, 20000)
    cells = O.cell_index3(pos, [0, 0, 0], [0.05] * 3, n)
    perm, start = O.sort_by_cell(cells, int(np.prod(n)))
    np.testing.assert_array_equal(bins.cell.to_array(), np.sort(cells, kind="stable"))
    np.testing.assert_array_equal(bins.perm.to_array(), perm)
    np.testing.assert_array_equal(bins.cell_start.to_array(), start)
    assert bins.outside == 0
    # deterministic across runs
    b2 = T.bin_particles(g, P, 20000)
    np.testing.assert_array_equal(b2.perm.to_array(), perm)


def test_binning_counts_outside_particles(flux, T):
    F = flux
    g = T.Grid((4, 4, 4), (1, 1, 1))
    pos = np.array([0.5, 0.5, 0.5, 5.0, 1.0, 1.0, -0.5, 2, 2], np.float32)
    bins = T.bin_particles(g, F.array_n("vec3", 3, pos), 3)
    assert bins.outside == 2


def test_seeded_dam_break_binning_and_beta(flux, T):
    """The seeded 64^3 dam break bins into exactly the liquid block the scene
    builder uses, and the GPU two-phase beta equals the oracle's."""
    F = flux
    n = 64
    pos, h = seeded(n)
    count = pos.size // 3
    assert count == scenes.particle_count(n)
    g = T.Grid((n, n, n), h)
    P = F.array_n("vec3", count, pos)
    bins = T.bin_particles(g, P, count)
    frac = F.array_n("float", (n, n, n))
    beta = T.mac_fields(g)
    T.two_phase_beta(g, bins, 1.0, 1000.0, frac, beta)
    np.testing.assert_array_equal(frac.to_array(), scenes.dam_break_fraction(n))
    ref = O.two_phase_beta3(scenes.dam_break_fraction(n), (n, n, n), 1.0, 1000.0)
    host = scenes.two_phase_beta(scenes.dam_break_fraction(n), (n, n, n))
    for a in range(3):
        np.testing.assert_array_equal(beta[a].to_array(), ref[a])
        np.testing.assert_array_equal(host[a], ref[a])


@pytest.mark.parametrize("comp", [0, 1, 2])
def test_p2g_gather_bit_exact(flux, T, comp):
    F = flux
    n = (20, 16, 12)
    h = (0.05, 0.05, 0.05)
    rng = np.random.default_rng(comp)
    L = np.array(n) * 0.05
    cnt = 30000
    pos = (rng.uniform(0.001, 0.999, (cnt, 3)) * L).astype(np.float32).ravel()
    vel = rng.standard_normal(3 * cnt).astype(np.float32)
    g = T.Grid(n, h)
    P = F.array_n("vec3", cnt, pos)
    V = F.array_n("vec3", cnt, vel)
    bins = T.bin_particles(g, P, cnt)
    fs = g.face_shape(comp)
    out = F.array_n("float", fs)
    valid = F.array_n("int", fs)
    T.p2g_gather(g, P, V, bins, comp, out, valid)
    cells = O.cell_index3(pos, [0, 0, 0], h, n)
    perm, start = O.sort_by_cell(cells, int(np.prod(n)))
    ref, rv = O.p2g_gather3(pos, vel, perm, start, n, h, [0, 0, 0], comp)
    np.testing.assert_array_equal(valid.to_array(), rv)
    np.testing.assert_array_equal(out.to_array(), ref)


@pytest.mark.parametrize("comp", [0, 1, 2])
def test_p2g_fixed_point_bit_exact(flux, T, comp):
    F = flux
    n = (16, 12, 10)
    h = (0.1, 0.1, 0.1)
    rng = np.random.default_rng(10 + comp)
    cnt = 20000
    pos = (rng.uniform(0.0, 1.0, (cnt, 3)) * np.array(n) * 0.1).astype(np.float32).ravel()
    vel = (rng.standard_normal(3 * cnt) * 3).astype(np.float32)
    g = T.Grid(n, h)
    P = F.array_n("vec3", cnt, pos)
    V = F.array_n("vec3", cnt, vel)
    fs = g.face_shape(comp)
    numer, denom = F.array_n("int", fs), F.array_n("int", fs)
    out, valid = F.array_n("float", fs), F.array_n("int", fs)
    T.p2g_fixed(g, P, V, cnt, comp, numer, denom, out, valid)
    ro, rv, rn, rd = O.p2g_fixed3(pos, vel, n, h, [0, 0, 0], comp)
    np.testing.assert_array_equal(numer.to_array(), rn)
    np.testing.assert_array_equal(denom.to_array(), rd)
    np.testing.assert_array_equal(valid.to_array(), rv)
    np.testing.assert_array_equal(out.to_array(), ro)


@pytest.mark.parametrize("comp", [0, 1])
def test_p2g_block_box_and_direct_atomics_are_bit_identical(flux, T, comp, monkeypatch):
    """The scatter sums a block's corners in LDS when they fit a small box
    (particles in seeded order) and adds them to the grid directly otherwise
    (the same particles shuffled, or FLUX_P2G_BOX=0): integer sums in any
    order, so all three give the oracle's bits."""
    F = flux
    n = (48, 40, 32)
    h = (1.0 / 48,) * 3
    pos = O.flip_box_seed3([0, 0, 0], [1.0, 1.0, 0.5], list(h), ppa=2, jitter=0.2)
    cnt = pos.size // 3
    rng = np.random.default_rng(20 + comp)
    vel = (rng.standard_normal(3 * cnt) * 2).astype(np.float32)
    perm = rng.permutation(cnt)
    shuffled_pos = pos.reshape(-1, 3)[perm].ravel()
    shuffled_vel = vel.reshape(-1, 3)[perm].ravel()
    g = T.Grid(n, h)
    fs = g.face_shape(comp)
    _, rv, rn, rd = O.p2g_fixed3(pos, vel, n, h, [0, 0, 0], comp)
    for name, p_, v_, box in (("ordered", pos, vel, "1"), ("shuffled", shuffled_pos, shuffled_vel, "1"),
                              ("direct", pos, vel, "0")):
        monkeypatch.setenv("FLUX_P2G_BOX", box)
        P = F.array_n("vec3", cnt, p_)
        V = F.array_n("vec3", cnt, v_)
        numer, denom = F.array_n("int", fs), F.array_n("int", fs)
        out, valid = F.array_n("float", fs), F.array_n("int", fs)
        T.p2g_fixed(g, P, V, cnt, comp, numer, denom, out, valid)
        np.testing.assert_array_equal(numer.to_array(), rn, err_msg=name)
        np.testing.assert_array_equal(denom.to_array(), rd, err_msg=name)
        np.testing.assert_array_equal(valid.to_array(), rv, err_msg=name)


def _random_mac(n, rng, scale=1.0):
    return [(rng.standard_normal(int(np.prod(O.face_shape(n, a)))) * scale).astype(np.float32) for a in range(3)]


def test_g2p_matches_oracle(flux, T):
    F = flux
    n = (12, 10, 8)
    h = (0.1, 0.1, 0.1)
    rng = np.random.default_rng(7)
    cnt = 5000
    pos = (rng.uniform(0.0, 1.0, (cnt, 3)) * np.array(n) * 0.1).astype(np.float32).ravel()
    vel = rng.standard_normal(3 * cnt).astype(np.float32)
    new, old = _random_mac(n, rng), _random_mac(n, rng)
    g = T.Grid(n, h)
    P = F.array_n("vec3", cnt, pos)
    V = F.array_n("vec3", cnt, vel)
    T.g2p(g, P, V, cnt, T.mac_fields(g, new), T.mac_fields(g, old), flip_ratio=0.95, damping=0.02)
    ref = O.g2p3(pos, vel, *new, *old, n, h, [0, 0, 0], flip_ratio=0.95, damping=0.02)
    np.testing.assert_allclose(V.to_array(), ref, rtol=0, atol=1e-6)


def test_particle_advection_matches_oracle(flux, T):
    F = flux
    n = (16, 16, 16)
    h = (1 / 16,) * 3
    rng = np.random.default_rng(8)
    cnt = 8000
    pos = rng.uniform(0.0, 1.0, (cnt, 3)).astype(np.float32).ravel()
    vel = _random_mac(n, rng, 0.5)
    g = T.Grid(n, h)
    P = F.array_n("vec3", cnt, pos)
    T.advect_particles(g, P, cnt, T.mac_fields(g, vel), 0.05)
    ref = O.advect_particles3(pos, *vel, n, h, [0, 0, 0], 0.05)
    got = P.to_array()
    assert (got >= 1e-4 - 1e-7).all() and (got <= 1 - 1e-4 + 1e-7).all()
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-6)


def test_face_advection_matches_oracle(flux, T):
    F = flux
    n = (16, 12, 10)
    h = (0.1, 0.1, 0.1)
    rng = np.random.default_rng(9)
    vel = _random_mac(n, rng, 0.3)
    g = T.Grid(n, h)
    vin = T.mac_fields(g, vel)
    vout = T.mac_fields(g)
    T.advect_faces(g, vin, vout, 0.1)
    ref = O.advect_face3(*vel, n, h, [0, 0, 0], 0.1)
    for a in range(3):
        np.testing.assert_allclose(vout[a].to_array(), ref[a], rtol=0, atol=1e-5)


def test_divergence_matches_oracle(flux, T):
    F = flux
    n = (10, 9, 8)
    h = (0.5, 0.25, 1.0)
    rng = np.random.default_rng(1)
    vel = _random_mac(n, rng)
    g = T.Grid(n, h)
    out = F.array_n("float", n)
    T.divergence(g, T.mac_fields(g, vel), out)
    ref = O.divergence3(n, h, vel, suffix="f32")
    np.testing.assert_array_equal(out.to_array(), ref)


def test_p2g_fixed_point_ties_round_half_to_even(flux, T):
    """Particles whose value * w * 65536 and w * 65536 are exact half
    integers (grid_flip_solver3.js:508-535, WGSL round = half to even): the
    GPU's integer sums equal the closed form and the oracle's."""
    F = flux
    halves = [0.5, 1.5, 2.5, 3.5, -0.5, -1.5, -2.5, 7.5]
    pos, vel = [], []
    for q, k in enumerate(halves):
        pos.append((float(1 + q % 4), 2 + q // 4 + 0.5, 3.5))
        vel.append((k / 65536.0, 0.0, 0.0))
    pos.append((1.0 + 2.5 / 65536.0, 1.5, 1.5))
    vel.append((0.0, 0.0, 0.0))
    pos = np.array(pos, np.float32).ravel()
    vel = np.array(vel, np.float32).ravel()
    n, h = (6, 6, 6), (1.0, 1.0, 1.0)
    cnt = len(pos) // 3
    g = T.Grid(n, h)
    P = F.array_n("vec3", cnt, pos)
    V = F.array_n("vec3", cnt, vel)
    fs = g.face_shape(0)
    numer, denom = F.array_n("int", fs), F.array_n("int", fs)
    out, valid = F.array_n("float", fs), F.array_n("int", fs)
    T.p2g_fixed(g, P, V, cnt, 0, numer, denom, out, valid)
    nu, de = numer.to_array(), denom.to_array()
    for q, k in enumerate(halves):
        f = (1 + q % 4) + fs[0] * ((2 + q // 4) + fs[1] * 3)
        assert nu[f] == round(k) and de[f] == 65536, (k, nu[f])
    assert de[1 + fs[0] * (1 + fs[1])] == 65534 and de[2 + fs[0] * (1 + fs[1])] == 2
    _, _, rn, rd = O.p2g_fixed3(pos, vel, n, h, [0, 0, 0], 0)
    np.testing.assert_array_equal(nu, rn)
    np.testing.assert_array_equal(de, rd)
