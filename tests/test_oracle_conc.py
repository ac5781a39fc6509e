"""Pins the oracle's carried concentration, density coupling, mass-weighted
transfer, pressure pin and phase-aware resampling of the FLIP solver
(grid_flip_solver3.js:265-357, 493-531, 848-880, 967-1049;
grid_two_phase_flip_solver2.js:133-169, 1131-1503) by hand-built cases whose
answers are worked out here, and by the physics they must keep: a stably
stratified closed box of fluid stays at rest, the unstable one overturns.
No GPU is used."""
import numpy as np

from oracle import pyoracle as O

f32 = np.float32


def cell_centre(n, i, j, k, h=1.0):
    return [(i + 0.5) * h, (j + 0.5) * h, (k + 0.5) * h]


def test_accumulate_and_resolve_by_hand():
    n = (4, 3, 2)
    cells = {(0, 0, 0): [0.25, 0.75, 1.0], (2, 1, 1): [-0.5, 1.7], (3, 2, 0): [0.3]}
    pos, conc = [], []
    for (i, j, k), cs in cells.items():
        for c in cs:
            pos.append(cell_centre(n, i, j, k))
            conc.append(c)
    accum, cnt = O.conc_accumulate3(np.array(pos, f32).ravel(), np.array(conc, f32), n, (1, 1, 1), (0, 0, 0))
    idx = lambda i, j, k: i + n[0] * (j + n[1] * k)
    assert cnt[idx(0, 0, 0)] == 3 and cnt[idx(2, 1, 1)] == 2 and cnt.sum() == 6
    # round(c * 4096) of the clamped concentrations: 1024 + 3072 + 4096; 0 + 4096; round(1228.8) = 1229
    assert accum[idx(0, 0, 0)] == 8192 and accum[idx(2, 1, 1)] == 4096 and accum[idx(3, 2, 0)] == 1229
    cc, rho = O.conc_resolve3(accum, cnt, density=(1.0, 1000.0))
    assert cc[idx(0, 0, 0)] == f32(f32(2.0) / f32(3.0))
    assert cc[idx(2, 1, 1)] == 0.5 and cc[idx(3, 2, 0)] == f32(1229 / 4096)
    assert cc[idx(1, 1, 1)] == 0.0 and rho[idx(1, 1, 1)] == 1.0  # empty cell: ambient
    assert rho[idx(2, 1, 1)] == f32(1 + 999 * 0.5)
    cc2, none = O.conc_resolve3(accum, cnt)
    np.testing.assert_array_equal(cc2, cc)
    assert none is None


def test_density_beta_cases():
    n = (3, 1, 1)
    rho = np.array([1.0, 1000.0, 500.0], f32)
    mask = np.array([1.0, 1.0, 0.0], f32)
    bu, bv, bw = O.density_beta3(rho, mask, n, 1.0, 1000.0)
    # U faces: both sides cell 0 (fluid, 1); cells 0|1 fluid: mean 500.5; 1 fluid | 2 air: 1000; 2|2 air: reference
    np.testing.assert_array_equal(bu, np.array([1000.0, f32(1000.0) / f32(500.5), 1.0, 1.0], f32))
    # V and W faces sit on one cell: fluid cells take their own density, the air cell the reference
    expect = np.array([1000.0, 1.0, 1.0], f32)
    np.testing.assert_array_equal(bv.reshape(2, 3)[0], expect)
    np.testing.assert_array_equal(bw.reshape(2, 3)[1], expect)
    # a zero density floors at 1e-4
    b0, _, _ = O.density_beta3(np.array([0.0, 0.0, 0.0], f32), np.ones(3, f32), n, 1.0, 2.0)
    assert b0[1] == f32(2.0) / f32(1e-4)


def test_reduced_gravity_by_hand():
    """u -= beta dt / rho_ref (g . x_face) d(rho)/dx: from grad p = grad p_r +
    rho g + (g . x) grad rho with p_r the reduced pressure"""
    n, h = (1, 2, 1), (1.0, 1.0, 1.0)
    rho = np.array([1.0, 3.0], f32)  # lighter below, heavier above
    vel = [np.zeros(int(np.prod(O.face_shape(n, a))), f32) for a in range(3)]
    beta = [np.full(v.size, 0.5, f32) for v in vel]
    out = O.reduced_gravity3(vel, beta, rho, n, h, (0, 0, 0), (0.0, -10.0, 0.0), 0.1, 1.0, 3.0)
    # the one interior V face at y = 1: 0.5 * 0.1 / 3 * (-10 * 1) * (3 - 1) / 1 = -1/3, subtracted
    np.testing.assert_allclose(out[1], [0.0, 1.0 / 3.0, 0.0], rtol=1e-6)
    assert not out[0].any() and not out[2].any()


def test_mix_and_fade():
    n = (2, 2, 2)
    pos = np.array([cell_centre(n, 0, 0, 0), cell_centre(n, 1, 1, 1), [1.0, 1.0, 1.0]], f32).ravel()
    conc = np.array([-0.2, 0.75, 1.3], f32)
    cell = np.full(8, 0.25, f32)
    np.testing.assert_array_equal(O.mix_fade3(pos, conc, cell, n, (1, 1, 1), (0, 0, 0), 0.0, 0.0), [0.0, 0.75, 1.0])
    assert not O.mix_fade3(pos, conc, cell, n, (1, 1, 1), (0, 0, 0), 0.3, 1.0).any()
    out = O.mix_fade3(pos, conc, cell, n, (1, 1, 1), (0, 0, 0), 0.5, 0.2)
    np.testing.assert_allclose(out, [(0 + 0.25 * 0.5) * 0.8, (0.75 - 0.25) * 0.8, (1 - 0.375) * 0.8], rtol=1e-6)
    # mixing and fade outside [0, 1] clamp; a non-finite cell value leaves c alone
    np.testing.assert_allclose(O.mix_fade3(pos, conc, cell, n, (1, 1, 1), (0, 0, 0), 7.0, -1.0), [0.25] * 3,
                               rtol=1e-6)
    nan_cell = np.full(8, np.nan, f32)
    np.testing.assert_array_equal(O.mix_fade3(pos, conc, nan_cell, n, (1, 1, 1), (0, 0, 0), 1.0, 0.0),
                                  [0.0, 0.75, 1.0])


def test_mass_weighted_scatter():
    n, h = (4, 4, 4), (1.0, 1.0, 1.0)
    p = [1.3, 2.1, 1.7]
    pos = np.array([p, p], f32).ravel()
    vel = np.array([[1.0, 0, 0], [4.0, 0, 0]], f32).ravel()
    conc = np.array([0.0, 1.0], f32)
    out, valid, _, _ = O.p2g_fixed3(pos, vel, n, h, (0, 0, 0), 0, mass=(conc, 1.0, 3.0))
    # masses 1/3 and 1: (1/3 + 4) / (4/3) = 3.25 on every face the pair reaches
    np.testing.assert_allclose(out[valid == 1], 3.25, atol=1e-3)  # fixed point at 65536
    plain, pv, _, _ = O.p2g_fixed3(pos, vel, n, h, (0, 0, 0), 0)
    np.testing.assert_allclose(plain[pv == 1], 2.5, atol=1e-3)
    # a uniform concentration weights every particle alike: the plain average
    rng = np.random.default_rng(1)
    pos = rng.uniform(0.2, 3.8, (200, 3)).astype(f32).ravel()
    vel = rng.uniform(-1, 1, 600).astype(f32)
    a, av, _, _ = O.p2g_fixed3(pos, vel, n, h, (0, 0, 0), 1, mass=(np.full(200, 0.5, f32), 1.0, 3.0))
    b, bv, _, _ = O.p2g_fixed3(pos, vel, n, h, (0, 0, 0), 1)
    np.testing.assert_array_equal(av, bv)
    np.testing.assert_allclose(a, b, atol=1e-3)


def test_phase_resample_by_hand():
    """two plus-shaped neighbourhoods whose empty centres are the only
    recipients (a cell off the arms sees at most 2 well-filled neighbours):
    the liquid-majority centre takes liquid donors, the gas-majority one gas
    donors, each pool ranked by particle index, and the donors keep their
    concentration"""
    n, h = (9, 9, 9), (1.0, 1.0, 1.0)
    pos, conc = [], []

    def fill(c, count, value):
        for _ in range(count):
            pos.append([c[0] + 0.25, c[1] + 0.25, c[2] + 0.25])
            conc.append(value)

    L, G = (2, 4, 4), (6, 4, 4)
    arms = [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]
    mixed_l = (L[0] - 1, L[1], L[2])   # over-full: 7 liquid, then 5 gas
    mixed_g = (G[0] + 1, G[1], G[2])   # over-full: 5 gas, then 5 liquid
    first = None
    for d in arms:
        c = (L[0] + d[0], L[1] + d[1], L[2] + d[2])
        if c == mixed_l:
            first = len(pos)
            fill(c, 7, 1.0)
            fill(c, 5, 0.0)
        else:
            fill(c, 4, 1.0)
    for d in arms:
        c = (G[0] + d[0], G[1] + d[1], G[2] + d[2])
        if c == mixed_g:
            fill(c, 5, 0.0)
            fill(c, 5, 1.0)
        else:
            fill(c, 4, 0.0)
    pos = np.array(pos, f32).ravel()
    conc = np.array(conc, f32)
    accum, cnt = O.conc_accumulate3(pos, conc, n, h, (0, 0, 0))
    mac = [np.full(int(np.prod(O.face_shape(n, a))), 0.5 * (a + 1), f32) for a in range(3)]
    p2, v2 = O.resample_phase3(pos, np.zeros_like(pos), conc, *mac, n, h, (0, 0, 0), 4, 8, cnt, accum)
    P, P2 = pos.reshape(-1, 3), p2.reshape(-1, 3)
    moved = np.flatnonzero(np.any(P2 != P, axis=1))
    to_l = [q for q in moved if tuple(P2[q]) == tuple(cell_centre(n, *L))]
    to_g = [q for q in moved if tuple(P2[q]) == tuple(cell_centre(n, *G))]
    assert to_l == list(range(first, first + 4))            # the four lowest liquid donors
    assert to_g == list(range(first + 7, first + 11))       # the four lowest gas donors
    assert len(moved) == 8
    np.testing.assert_allclose(v2.reshape(-1, 3)[moved], [[0.5, 1.0, 1.5]] * 8, rtol=1e-6)
    # the single-phase pass would hand the gas-majority centre liquid particles
    p1, _ = O.resample3(pos, np.zeros_like(pos), *mac, n, h, (0, 0, 0), 4, 8)
    to_g1 = [q for q in range(P.shape[0]) if tuple(p1.reshape(-1, 3)[q]) == tuple(cell_centre(n, *G))]
    assert any(conc[q] > 0.5 for q in to_g1)


def stratified_box(heavy_below, n=(8, 8, 8), per_axis=2, seed=3, bump=0.0):
    """a closed box full of fluid, concentration 1 in one half (y); the
    interface height varies by +-bump cells along x"""
    g = np.stack(np.meshgrid(*[(np.arange(n[a] * per_axis) + 0.5) / per_axis for a in range(3)], indexing="ij"),
                 -1).reshape(-1, 3)
    rng = np.random.default_rng(seed)
    pos = (g + rng.uniform(-0.1, 0.1, g.shape)).astype(f32)
    low = pos[:, 1] < n[1] / 2 + bump * np.sin(2 * np.pi * pos[:, 0] / n[0])
    conc = (low if heavy_below else ~low).astype(f32)
    return pos.ravel(), conc


def run_box(heavy_below, steps, n=(8, 8, 8), bump=0.0):
    pos, conc = stratified_box(heavy_below, n, bump=bump)
    cells = int(np.prod(n))
    st = {"pos": pos, "vel": np.zeros_like(pos), "conc": conc, "pressure": np.zeros(cells, f32),
          "p_accepted": np.zeros(cells, f32)}
    for a in range(3):
        st["uvw"[a]] = np.zeros(int(np.prod(O.face_shape(n, a))), f32)
    for _ in range(steps):
        d = O.flip3_step(st, n, (1.0, 1.0, 1.0), (0, 0, 0), 0.05, levels=2, carry_concentration=True,
                         density=(1.0, 3.0), pressure_pin="topCenter", resample=False)
        assert d["converged"] and not d["rejected"]
    return st


def test_stable_stratification_stays_at_rest():
    st = run_box(heavy_below=True, steps=3)
    assert np.abs(st["vel"]).max() < 1e-3
    assert set(np.unique(st["conc"])) <= {0.0, 1.0}


def test_unstable_stratification_overturns():
    """heavy above light with a wavy interface: the pinned, density-coupled
    solve lets the heavy fluid sink where the interface dips (mean vertical
    velocity of the heavy particles there negative, of the light ones where
    it bulges positive)"""
    pos, conc = stratified_box(heavy_below=False, bump=1.0)
    st = run_box(heavy_below=False, steps=3, bump=1.0)
    v = st["vel"].reshape(-1, 3)
    heavy = conc > 0.5
    assert np.abs(v).max() > 1e-2
    assert v[heavy, 1].mean() < 0 < v[~heavy, 1].mean()


def test_pin_only_without_dirichlet_cells():
    assert O.pin_cell3((8, 6, 4), "topCenter") == (4, 5, 2)
    assert O.pin_cell3((8, 6, 4), None) is None
    assert O.pin_cell3((8, 6, 4), [1, 2, 3]) == (1, 2, 3)
