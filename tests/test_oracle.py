"""Pins the CPU oracle (oracle/) before anything is compared against it.

Every check here is independent of the oracle's own formulas: the reference's
captured systems and generator output, an edge-based sparse assembly of the
operator (scipy), closed forms, adjoint identities and brute force on tiny
grids.  No GPU is used.
"""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import pyoracle as O

from conftest import GOLDEN


# ------------------------------------------------------------------ fixtures
def test_frame110_residual_matches_reported(frame110):
    """FP64 ||b - A p|| on the captured pressure reproduces the GPU-reported
    residual (stalled-system/data/frame-110/meta.json) -- the agreement test of
    cpu_reference.mjs:145-159 (within a factor 2)."""
    c = frame110
    r = O.true_residual(c["shape"], c["spacing"], c["b"], c["pressure"], c["mask"], c["weights"])
    rep = c["meta"]["diagnostics"]["residual"]
    assert rep / 2 < r < rep * 2, (r, rep)
    assert c["mask"].sum() == 2016 and c["vent"].sum() == 1152


def test_frame112_residual_matches_reported(frame112):
    c = frame112
    r = O.true_residual(c["shape"], c["spacing"], c["b"], c["pressure"], c["mask"], c["weights"])
    rep = c["meta"]["diagnostics"]["residual"]
    assert rep / 2 < r < rep * 2, (r, rep)


def test_symmetric_vcycle_solves_frame112(frame112):
    """The captured system the reference stalled on (pAp-growth at 38) is
    solved by a symmetric V-cycle with a correct dot (stalled-system/main.js:12-23)."""
    c = frame112
    mg = O.Multigrid(c["shape"], c["spacing"], levels=4, nu_down=3, nu_up=3, n_coarsest=30,
                     mask=c["mask"], coarse_mask=c["vent"], weights=c["weights"])
    x, st = O.pcg_solve(c["shape"], c["spacing"], c["b"], mask=c["mask"], weights=c["weights"], mg=mg,
                        tol=1e-5, maxiter=100)
    assert st["converged"] and st["stoppedBy"] == "converged"
    assert st["iterations"] < 38
    r = O.true_residual(c["shape"], c["spacing"], c["b"], x, c["mask"], c["weights"])
    assert r < st["threshold"]


def test_mulberry32_matches_reference_vectors():
    d = json.load(open(os.path.join(GOLDEN, "mulberry32.json")))
    assert d["defaultSeed"] == 0x9E3779B9
    for v in d["vectors"]:
        np.testing.assert_array_equal(O.rng(v["seed"], len(v["values"])), np.array(v["values"]))


# -------------------------------------------- independent operator assembly
def _cidx(n, i, j, k):
    return i + n[0] * (j + n[1] * k)


def edge_assembly(shape, spacing, mask=None, weights=None):
    """A assembled face by face (graph-Laplacian form), independent of the
    oracle's row stencil: every interior face between cells a and b with
    weight w adds -w/h^2 to both diagonals and +w/h^2 to the couplings; a
    coupling to a Dirichlet cell is then dropped (elimination, diagonal
    kept) and a Dirichlet cell's row becomes -e (multigrid.js:273-339)."""
    n = list(shape) + [1] * (3 - len(shape))
    N = int(np.prod(n))
    rows, cols, vals = [], [], []
    for a in range(len(shape)):
        h2 = spacing[a] ** 2
        fs = list(n)
        fs[a] += 1
        for k in range(n[2]):
            for j in range(n[1]):
                for i in range(n[0]):
                    I = [i, j, k]
                    if I[a] == n[a] - 1:
                        continue
                    J = list(I)
                    J[a] += 1
                    ca, cb = _cidx(n, *I), _cidx(n, *J)
                    w = 1.0 if weights is None else float(weights[a][J[0] + fs[0] * (J[1] + fs[1] * J[2])])
                    rows += [ca, cb, ca, cb]
                    cols += [ca, cb, cb, ca]
                    vals += [-w / h2, -w / h2, w / h2, w / h2]
    A = sp.coo_matrix((vals, (rows, cols)), shape=(N, N)).tocsr().toarray()
    if mask is not None:
        m = mask.astype(bool)
        off = ~np.eye(N, dtype=bool)
        A[np.ix_(~m, m)] = np.where(off[np.ix_(~m, m)], 0.0, A[np.ix_(~m, m)])
        A[m, :] = 0.0
        A[m, m] = -1.0
    return A


def _random_system(shape, seed, with_mask=True, with_weights=True):
    rng = np.random.default_rng(seed)
    n = list(shape) + [1] * (3 - len(shape))
    N = int(np.prod(n))
    mask = (rng.random(N) < 0.2).astype(np.uint8) if with_mask else None
    weights = None
    if with_weights:
        weights = []
        for a in range(len(shape)):
            fs = list(n)
            fs[a] += 1
            weights.append(rng.uniform(0.1, 3.0, int(np.prod(fs))))
    return mask, weights


@pytest.mark.parametrize("shape,spacing", [((5, 4, 3), (1.0, 0.5, 2.0)), ((6, 5), (0.7, 1.3)), ((4, 4, 4), (0.25,) * 3)])
@pytest.mark.parametrize("with_mask,with_weights", [(False, False), (True, False), (False, True), (True, True)])
def test_operator_equals_edge_assembly(shape, spacing, with_mask, with_weights):
    mask, weights = _random_system(shape, 7, with_mask, with_weights)
    A = edge_assembly(shape, spacing, mask, weights)
    N = A.shape[0]
    rng = np.random.default_rng(3)
    for _ in range(3):
        x = rng.standard_normal(N)
        np.testing.assert_allclose(O.apply(shape, spacing, x, mask, weights), A @ x, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(O.diag(shape, spacing, mask, weights), np.diag(A), rtol=1e-13)
    # unmasked rows are symmetric; the f32 arm is the same operator rounded
    np.testing.assert_allclose(A, A.T, atol=1e-12) if mask is None else None
    x = rng.standard_normal(N)
    ref = A @ x
    got32 = O.apply(shape, spacing, x, mask, weights, suffix="f32")
    assert np.linalg.norm(got32 - ref) <= 1e-5 * np.linalg.norm(ref)


def test_unweighted_operator_is_kronecker_laplacian():
    """No mask, no weights: A is the Kronecker sum of 1-D Neumann Laplacians."""
    shape, spacing = (6, 5, 4), (1.0, 0.5, 2.0)

    def lap1(n, h):
        d = np.zeros((n, n))
        for i in range(n - 1):
            d[i, i] -= 1; d[i + 1, i + 1] -= 1; d[i, i + 1] += 1; d[i + 1, i] += 1
        return d / h ** 2
    L0, L1, L2 = (lap1(n, h) for n, h in zip(shape, spacing))
    I0, I1, I2 = (np.eye(n) for n in shape)
    A = np.kron(I2, np.kron(I1, L0)) + np.kron(I2, np.kron(L1, I0)) + np.kron(L2, np.kron(I1, I0))
    x = np.random.default_rng(1).standard_normal(A.shape[0])
    np.testing.assert_allclose(O.apply(shape, spacing, x), A @ x, rtol=1e-12, atol=1e-12)


def test_masked_free_block_is_symmetric_negative_definite():
    shape = (5, 4, 3)
    mask, weights = _random_system(shape, 11)
    A = edge_assembly(shape, (1.0, 1.0, 1.0), mask, weights)
    free = ~mask.astype(bool)
    Af = A[np.ix_(free, free)]
    np.testing.assert_allclose(Af, Af.T, atol=1e-12)
    assert np.linalg.eigvalsh(Af).max() < 0


# ------------------------------------------------------ multigrid identities
@pytest.mark.parametrize("shape", [(8, 8, 8), (16, 8), (8, 4, 12)])
def test_restriction_prolongation_adjoint(shape):
    """(u, P v) / (R u, v) = 2^d for the taps of multigrid.js:766-780, 921-934
    (examples/31-conjugate-gradient-3d/main.js:105-166)."""
    rng = np.random.default_rng(5)
    nf = int(np.prod(shape))
    nc = int(np.prod([s // 2 for s in shape]))
    ratios = []
    for _ in range(3):
        u = rng.standard_normal(nf)
        v = rng.standard_normal(nc)
        Pv = O.correct(shape, v, np.zeros(nf))
        Ru = O.restrict(shape, u)
        ratios.append(np.dot(u, Pv) / np.dot(Ru, v))
    np.testing.assert_allclose(ratios, 2.0 ** len(shape), rtol=1e-12)


def test_vcycle_is_symmetric():
    """<u, M v> = <M u, v>: post-smoothing runs the colours in reverse
    (multigrid.js:1562-1566, 1640).  Masks, vent coarsening and weights on.
    The reference's coarsest sweep (colour 0 then 1, 679-737) is symmetric
    only once converged on a definite coarse system, so the Dirichlet slab is
    thick enough to survive ALL-children coarsening and the coarse sweep runs
    to convergence; an unreversed post-smooth breaks the identity anyway."""
    n = (16, 8, 8)
    N = 16 * 8 * 8
    rng = np.random.default_rng(21)
    mask = np.zeros((8, 8, 16), np.uint8)
    mask[:, :4, :] = 1  # y < 4 Dirichlet (and vent)
    mask = mask.ravel()
    weights = [rng.uniform(0.5, 2.0, int(np.prod(O.face_shape(n, a)))) for a in range(3)]
    mg = O.Multigrid(n, (1, 1, 1), levels=3, n_coarsest=3000, mask=mask, coarse_mask=mask, weights=weights)
    free = mask == 0
    u = rng.standard_normal(N) * free
    v = rng.standard_normal(N) * free
    lhs, rhs = np.dot(u, mg.apply(v)), np.dot(mg.apply(u), v)
    assert abs(lhs - rhs) <= 1e-9 * abs(lhs)
    # and definite (negative, like A) on the free cells
    assert np.dot(u, mg.apply(u)) < 0


def test_coarse_masks_and_weights(frame110):
    """Coarsening rules of multigrid.js:1066-1184 checked by brute force on level 1."""
    c = frame110
    mg = O.Multigrid(c["shape"], c["spacing"], levels=2, mask=c["mask"], coarse_mask=c["vent"], weights=c["weights"])
    dims, cmask, cw = mg.level(1)
    nx, ny, nz = c["shape"]
    U = c["weights"][0].reshape(nz, ny, nx + 1)
    ref_u = 0.25 * (U[0::2, 0::2, 0::2] + U[0::2, 1::2, 0::2] + U[1::2, 0::2, 0::2] + U[1::2, 1::2, 0::2])
    np.testing.assert_allclose(cw[0].reshape(nz // 2, ny // 2, nx // 2 + 1), ref_u, rtol=1e-12)
    vent = c["vent"].reshape(nz, ny, nx).astype(bool)
    all_children = np.ones((nz // 2, ny // 2, nx // 2), bool)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                all_children &= vent[dz::2, dy::2, dx::2]
    cn = (nx // 2, ny // 2, nz // 2)
    tot = np.zeros(cn[::-1])
    for a in range(3):
        fs = list(cn); fs[a] += 1
        W = cw[a].reshape(fs[::-1])
        sl_lo = [slice(None)] * 3; sl_hi = [slice(None)] * 3
        sl_lo[2 - a] = slice(0, -1); sl_hi[2 - a] = slice(1, None)
        tot += W[tuple(sl_lo)] + W[tuple(sl_hi)]
    ref_mask = all_children | (tot < 1e-6)
    np.testing.assert_array_equal(cmask.reshape(cn[::-1]).astype(bool), ref_mask)


def test_analytic_poisson_16cubed():
    """examples/31: zero-flux 16^3 MGPCG against sin*cos*cos, 4 levels."""
    N = 16
    i, j, k = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    xs = (np.sin(2 * np.pi * i / N) * np.cos(2 * np.pi * j / N) * np.cos(2 * np.pi * k / N)).ravel(order="F")
    xs -= xs.mean()
    shape, h = (N, N, N), (1.0, 1.0, 1.0)
    b = O.apply(shape, h, xs)
    mg = O.Multigrid(shape, h, levels=4)
    x, st = O.pcg_solve(shape, h, b, mg=mg, tol=1e-10, maxiter=200)
    assert st["converged"]
    x -= x.mean()
    assert np.abs(x - xs).max() < 1e-6


# ---------------------------------------------------- pressure assembly
def test_one_dimensional_worked_example():
    """grid_pressure_solver2.js:47-53: u* = [0,1,1,0] -> b = [1,0,-1] ->
    p = [0,1,2] + c -> projected u = 0, here on a 3x1x1 MAC grid."""
    n = (3, 1, 1)
    U = np.array([0, 1, 1, 0], np.float64)
    V = np.zeros(3 * 2)
    W = np.zeros(3 * 2)
    A = O.pressure3_assemble(n, (1, 1, 1), [U, V, W])
    np.testing.assert_allclose(A["b"], [1, 0, -1])
    x, st = O.pcg_solve(n, (1, 1, 1), A["b"], precond="none", tol=1e-12, maxiter=10)
    np.testing.assert_allclose(x - x[0], [0, 1, 2], atol=1e-12)
    out = O.pressure3_correct(n, (1, 1, 1), x, [U, V, W])
    np.testing.assert_allclose(out[0], 0, atol=1e-12)


def test_projection_is_divergence_free_with_weights_and_dirichlet():
    """Variable-density projection (variable_density_projection.test.js:239-258):
    after u -= beta grad p the collider-free divergence equals the Dirichlet
    contribution only (zero where no Dirichlet neighbour)."""
    n = (8, 8, 8)
    rng = np.random.default_rng(9)
    vel = [rng.standard_normal(int(np.prod(O.face_shape(n, a)))) for a in range(3)]
    beta = [rng.uniform(1, 100, int(np.prod(O.face_shape(n, a)))) for a in range(3)]
    active = np.zeros(512, np.uint8)
    active[:64] = 1  # bottom slab Dirichlet, target 0
    A = O.pressure3_assemble(n, (1, 1, 1), vel, active=active, target=np.zeros(512), beta=beta)
    mg = O.Multigrid(n, (1, 1, 1), levels=2, mask=A["mask"], coarse_mask=A["vent"], weights=A["weights"])
    x, st = O.pcg_solve(n, (1, 1, 1), A["b"], x0=A["pressure"], mask=A["mask"], weights=A["weights"], mg=mg,
                        tol=1e-12, maxiter=200)
    assert st["converged"]
    out = O.pressure3_correct(n, (1, 1, 1), x, vel, beta=beta)
    # walls are left alone by the correction: close them to measure the interior
    for a, s in enumerate(out):
        fs = O.face_shape(n, a)
        arr = s.reshape(fs[::-1])
        sl0 = [slice(None)] * 3; sl1 = [slice(None)] * 3
        sl0[2 - a] = 0; sl1[2 - a] = -1
        arr[tuple(sl0)] = 0; arr[tuple(sl1)] = 0
    # rebuild with the closed walls: the free cells not touching the wall-slab see zero divergence
    div = O.divergence3(n, (1, 1, 1), out)
    free = (A["mask"] == 0).reshape(8, 8, 8)
    interior = np.zeros((8, 8, 8), bool)
    interior[2:7, 1:7, 1:7] = True
    assert np.abs(div.reshape(8, 8, 8)[free & interior]).max() < 1e-6 * np.abs(A["b"]).max()


def test_uniform_beta_is_constant_coefficient_case():
    n = (6, 6, 6)
    rng = np.random.default_rng(4)
    x = rng.standard_normal(216)
    ones = [np.ones(int(np.prod(O.face_shape(n, a)))) for a in range(3)]
    np.testing.assert_allclose(O.apply(n, (1, 1, 1), x, weights=ones), O.apply(n, (1, 1, 1), x), rtol=1e-14)
    twos = [2 * w for w in ones]
    np.testing.assert_allclose(O.apply(n, (1, 1, 1), x, weights=twos), 2 * O.apply(n, (1, 1, 1), x), rtol=1e-14)


def test_bad_cell_count():
    p = np.array([0, 1, np.nan, np.inf, -2e6, 5e5], np.float64)
    assert O.count_bad(p, 1e6) == 3


def test_mixed_arm_matches_f64_solution(frame110):
    """The HIP path's arithmetic reaches the same tolerance as FP64 on the
    captured system and its solution agrees with the FP64 one."""
    c = frame110
    mg32 = O.Multigrid(c["shape"], c["spacing"], levels=4, mask=c["mask"], coarse_mask=c["vent"],
                       weights=c["weights"], suffix="f32")
    x32, x64, st = O.pcg_solve_mixed(c["shape"], c["spacing"], c["b"], mask=c["mask"], weights=c["weights"],
                                     mg=mg32, tol=1e-8, maxiter=100)
    assert st["converged"]
    mg = O.Multigrid(c["shape"], c["spacing"], levels=4, mask=c["mask"], coarse_mask=c["vent"], weights=c["weights"])
    xr, st2 = O.pcg_solve(c["shape"], c["spacing"], c["b"], mask=c["mask"], weights=c["weights"], mg=mg,
                          tol=1e-12, maxiter=200)
    r64 = O.true_residual(c["shape"], c["spacing"], c["b"], x64, c["mask"], c["weights"])
    assert r64 <= st["threshold"] * 1.0000001
    assert np.linalg.norm(x64 - xr) <= 1e-6 * np.linalg.norm(xr)


# ------------------------------------------------------------- transfers
def test_flip_box_seed_counts_and_bounds():
    """computeFlipBoxSeed (grid_flip_solver2.test.js:24-67) in 3-D."""
    pos = O.flip_box_seed3([0, 0, 0], [4, 2, 1], [1, 1, 1], ppa=2, jitter=0)
    assert pos.size // 3 == 4 * 2 * 1 * 8
    pos = O.flip_box_seed3([1, 1, 1], [3, 4, 2], [1, 1, 1], ppa=3, jitter=0).reshape(-1, 3)
    assert (pos > [1, 1, 1]).all() and (pos < [3, 4, 2]).all()
    # jitter bounded by 0.2 sub-spacing around the sub-cell centres
    pos = O.flip_box_seed3([0, 0, 0], [2, 2, 2], [1, 1, 1], ppa=2, jitter=0.2).reshape(-1, 3)
    frac = (pos * 2) % 1.0
    assert (np.abs(frac - 0.5) <= 0.2 * 1.0 + 1e-6).all()


def test_seed_jitter_uses_mulberry_sequence():
    """The first particle's jitter is the first three generator draws."""
    d = json.load(open(os.path.join(GOLDEN, "mulberry32.json")))
    r = np.array(d["vectors"][0]["values"][:3])
    pos = O.flip_box_seed3([0, 0, 0], [1, 1, 1], [1, 1, 1], ppa=2, jitter=0.2)
    expect = 0.25 + (r * 2 - 1) * 0.2 * 0.5
    np.testing.assert_array_equal(pos[:3], expect.astype(np.float32))


def test_binning_and_stable_sort():
    rng = np.random.default_rng(0)
    n = (7, 5, 6)
    pos = (rng.random((1000, 3)) * np.array(n)).astype(np.float32).ravel()
    cells = O.cell_index3(pos, [0, 0, 0], [1, 1, 1], n)
    p3 = pos.reshape(-1, 3)
    expect = np.floor(p3[:, 0]) + n[0] * (np.floor(p3[:, 1]) + n[1] * np.floor(p3[:, 2]))
    np.testing.assert_array_equal(cells, expect.astype(np.int32))
    perm, start = O.sort_by_cell(cells, int(np.prod(n)))
    np.testing.assert_array_equal(perm, np.argsort(cells, kind="stable"))
    np.testing.assert_array_equal(start[1:] - start[:-1], np.bincount(cells, minlength=int(np.prod(n))))


def test_p2g_modes_agree_and_reproduce_uniform_field():
    """A uniform particle velocity transfers exactly (weights sum to one) in
    every mode; the gather mode equals the double-precision scatter to f32
    rounding, and the fixed-point mode to its 1/65536 quantum."""
    n = (6, 5, 4)
    h = (1.0 / 6, 1.0 / 6, 1.0 / 6)
    L = np.array(n) * np.array(h)
    rng = np.random.default_rng(1)
    pos = (rng.uniform(0.01, 0.99, (4000, 3)) * L).astype(np.float32).ravel()
    vel = np.tile(np.array([1.5, -0.25, 3.0], np.float32), 4000)
    cells = O.cell_index3(pos, [0, 0, 0], h, n)
    perm, start = O.sort_by_cell(cells, int(np.prod(n)))
    for comp in range(3):
        g, gv = O.p2g_gather3(pos, vel, perm, start, n, h, [0, 0, 0], comp)
        f, fv = O.p2g_float3(pos, vel, n, h, [0, 0, 0], comp)
        x, xv, numer, denom = O.p2g_fixed3(pos, vel, n, h, [0, 0, 0], comp)
        np.testing.assert_array_equal(gv, fv)
        ok = gv == 1
        np.testing.assert_allclose(g[ok], vel[comp], rtol=2e-6)
        np.testing.assert_allclose(g[ok], f[ok], rtol=2e-6)
        np.testing.assert_allclose(x[xv == 1], vel[comp], atol=4 / 65536 * abs(vel[comp]) + 1e-4)
    # random velocities: gather == double scatter to f32 accumulation error
    vel = rng.standard_normal(12000).astype(np.float32)
    for comp in range(3):
        g, gv = O.p2g_gather3(pos, vel, perm, start, n, h, [0, 0, 0], comp)
        f, fv = O.p2g_float3(pos, vel, n, h, [0, 0, 0], comp)
        np.testing.assert_allclose(g, f, atol=1e-5)


def test_trilinear_and_cubic_reproduce_linear_fields():
    """Trilinear (grid_math3.js:61-113) and monotone cubic (137-207) samples of
    a linear field are exact away from the clamped border."""
    n = (8, 8, 8)
    h = (0.5, 0.5, 0.5)
    ou = O.mac_origin([0, 0, 0], h, 0)
    fs = O.face_shape(n, 0)
    ii, jj, kk = np.meshgrid(*[np.arange(s) for s in fs], indexing="ij")
    X = ou[0] + ii * h[0]; Y = ou[1] + jj * h[1]; Z = ou[2] + kk * h[2]
    data = (0.5 * X - 2 * Y + 0.25 * Z).ravel(order="F").astype(np.float32)
    rng = np.random.default_rng(2)
    for _ in range(50):
        p = rng.uniform(1.0, 3.0, 3).astype(np.float32)
        val = O.lib().orc_sample_trilinear3(O._p(data), O._i3(fs), O._f3(ou), O._f3(h), O._f3(p))
        assert abs(val - (0.5 * p[0] - 2 * p[1] + 0.25 * p[2])) < 2e-5
    # the face advection with zero velocity is the identity for any data
    U = rng.standard_normal(int(np.prod(fs))).astype(np.float32)
    V = np.zeros(int(np.prod(O.face_shape(n, 1))), np.float32)
    W = np.zeros(int(np.prod(O.face_shape(n, 2))), np.float32)
    u2, v2, w2 = O.advect_face3(np.zeros_like(U), V, W, n, h, [0, 0, 0], 0.1)
    np.testing.assert_array_equal(u2, 0)
    uo, vo, wo = O.advect_face3(U * 0, V, W, n, h, [0, 0, 0], 0.1)
    np.testing.assert_array_equal(uo, 0)


def test_g2p_closed_forms():
    n = (4, 4, 4)
    h = (0.25,) * 3
    rng = np.random.default_rng(3)
    pos = rng.uniform(0.1, 0.9, (50, 3)).astype(np.float32).ravel()
    vel = rng.standard_normal(150).astype(np.float32)
    grids = [rng.standard_normal(int(np.prod(O.face_shape(n, a)))).astype(np.float32) for a in range(3)]
    # new == old and pure FLIP: velocities unchanged
    out = O.g2p3(pos, vel, *grids, *grids, n, h, [0, 0, 0], flip_ratio=1.0, damping=0.0)
    np.testing.assert_array_equal(out, vel)
    # pure PIC: velocity = the grid sample, independent of the particle's own
    out0 = O.g2p3(pos, vel, *grids, *grids, n, h, [0, 0, 0], flip_ratio=0.0, damping=0.0)
    out1 = O.g2p3(pos, vel * 7, *grids, *grids, n, h, [0, 0, 0], flip_ratio=0.0, damping=0.0)
    np.testing.assert_array_equal(out0, out1)
    # clamp and non-finite guard (grid_flip_solver3.js:903-907)
    big = np.full(150, 1e4, np.float32)
    big[3] = np.nan
    out = O.g2p3(pos, big, *grids, *grids, n, h, [0, 0, 0], flip_ratio=1.0, damping=0.0)
    assert np.abs(out).max() <= 500
    assert (out[3:6] == 0).all() if False else out[3] == 0


def test_two_phase_beta_range():
    n = (6, 6, 6)
    frac = np.zeros(216, np.float32)
    frac[:100] = 1
    bu, bv, bw = O.two_phase_beta3(frac, n, 1.0, 1000.0)
    for b in (bu, bv, bw):
        assert b.min() >= 1.0 - 1e-6 and b.max() <= 1000.0 + 1e-3
    # an all-liquid column gives beta 1, an all-gas one 1000, the interface ~2
    assert np.isclose(bu.min(), 1.0) and np.isclose(bu.max(), 1000.0)
    assert np.any(np.isclose(bv, 1000.0 / 500.5, rtol=1e-6))


# ------------------------------------------------------- learned-guess oracle
def test_unet_conv_matches_scipy_correlate():
    from scipy.signal import correlate
    from oracle import unet_ref as U
    rng = np.random.default_rng(0)
    x = rng.standard_normal((3, 6, 5, 7))
    w = rng.standard_normal((4, 3, 3, 3, 3))
    b = rng.standard_normal(4)
    got = U.conv3d(x, w, b)
    for co in range(4):
        ref = sum(correlate(x[ci], w[co, ci], mode="same", method="direct") for ci in range(3)) + b[co]
        np.testing.assert_allclose(got[co], ref, rtol=1e-12, atol=1e-12)


def test_bf16_rounding_matches_torch():
    """the oracle's round-to-nearest-even BF16 against torch's CPU cast, on
    random values over many binades and on exact ties (low half 0x8000)"""
    import torch
    from oracle import unet_ref as U
    rng = np.random.default_rng(4)
    v = (rng.standard_normal(20000) * np.exp2(rng.integers(-30, 30, 20000))).astype(np.float32)
    bits = v.view(np.uint32)
    ties = ((bits[:2000] & 0xFFFF0000) | 0x8000).view(np.float32)
    for a in (v, ties, np.float32([0.0, -0.0, 1.0, -2.5, 65504.0])):
        ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float64).numpy()
        np.testing.assert_array_equal(U.to_bf16(a), ref)


def test_unet_bf16_forward_is_the_fp32_forward_rounded():
    """the BF16 restatement stays within BF16 rounding of the float64
    forward and differs from it (the rounding happens)"""
    from oracle import unet_ref as U
    rng = np.random.default_rng(6)
    p = U.init_params(2)
    for k in p:
        if k.endswith(".b"):
            p[k] = rng.standard_normal(p[k].shape) * 0.1
    x = rng.standard_normal((3, 8, 8, 16))
    y, _ = U.forward(p, x)
    yb = U.forward_bf16(p, x)
    r = np.linalg.norm(yb - y) / np.linalg.norm(y)
    assert 1e-5 < r < 2e-2, r


def test_unet_gradients_match_finite_differences():
    from oracle import unet_ref as U
    assert 600_000 < U.param_count() < 720_000
    rng = np.random.default_rng(1)
    p = U.init_params(3)
    for k in p:
        if k.endswith(".b"):
            p[k] = rng.standard_normal(p[k].shape) * 0.1
    x = rng.standard_normal((3, 8, 8, 8))
    t = rng.standard_normal((1, 8, 8, 8))
    y, cache = U.forward(p, x)
    loss, gy = U.mse_loss(y, t)
    g, gx = U.backward(p, cache, gy)
    eps = 1e-6
    for name in ["enc0a.w", "mid.w", "dec2.b", "dec0a.w", "out.w", "enc1b.b"]:
        arr = p[name]
        for _ in range(3):
            idx = tuple(rng.integers(0, s) for s in arr.shape)
            old = arr[idx]
            arr[idx] = old + eps
            lp = U.mse_loss(U.forward(p, x)[0], t)[0]
            arr[idx] = old - eps
            lm = U.mse_loss(U.forward(p, x)[0], t)[0]
            arr[idx] = old
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - g[name][idx]) <= 1e-6 + 1e-4 * abs(fd), (name, idx, fd, g[name][idx])
    idx = (1, 3, 4, 5)
    old = x[idx]
    x[idx] = old + eps
    lp = U.mse_loss(U.forward(p, x)[0], t)[0]
    x[idx] = old - eps
    lm = U.mse_loss(U.forward(p, x)[0], t)[0]
    x[idx] = old
    assert abs((lp - lm) / (2 * eps) - gx[idx]) <= 1e-6 + 1e-4 * abs(gx[idx])
