"""Pins of the oracle's brick smoother (orc_brick_relax: red-black sweeps
brick by brick, values outside a brick held at their values on entry; the
library's one-pass smoother, no counterpart in the reference).

* One brick covering the grid is the reference's relax (multigrid.js:626-640):
  `sweeps` colour pairs of orc_relax, bit for bit in f64 (both colour orders).
* Bricks of one cell and one sweep are one damped-Jacobi sweep from the
  field on entry (every neighbour lies outside the cell's brick): bit for
  bit against orc_jacobi, itself pinned against a face-by-face assembly.
* The shift: bricks twice the grid's size shifted by half a brick again
  make one brick covering the grid, so the shifted pass equals global
  red-black too; any other origin would cut the grid.
* A brick's cells see other bricks only through the field on entry:
  changing the entry field away from a brick's faces changes no cell of the
  neighbouring brick; changing it on the face layer reaches the neighbour's
  face layer and, through its sweeps, the cells behind it.
* The V-cycle with two brick passes per smoothing (the second shifted, the
  post-smoother the adjoint) is symmetric and negative definite on the free
  cells, and it stops being symmetric when the post-smoothing is not
  reversed.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as O
from test_oracle import _random_system


@pytest.fixture(autouse=True)
def _restore_brick():
    yield
    O.set_brick()
    O.set_post_forward(False)


def _brick(shape, spacing, xin, b, mask, weights, sweeps, reversed, dims, shifted):
    g = O.make_grid(shape, spacing)
    out = np.zeros_like(b)
    w = O._w([np.asarray(a, np.float64) for a in weights])
    B = (C.c_int * 3)(*dims)
    O.lib().orc_brick_relax_shape_f64(O.C.byref(g), O._p(xin), O._p(out), O._p(b), O._p(mask), w[1],
                                      O.C.c_double(1.0), int(sweeps), int(reversed), B, int(shifted))
    return out


def _global_rbgs(shape, spacing, x, b, mask, weights, sweeps, reversed):
    x = x.copy()
    for _ in range(sweeps):
        for col in ((1, 0) if reversed else (0, 1)):
            x = O.relax(shape, spacing, x, b, col, mask=mask, weights=weights)
    return x


@pytest.mark.parametrize("reversed", [0, 1])
def test_one_brick_is_global_red_black(reversed):
    shape, spacing = (12, 6, 10), (1.0, 0.5, 2.0)
    mask, weights = _random_system(shape, 3, True, True)
    rng = np.random.default_rng(8)
    N = int(np.prod(shape))
    x, b = rng.standard_normal(N), rng.standard_normal(N)
    got = _brick(shape, spacing, x, b, mask, weights, 3, reversed, (64, 8, 16), 0)
    want = _global_rbgs(shape, spacing, x, b, mask, weights, 3, reversed)
    assert np.array_equal(got, want)
    # the shifted pass: bricks of 2n shifted by n cover the grid with one brick
    got = _brick(shape, spacing, x, b, mask, weights, 3, reversed, (24, 12, 20), 1)
    assert np.array_equal(got, want)
    # a real cut differs
    cut = _brick(shape, spacing, x, b, mask, weights, 3, reversed, (6, 6, 10), 0)
    assert not np.array_equal(cut, want)


def test_single_cell_bricks_are_one_jacobi_sweep():
    shape, spacing = (7, 5, 4), (0.7, 1.0, 1.3)
    mask, weights = _random_system(shape, 6, True, True)
    rng = np.random.default_rng(9)
    N = int(np.prod(shape))
    x, b = rng.standard_normal(N), rng.standard_normal(N)
    got = _brick(shape, spacing, x, b, mask, weights, 1, 0, (1, 1, 1), 0)
    g = O.make_grid(shape, spacing)
    want = np.zeros(N)
    w = O._w([np.asarray(a, np.float64) for a in weights])
    O.lib().orc_jacobi_f64(O.C.byref(g), O._p(x), O._p(want), O._p(b), O._p(mask), w[1], O.C.c_double(1.0))
    assert np.array_equal(got, want)


def test_bricks_only_see_each_other_through_the_entry_field():
    shape, spacing = (16, 8, 8), (1.0, 1.0, 1.0)
    mask, weights = _random_system(shape, 2, False, True)
    rng = np.random.default_rng(4)
    N = int(np.prod(shape))
    x, b = rng.standard_normal(N), rng.standard_normal(N)
    dims = (8, 8, 8)  # two bricks along x: i < 8 and i >= 8
    base = _brick(shape, spacing, x, b, mask, weights, 2, 0, dims, 0)
    x2 = x.reshape(8, 8, 16).copy()
    x2[:, :, 2:6] += 1.0  # entry values deep inside the first brick
    got = _brick(shape, spacing, x2.ravel(), b, mask, weights, 2, 0, dims, 0).reshape(8, 8, 16)
    assert np.array_equal(got[:, :, 8:], base.reshape(8, 8, 16)[:, :, 8:])
    x3 = x.reshape(8, 8, 16).copy()
    x3[:, :, 7] += 1.0  # the first brick's face layer: only the second brick's face layer sees it
    got = _brick(shape, spacing, x3.ravel(), b, mask, weights, 2, 0, dims, 0).reshape(8, 8, 16)
    ref = base.reshape(8, 8, 16)
    assert not np.array_equal(got[:, :, 8], ref[:, :, 8])
    assert not np.array_equal(got[:, :, 9:], ref[:, :, 9:])  # ... and it spreads inside that brick


def test_brick_vcycle_is_symmetric_and_definite():
    n = (128, 16, 16)
    N = int(np.prod(n))
    rng = np.random.default_rng(31)
    mask = np.zeros((16, 16, 128), np.uint8)
    mask[:, 12:, :] = 1
    mask = mask.ravel()
    weights = [rng.uniform(0.5, 2.0, int(np.prod(O.face_shape(n, a)))) for a in range(3)]
    mg = O.Multigrid(n, (1, 1, 1), levels=3, nu_down=3, nu_up=3, n_final=3, n_coarsest=200, mask=mask,
                     coarse_mask=mask, weights=weights, smoother="brick")
    free = mask == 0
    u = rng.standard_normal(N) * free
    v = rng.standard_normal(N) * free
    lhs, rhs = np.dot(u, mg.apply(v)), np.dot(mg.apply(u), v)
    assert abs(lhs - rhs) <= 1e-9 * abs(lhs)  # the coarsest level is solved to convergence
    assert np.dot(u, mg.apply(u)) < 0 and np.dot(v, mg.apply(v)) < 0
    # red-black on the whole level differs (the level is cut into bricks)
    rb = O.Multigrid(n, (1, 1, 1), levels=3, nu_down=3, nu_up=3, n_final=3, n_coarsest=200, mask=mask,
                     coarse_mask=mask, weights=weights)
    assert np.linalg.norm(rb.apply(v) - mg.apply(v)) > 1e-3 * np.linalg.norm(rb.apply(v))
    # post-smoothing in the pre-smoothing order breaks the symmetry
    O.set_post_forward(True)
    fw = O.Multigrid(n, (1, 1, 1), levels=3, nu_down=3, nu_up=3, n_final=3, n_coarsest=200, mask=mask,
                     coarse_mask=mask, weights=weights, smoother="brick")
    lhs, rhs = np.dot(u, fw.apply(v)), np.dot(fw.apply(u), v)
    assert abs(lhs - rhs) > 1e-4 * abs(lhs)


def test_brick_pcg_converges_on_dam_break():
    from tsl_array_n_amd import scenes
    sc = scenes.dam_break_system(32)
    sh, h = sc["shape"], sc["spacing"]
    A = O.pressure3_assemble(sh, h, sc["vel"], active=sc["active"], target=sc["target"], beta=sc["beta"],
                             suffix="f32")
    its = {}
    for sm in ("rbgs", "brick"):
        mg = O.Multigrid(sh, h, levels=3, nu_down=4, nu_up=4, n_final=4, mask=A["mask"], coarse_mask=A["vent"],
                         weights=A["weights"], suffix="f32", smoother=sm)
        _, _, st = O.pcg_solve_mixed(sh, h, A["b"], mask=A["mask"], weights=A["weights"], mg=mg, tol=1e-6,
                                     maxiter=200)
        assert st["converged"], (sm, st)
        its[sm] = st["iterations"]
    assert its["brick"] <= its["rbgs"] + 4, its
