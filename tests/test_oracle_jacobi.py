"""Pins of the oracle's damped-Jacobi smoother (row N1:
sandbox/jacobi-smoother/multigrid_jacobi.js:622-642, 1271-1324).

* One sweep equals x + omega (b - A x) / diag(A) with A assembled face by
  face (the edge assembly of test_oracle, independent of the row stencil).
* The Jacobi V-cycle is symmetric (the reference's reason for wanting it:
  "symmetric by construction") and negative definite on the free cells.
* Odd sweep counts are rejected (multigrid_jacobi.js:1273-1276).
* The iteration counts the reference recorded for its operator (omega = 2/3
  and 0.8 both close to red-black's count, undamped far worse,
  multigrid_jacobi.js:629-631) hold in kind here: on a 32^3 constant-
  coefficient Poisson problem, omega = 2/3 and 0.8 stay within a few
  iterations of red-black and omega = 1 needs several times as many.
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from test_oracle import edge_assembly, _random_system


@pytest.mark.parametrize("shape,spacing", [((5, 4, 3), (1.0, 0.5, 2.0)), ((6, 5), (0.7, 1.3))])
def test_jacobi_sweep_matches_assembled_operator(shape, spacing):
    mask, weights = _random_system(shape, 4, True, True)
    A = edge_assembly(shape, spacing, mask, weights)
    N = A.shape[0]
    rng = np.random.default_rng(5)
    x = rng.standard_normal(N)
    b = rng.standard_normal(N)
    g = O.make_grid(shape, spacing)
    out = np.zeros(N)
    w = O._w([np.asarray(a, np.float64) for a in weights])
    O.lib().orc_jacobi_f64(O.C.byref(g), O._p(x), O._p(out), O._p(b), O._p(mask), w[1], O.C.c_double(0.7))
    want = x + 0.7 * (b - A @ x) / np.diag(A)
    np.testing.assert_allclose(out, want, rtol=1e-12, atol=1e-12)


def test_jacobi_vcycle_is_symmetric_and_definite():
    n = (16, 8, 8)
    N = 16 * 8 * 8
    rng = np.random.default_rng(22)
    mask = np.zeros((8, 8, 16), np.uint8)
    mask[:, :4, :] = 1
    mask = mask.ravel()
    weights = [rng.uniform(0.5, 2.0, int(np.prod(O.face_shape(n, a)))) for a in range(3)]
    mg = O.Multigrid(n, (1, 1, 1), levels=3, n_coarsest=20, mask=mask, coarse_mask=mask, weights=weights,
                     smoother="jacobi")
    free = mask == 0
    u = rng.standard_normal(N) * free
    v = rng.standard_normal(N) * free
    lhs, rhs = np.dot(u, mg.apply(v)), np.dot(mg.apply(u), v)
    assert abs(lhs - rhs) <= 1e-9 * abs(lhs)
    assert np.dot(u, mg.apply(u)) < 0
    # red-black with an unreversed order would not be; Jacobi differs from it
    rb = O.Multigrid(n, (1, 1, 1), levels=3, n_coarsest=20, mask=mask, coarse_mask=mask, weights=weights)
    assert np.linalg.norm(rb.apply(v) - mg.apply(v)) > 1e-3 * np.linalg.norm(rb.apply(v))


def test_jacobi_rejects_odd_sweep_counts():
    with pytest.raises(ValueError, match="must be even"):
        O.Multigrid((8, 8, 8), (1, 1, 1), levels=2, nu_down=3, smoother="jacobi")


def test_jacobi_damping_factors_behave_as_reference_recorded():
    n = (32, 32, 32)
    h = (1.0 / 32,) * 3
    N = 32 ** 3
    rng = np.random.default_rng(3)
    b = rng.standard_normal(N)
    b -= b.mean()
    mask = np.zeros(N, np.uint8)
    mask[0] = 1  # one pinned cell: an otherwise all-Neumann box
    b[0] = 0.0
    its = {}
    for name, kw in [("rbgs", {}), ("j2/3", {"smoother": "jacobi"}), ("j0.8", {"smoother": "jacobi", "omega": 0.8}),
                     ("j1", {"smoother": "jacobi", "omega": 1.0})]:
        mg = O.Multigrid(n, h, levels=4, mask=mask, **kw)
        _, st = O.pcg_solve(n, h, b, mask=mask, mg=mg, tol=1e-6, maxiter=400)
        assert st["converged"] or name == "j1", (name, st)
        its[name] = st["iterations"]
    print(its)
    assert its["j2/3"] <= its["rbgs"] * 2 and its["j0.8"] <= its["rbgs"] * 2
    assert its["j1"] >= 2 * max(its["j2/3"], its["j0.8"])
