"""Pins of the coarse-space PCG restatement (pyoracle.pcg_coarse_space).

* With no coarse vectors it is the oracle's mixed PCG (the GPU plain path's
  arithmetic): same iteration count and solution on the captured frame.
* On a system pinned at one cell (an otherwise all-Neumann box), the
  free-cell indicator is the near-null mode: the corrected preconditioner
  needs fewer iterations, reaches the same FP64 tolerance and the same
  solution.
* The correction alone (a V-cycle of no sweeps, so M^-1 = Phi E^-1 Phi^T)
  with the exact solution in the space: M^-1 b = x* (x*.b)/(x*.A x*) = x*,
  and PCG stops after one iteration; the same space without x* does not.
"""
import numpy as np

from oracle import pyoracle as O


def test_no_vectors_is_the_mixed_pcg(frame110):
    c = frame110
    mg = O.Multigrid(c["shape"], c["spacing"], levels=4, mask=c["mask"], coarse_mask=c["vent"],
                     weights=c["weights"], suffix="f32")
    x, st = O.pcg_coarse_space(c["shape"], c["spacing"], c["b"], c["mask"], c["weights"], mg, [],
                               free_indicator=False, tol=1e-6)
    xr, _, ref = O.pcg_solve_mixed(c["shape"], c["spacing"], c["b"], mask=c["mask"], weights=c["weights"], mg=mg,
                                   tol=1e-6, maxiter=200)
    assert st["converged"] and abs(st["iterations"] - ref["iterations"]) <= 1
    assert np.linalg.norm(x - xr) <= 1e-5 * np.linalg.norm(xr)


def _pinned_box(n=32, seed=4):
    rng = np.random.default_rng(seed)
    shape, h = (n, n, n), (1.0 / n,) * 3
    N = n ** 3
    mask = np.zeros(N, np.uint8)
    mask[(n // 2) + n * ((n - 1) + n * (n // 2))] = 1
    w = []
    heavy = np.zeros((n, n, n), bool)
    heavy[:, : n // 2, : n // 3] = True  # [k][j][i]: a dense block
    for a in range(3):
        fs = list(shape)
        fs[a] += 1
        w.append(np.ones(int(np.prod(fs)), np.float32))
    rho = np.where(heavy, 1000.0, 1.0).astype(np.float32)
    for a in range(3):
        ax = 2 - a
        m = rho.shape[ax]
        lo = np.take(rho, np.clip(np.arange(m + 1) - 1, 0, m - 1), axis=ax)
        hi = np.take(rho, np.clip(np.arange(m + 1), 0, m - 1), axis=ax)
        w[a] = (1000.0 / np.maximum(0.5 * (lo + hi), 1e-4)).astype(np.float32).ravel()
    b = rng.standard_normal(N).astype(np.float32)
    b[mask == 1] = 0
    b -= np.float32(b[mask == 0].mean()) * (mask == 0)
    return shape, h, mask, w, b


def test_free_indicator_cuts_iterations_on_a_pinned_box():
    shape, h, mask, w, b = _pinned_box()
    mg = O.Multigrid(shape, h, levels=4, mask=mask, weights=w, suffix="f32")
    x0, s0 = O.pcg_coarse_space(shape, h, b, mask, w, mg, [], free_indicator=False)
    x1, s1 = O.pcg_coarse_space(shape, h, b, mask, w, mg, [], free_indicator=True)
    print(s0, s1)
    assert s0["converged"] and s1["converged"]
    assert s1["iterations"] < s0["iterations"]
    assert np.linalg.norm(x1 - x0) <= 1e-4 * np.linalg.norm(x0)


def test_exact_direction_in_the_space_solves_in_one_iteration():
    shape, h, mask, w, _ = _pinned_box(16)
    rng = np.random.default_rng(1)
    xs = rng.standard_normal(16 ** 3) * (mask == 0)
    b = O.apply(shape, h, xs, mask, w, suffix="f64").astype(np.float32)
    none = O.Multigrid(shape, h, levels=1, nu_down=0, nu_up=0, n_coarsest=0, n_final=0, mask=mask, weights=w,
                       suffix="f32")
    assert not none.apply(b).any()
    x, st = O.pcg_coarse_space(shape, h, b, mask, w, none, [xs], free_indicator=True, tol=1e-5)
    assert st["converged"] and st["iterations"] == 1, st
    assert np.linalg.norm(x - xs) <= 1e-5 * np.linalg.norm(xs)
    _, st2 = O.pcg_coarse_space(shape, h, b, mask, w, none, [], free_indicator=True, tol=1e-5, maxiter=3)
    assert not st2["converged"]
