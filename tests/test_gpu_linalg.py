"""createDotReducer and createConjugateGradientSolver on the GPU.

* The dot reducer's probes of examples/31-conjugate-gradient-3d/main.js:168-238
  on a 6x4x5 field: a single cell set to one, reduced against ones, gives 1 in
  the lane of its first index and 0 in every other lane; a ramp holding each
  cell's flat index sums to 119*120/2 = 7140.  Plus 1-D (64 strided lanes)
  and 2-D shapes against numpy, exact on integer-valued data.
* CG: examples/04-conjugate-gradient (A = diag(1..8), b = 1, x -> 1/(i+1)),
  and a Laplacian system whose answer is checked against the oracle's FP64
  solve (CG in FP32 reaches the requested absolute tolerance).
"""
import numpy as np
import pytest

from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def test_dot_reducer_single_cell_probes_and_ramp(flux):
    F = flux
    shape = (6, 4, 5)
    ones = F.array_n("float", shape).fill(1.0)
    delta = F.array_n("float", shape)
    red = F.create_dot_reducer(shape, delta, ones)
    assert red.lanes == 6
    for (pi, pj, pk) in [(0, 0, 0), (0, 0, 1), (0, 1, 0), (5, 3, 4), (2, 3, 1)]:
        d = np.zeros(120, np.float32)
        d[pi + 6 * (pj + 4 * pk)] = 1.0
        delta.from_array(d)
        assert red.read() == 1.0
        lanes = red.partial.to_array()
        lit = [(q, v) for q, v in enumerate(lanes) if v != 0]
        assert lit == [(pi, 1.0)], (pi, pj, pk, lanes)
    ramp = F.array_n("float", shape, np.arange(120, dtype=np.float32))
    assert F.create_dot_reducer(shape, ramp, ones).read() == 7140.0


@pytest.mark.parametrize("shape", [(1000,), (37,), (300, 7), (513, 3, 4), (256, 256, 17)])
def test_dot_reducer_matches_numpy(flux, shape):
    F = flux
    n = int(np.prod(shape))
    rng = np.random.default_rng(n)
    a = rng.integers(-50, 50, n).astype(np.float32)
    b = rng.integers(-50, 50, n).astype(np.float32)
    A, B = F.array_n("float", shape, a), F.array_n("float", shape, b)
    red = F.create_dot_reducer(shape, A, B)
    assert red.read() == float(np.dot(a.astype(np.int64), b.astype(np.int64)))
    lanes = red.lanes
    if len(shape) == 1:
        assert lanes == min(shape[0], 64)
        want = np.array([np.dot(a[l::lanes].astype(np.int64), b[l::lanes].astype(np.int64)) for l in range(lanes)])
    else:
        assert lanes == shape[0]
        m = (a.astype(np.int64) * b.astype(np.int64)).reshape(-1, shape[0])
        want = m.sum(axis=0)
    np.testing.assert_array_equal(red.partial.to_array(), want.astype(np.float32))


def test_dot_reducer_rejects_bad_shapes(flux):
    F = flux
    a = F.array_n("float", (8,))
    with pytest.raises(F.FluxError, match="1D/2D/3D"):
        F.create_dot_reducer((2, 2, 2, 1), a, a)
    with pytest.raises(F.FluxError):
        F.create_dot_reducer((9,), a, a)


def test_cg_diagonal_example(flux):
    F = flux
    N = 8
    b = F.array_n("float", N, np.ones(N, np.float32))
    x = F.array_n("float", N, np.zeros(N, np.float32))
    d = F.array_n("float", N, np.arange(1, N + 1, dtype=np.float32))
    cg = F.create_cg_solver(d, b, x)
    assert cg.solve(1e-5, 20)
    np.testing.assert_allclose(x.to_array(), 1.0 / np.arange(1, N + 1), atol=1e-3)


def test_cg_laplacian_against_oracle(flux, frame110):
    F, c = flux, frame110
    sh = c["shape"]
    m = F.array_n("u8", sh, c["mask"])
    w = [F.array_n("float", F.face_shape(sh, a), c["weights"][a]) for a in range(3)]
    op = F.create_laplacian_operator(sh, c["spacing"], m, w)
    b = F.array_n("float", sh, c["b"])
    x = F.array_n("float", sh)
    cg = F.create_cg_solver(op, b, x)
    ok = cg.solve(1e-3, 2000)
    got = x.to_array().astype(np.float64)
    r = np.linalg.norm(c["b"] - O.apply(sh, c["spacing"], got, c["mask"], c["weights"]))
    print("cg iterations", cg.iterations, "residual", r)
    assert ok and r < 2e-3
    mg = O.Multigrid(sh, c["spacing"], levels=4, mask=c["mask"], coarse_mask=c["vent"], weights=c["weights"])
    xt, _ = O.pcg_solve(sh, c["spacing"], c["b"], mask=c["mask"], weights=c["weights"], mg=mg, tol=1e-12,
                        maxiter=400)
    assert np.linalg.norm(got - xt) / np.linalg.norm(xt) < 1e-2
    with pytest.raises(F.FluxError, match="shape mismatch"):
        F.create_cg_solver(op, b, F.array_n("float", (4, 4, 4)))
