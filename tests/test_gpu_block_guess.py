"""The coarse correction along a learned guess (flux_pcg_block_start,
no reference counterpart) on the GPU against the oracle's dense Galerkin
solve (pyoracle.block_galerkin_start), and its exactness: a guess equal to
the solution up to one scale per block is corrected to the solution, so the
PCG that follows needs (almost) no iterations."""
import numpy as np
import pytest

from oracle import pyoracle as O
from tsl_array_n_amd import scenes

pytestmark = pytest.mark.gpu


def system(F, n, seed, tol=1e-6):
    sc = scenes.random_scene(n, seed)
    shape = sc["shape"]
    beta = [F.array_n("float", F.face_shape(shape, a), sc["beta"][a]) for a in range(3)]
    act = F.array_n("u8", shape, sc["active"])
    tgt = F.array_n("float", shape, sc["target"])
    ps = F.create_grid_pressure_solver3(resolution=shape, grid_spacing=sc["spacing"], dirichlet=(act, tgt),
                                        face_weights=beta, multigrid={"levels": 3}, tolerance=tol)
    vel = [F.array_n("float", F.face_shape(shape, a), sc["vel"][a]) for a in range(3)]
    return sc, ps, vel


def block_ids(shape, block):
    n0, n1, n2 = shape
    nbx, nby = -(-n0 // block), -(-n1 // block)
    idx = np.arange(n0 * n1 * n2)
    return (idx % n0) // block + nbx * (((idx // n0) % n1) // block + nby * ((idx // (n0 * n1)) // block))


@pytest.mark.parametrize("block", [4, 5, 16])
def test_block_guess_matches_oracle(flux, block):
    F = flux
    sc, ps, vel = system(F, 16, 303)
    shape, h = sc["shape"], sc["spacing"]
    _, ref, rvel = system(F, 16, 303, tol=1e-9)
    ref.project(rvel)
    p_star = ref.pressure.to_array().astype(np.float64)
    ps.assemble(vel)
    mask = ps.dirichlet_mask.to_array()
    weights = [w.to_array() for w in ps.weights]
    b = ps.b.to_array().astype(np.float64)
    x0 = ps.pressure.to_array().astype(np.float64)
    g = (p_star * (1 + 0.3 * np.random.default_rng(1).standard_normal(p_star.size))).astype(np.float32)
    gf = F.array_n("float", shape, g)
    it = ps.block_guess(gf, block)
    x_gpu = ps.pressure.to_array().astype(np.float64)
    x_ref, alpha = O.block_galerkin_start(shape, h, b, x0, g, block, mask, weights)
    err = np.linalg.norm(x_gpu - x_ref) / np.linalg.norm(x_ref)
    print("block", block, "coarse iterations", it, "rel", err)
    assert err <= 1e-5
    d = ps.solve(vel)
    assert d["converged"] and d["residual"] <= 1e-6 * d["bNorm"]


def test_blockwise_scaled_solution_is_recovered(flux):
    F = flux
    n, block = 64, 8
    sc, cold, vel = system(F, n, 404)
    d0 = cold.project(vel)
    assert d0["converged"]
    p = cold.pressure.to_array().astype(np.float64)
    shape = sc["shape"]
    bid = block_ids(shape, block)
    s = np.random.default_rng(5).uniform(0.5, 2.0, bid.max() + 1)
    g = F.array_n("float", shape, (p / s[bid]).astype(np.float32))
    _, ps, vel = system(F, n, 404)
    ps.assemble(vel)
    it = ps.block_guess(g, block)
    d = ps.solve(vel)
    print("cold", d0["iterations"], "after the block correction", d["iterations"], "coarse iterations", it)
    assert d["converged"] and d["iterations"] <= 2


def test_block_guess_checks_its_arguments(flux):
    F = flux
    sc, ps, vel = system(F, 16, 303)
    ps.assemble(vel)
    with pytest.raises(F.FluxError, match="block edge"):
        ps.block_guess(F.array_n("float", sc["shape"]), 0)
    with pytest.raises(F.FluxError, match="blockStart"):
        ps.block_guess(F.array_n("float", (8, 8, 8)), 4)
