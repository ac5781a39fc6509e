"""The fused one-pass red-black sweeps against the two-launch form (GPU tests).

Both fused kernels (FLUX_FUSED_SWEEP=1: LDS tiles on every wide level; =2:
the register form on whole rows of the finest level) do the same arithmetic
per cell as two single-colour launches, so one V-cycle must agree bit for
bit (FP32 and FP16 storage); whole projections agree to the PCG tolerance
(their r.z partials are summed over different block partitions) in the same
number of iterations.
"""
import os

import numpy as np
import pytest

import tsl_array_n_amd as F
from tsl_array_n_amd import scenes

pytestmark = pytest.mark.gpu


def build_mg(F, sc, levels, storage, fused):
    old = os.environ.get("FLUX_FUSED_SWEEP")
    os.environ["FLUX_FUSED_SWEEP"] = fused if isinstance(fused, str) else ("1" if fused else "0")
    try:
        n = sc["shape"]
        vel = [F.array_n("float", F.face_shape(n, a), sc["vel"][a]) for a in range(3)]
        beta = [F.array_n("float", F.face_shape(n, a), sc["beta"][a]) for a in range(3)]
        act = F.array_n("u8", n, sc["active"])
        tgt = F.array_n("float", n, sc["target"])
        ps = F.create_grid_pressure_solver3(resolution=n, grid_spacing=sc["spacing"], dirichlet=(act, tgt),
                                            face_weights=beta, multigrid={"levels": levels, "storage": storage},
                                            tolerance=1e-6)
        ps.assemble(vel)
        mg = F.create_multigrid_preconditioner(n, sc["spacing"], dirichlet_mask=ps.dirichlet_mask,
                                               coarse_dirichlet_mask=ps.vent_mask, face_weights=ps.weights,
                                               levels=levels, storage=storage)
    finally:
        if old is None:
            del os.environ["FLUX_FUSED_SWEEP"]
        else:
            os.environ["FLUX_FUSED_SWEEP"] = old
    return ps, vel, mg


def box_scene(shape, seed):
    """random two-phase blobs, random velocity with closed walls, one pinned cell"""
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    frac = np.zeros((nz, ny, nx), np.float32)
    for _ in range(4):
        c = rng.uniform(0.2, 0.8, 3) * np.array(shape)
        r = rng.uniform(5, 15)
        frac[(i - c[0]) ** 2 + (j - c[1]) ** 2 + (k - c[2]) ** 2 < r * r] = 1.0
    beta = scenes.two_phase_beta(frac.ravel(), shape, 1.0, 300.0)
    vel = []
    for a in range(3):
        fs = list(shape)
        fs[a] += 1
        v = rng.standard_normal(fs[::-1]).astype(np.float32)
        sl0 = [slice(None)] * 3
        sl1 = [slice(None)] * 3
        sl0[2 - a] = 0
        sl1[2 - a] = -1
        v[tuple(sl0)] = 0
        v[tuple(sl1)] = 0
        vel.append(v.ravel())
    active = np.zeros(nx * ny * nz, np.uint8)
    active[nx // 2 + nx * (ny - 1)] = 1
    return {"shape": shape, "spacing": (1.0 / nx,) * 3, "vel": vel, "beta": beta, "active": active,
            "target": np.zeros(nx * ny * nz, np.float32)}


@pytest.mark.parametrize("mode", ["1", "2"])
@pytest.mark.parametrize("storage", ["f32", "f16"])
@pytest.mark.parametrize("shape", [(64, 64, 64), (128, 64, 48), (256, 24, 20), (96, 20, 12)])
def test_fused_vcycle_is_bit_identical(flux, storage, shape, mode):
    """including whole 256-cell rows, a partial last row tile (20 rows in
    tiles of 8) and a short last z-chunk (12 planes in chunks of 8)"""
    F = flux
    sc = scenes.dam_break_system(64) if shape == (64, 64, 64) else box_scene(shape, 5)
    outs = []
    for fused in ("0", mode):
        ps, vel, mg = build_mg(F, sc, 3, storage, fused)
        z = F.array_n("float", shape)
        mg(ps.b, z)
        outs.append(z.to_array())
    assert np.all(np.isfinite(outs[0])) and np.any(outs[0] != 0)
    np.testing.assert_array_equal(outs[1], outs[0])


@pytest.mark.parametrize("mode", ["1", "2"])
def test_fused_projection_matches_two_launch(flux, mode):
    F = flux
    sc = scenes.dam_break_system(128)
    res = []
    for fused in ("0", mode):
        os.environ["FLUX_FUSED_SWEEP"] = fused
        try:
            n = sc["shape"]
            vel = [F.array_n("float", F.face_shape(n, a), sc["vel"][a]) for a in range(3)]
            beta = [F.array_n("float", F.face_shape(n, a), sc["beta"][a]) for a in range(3)]
            act = F.array_n("u8", n, sc["active"])
            tgt = F.array_n("float", n, sc["target"])
            ps = F.create_grid_pressure_solver3(resolution=n, grid_spacing=sc["spacing"], dirichlet=(act, tgt),
                                                face_weights=beta, multigrid={"levels": 5}, tolerance=1e-6)
            d = ps.project(vel)
        finally:
            del os.environ["FLUX_FUSED_SWEEP"]
        res.append((d, ps.pressure.to_array()))
    print(res[0][0], res[1][0])
    assert res[0][0]["converged"] and res[1][0]["converged"]
    assert abs(res[0][0]["iterations"] - res[1][0]["iterations"]) <= 1
    p0, p1 = res[0][1], res[1][1]
    assert np.linalg.norm(p1 - p0) <= 1e-5 * np.linalg.norm(p0)


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
@pytest.mark.parametrize("shape,levels", [((64, 64, 64), 4), ((64, 32, 48), 4), ((96, 64, 64), 5)])
@pytest.mark.parametrize("env", ["FLUX_VCYCLE_TAIL", "FLUX_CORRECT_PAIRS"])
def test_vcycle_variants_are_bit_identical(flux, storage, shape, levels, env):
    """The last two levels in one workgroup (k_vcycle_tail) against their
    separate launches, and the paired prolongation-correction against the
    one-cell-per-thread kernel: one V-cycle agrees bit for bit."""
    F = flux
    sc = scenes.dam_break_system(64) if shape == (64, 64, 64) else box_scene(shape, 7)
    outs = []
    for flag in ("0", "1"):
        old = os.environ.get(env)
        os.environ[env] = flag
        try:
            ps, vel, mg = build_mg(F, sc, levels, storage, False)
        finally:
            if old is None:
                del os.environ[env]
            else:
                os.environ[env] = old
        z = F.array_n("float", shape)
        mg(ps.b, z)
        outs.append(z.to_array())
    assert np.all(np.isfinite(outs[0])) and np.any(outs[0] != 0)
    np.testing.assert_array_equal(outs[1], outs[0])
