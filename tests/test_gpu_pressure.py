"""createGridPressureSolver3 on the GPU against the oracle (GPU tests).

The oracle composes the same stages (updateDirichletFields, build system,
MGPCG in the HIP path's arithmetic, circuit breaker, correction) on the CPU.
Tolerance: relative L2 <= 1e-5 for b, pressure and the corrected velocity;
the mask and vent mask are exact.
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from tsl_array_n_amd import scenes

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def oracle_project(sc, levels, tol, maxiter, collider=None):
    n, h = sc["shape"], sc["spacing"]
    A = O.pressure3_assemble(n, h, sc["vel"], active=sc["active"], target=sc["target"], beta=sc["beta"], cw=collider,
                             suffix="f32")
    mg = O.Multigrid(n, h, levels=levels, mask=A["mask"], coarse_mask=A["vent"], weights=A["weights"], suffix="f32")
    x, x64, st = O.pcg_solve_mixed(n, h, A["b"], x0=A["pressure"], mask=A["mask"], weights=A["weights"], mg=mg,
                                   tol=tol, maxiter=maxiter)
    out = O.pressure3_correct(n, h, x, sc["vel"], beta=sc["beta"], cw=collider, suffix="f32")
    return A, x, x64, st, out


def gpu_project(F, sc, levels, tol, collider=None, **kw):
    n, h = sc["shape"], sc["spacing"]
    vel = [F.array_n("float", F.face_shape(n, a), sc["vel"][a]) for a in range(3)]
    beta = [F.array_n("float", F.face_shape(n, a), sc["beta"][a]) for a in range(3)]
    cw = None
    if collider is not None:
        cw = [F.array_n("float", F.face_shape(n, a), collider[a]) for a in range(3)]
    act = F.array_n("u8", n, sc["active"])
    tgt = F.array_n("float", n, sc["target"])
    ps = F.create_grid_pressure_solver3(resolution=n, grid_spacing=h, dirichlet=(act, tgt), face_weights=beta,
                                        collider_weights=cw, multigrid={"levels": levels}, tolerance=tol, **kw)
    diag = ps.project(vel)
    return ps, vel, diag


@pytest.mark.parametrize("n,levels", [(32, 4), (64, 5)])
def test_dam_break_projection_matches_oracle(flux, n, levels):
    F = flux
    sc = scenes.dam_break_system(n)
    ps, vel, diag = gpu_project(F, sc, levels, 1e-6)
    A, x, x64, st, out = oracle_project(sc, levels, 1e-6, int(np.ceil(1.5 * n)))
    print("gpu", diag, "oracle", st)
    assert diag["converged"] and st["converged"]
    assert not diag["rejected"]
    assert abs(diag["iterations"] - st["iterations"]) <= 1
    assert diag["residual"] <= 1e-6 * diag["bNorm"]
    np.testing.assert_array_equal(ps.dirichlet_mask.to_array(), A["mask"])
    np.testing.assert_array_equal(ps.vent_mask.to_array(), A["vent"])
    assert rel(ps.b.to_array(), A["b"]) <= 1e-6
    p_gpu = ps.pressure.to_array()
    assert rel(p_gpu, x) <= 1e-5
    # the projected dam-break velocity is ~0 (hydrostatic balance), so the
    # correction is checked against the oracle's correction of the GPU's own
    # pressure, relative to the size of the input velocity
    own = O.pressure3_correct(sc["shape"], sc["spacing"], p_gpu, sc["vel"], beta=sc["beta"], suffix="f32")
    vin = np.concatenate(sc["vel"])
    for a in range(3):
        got = vel[a].to_array()
        assert np.linalg.norm(got - own[a]) <= 1e-6 * np.linalg.norm(vin)
        assert np.linalg.norm(got - out[a]) <= 1e-3 * np.linalg.norm(vin)


def test_random_scene_with_collider_matches_oracle(flux):
    F = flux
    n = 32
    sc = scenes.random_scene(n, 3)
    rng = np.random.default_rng(8)
    # collider open fractions: a solid slab at x < 4 (weight 0) and partial faces around it
    collider = []
    for a in range(3):
        fs = list((n, n, n))
        fs[a] += 1
        wgt = np.ones(fs[::-1], np.float32)
        if a == 0:
            wgt[:, :, :5] = 0.0
            wgt[:, :, 5] = rng.uniform(0.01, 1.0, wgt[:, :, 5].shape)
        else:
            wgt[:, :, :4] = 0.0
        collider.append(wgt.ravel())
    ps, vel, diag = gpu_project(F, sc, 4, 1e-6, collider=collider)
    A, x, x64, st, out = oracle_project(sc, 4, 1e-6, 48, collider=collider)
    print("gpu", diag, "oracle", st)
    assert diag["converged"] == st["converged"]
    np.testing.assert_array_equal(ps.dirichlet_mask.to_array(), A["mask"])
    assert rel(ps.b.to_array(), A["b"]) <= 1e-6
    assert rel(ps.pressure.to_array(), x) <= 1e-5
    for a in range(3):
        assert rel(vel[a].to_array(), out[a]) <= 1e-5


def test_circuit_breaker_restores_snapshot(flux):
    """maxPlausiblePressure tiny -> rejected, pressure restored to the last
    accepted snapshot (zeros), velocity corrected with it (grid_pressure_solver3.js:545-554)."""
    F = flux
    sc = scenes.dam_break_system(16)
    ps, vel, diag = gpu_project(F, sc, 3, 1e-6, max_plausible_pressure=1e-12)
    assert diag["rejected"] and diag["badCells"] > 0
    p = ps.pressure.to_array()
    assert np.all(p == 0)
    for a in range(3):
        np.testing.assert_array_equal(vel[a].to_array(), sc["vel"][a])


def test_warm_start_from_converged_pressure_is_immediate(flux):
    """The pressure field is the warm start (grid_pressure_solver3.js:519-524):
    solving the same assembled system again from the converged pressure needs
    no iteration."""
    F = flux
    sc = scenes.dam_break_system(32)
    n, h = sc["shape"], sc["spacing"]
    vel = [F.array_n("float", F.face_shape(n, a), sc["vel"][a]) for a in range(3)]
    out = [F.array_n("float", F.face_shape(n, a)) for a in range(3)]
    beta = [F.array_n("float", F.face_shape(n, a), sc["beta"][a]) for a in range(3)]
    act = F.array_n("u8", n, sc["active"])
    tgt = F.array_n("float", n, sc["target"])
    ps = F.create_grid_pressure_solver3(resolution=n, grid_spacing=h, dirichlet=(act, tgt), face_weights=beta,
                                        multigrid={"levels": 4}, tolerance=1e-5)
    ps.assemble(vel)
    d1 = ps.solve(vel, out)
    assert d1["converged"] and d1["iterations"] > 3
    d2 = ps.solve(vel, out)
    assert d2["converged"] and d2["iterations"] <= 1


def test_large_projection_reaches_tolerance(flux):
    """128^3 two-phase (1000:1): the FP64 true residual of the solve meets
    1e-6 relative, and the FP32 pressure the oracle recomputes sits at the FP32
    representability floor, not above it."""
    F = flux
    n = 128
    sc = scenes.dam_break_system(n)
    ps, vel, diag = gpu_project(F, sc, 6, 1e-6)
    print(diag)
    assert diag["converged"] and diag["residual"] <= 1e-6 * diag["bNorm"]
    p = ps.pressure.to_array()
    A = O.pressure3_assemble(sc["shape"], sc["spacing"], sc["vel"], active=sc["active"], target=sc["target"],
                             beta=sc["beta"], suffix="f32")
    r32 = O.true_residual(sc["shape"], sc["spacing"], A["b"], p, A["mask"], A["weights"], suffix="f32")
    print("f32-rounded pressure relative residual:", r32 / diag["bNorm"])
    assert r32 / diag["bNorm"] < 1e-3


@pytest.mark.parametrize("storage,vtol", [("f16", 5e-3), ("bf16", 5e-2)])
def test_mixed_precision_vcycle_tracks_fp32(flux, frame110, storage, vtol):
    """FP16/BF16 storage with FP32 arithmetic: one V-cycle stays within the
    storage precision of the FP32 oracle's V-cycle (a wrong residual scale or
    conversion would be off by O(1))."""
    F, c = flux, frame110
    sh = c["shape"]
    m = F.array_n("u8", sh, c["mask"])
    v = F.array_n("u8", sh, c["vent"])
    w = [F.array_n("float", s, c["weights"][a]) for a, s in enumerate(
        [(sh[0] + 1, sh[1], sh[2]), (sh[0], sh[1] + 1, sh[2]), (sh[0], sh[1], sh[2] + 1)])]
    mg = F.create_multigrid_preconditioner(sh, c["spacing"], dirichlet_mask=m, coarse_dirichlet_mask=v,
                                           face_weights=w, levels=4, storage=storage)
    z = F.array_n("float", sh)
    mg(F.array_n("float", sh, c["b"]), z)
    ref = O.Multigrid(sh, c["spacing"], levels=4, mask=c["mask"], coarse_mask=c["vent"], weights=c["weights"],
                      suffix="f32").apply(c["b"])
    err = rel(z.to_array(), ref)
    print(storage, "V-cycle rel L2 vs FP32 oracle:", err)
    assert err <= vtol


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_mixed_precision_projection_reaches_same_tolerance(flux, storage):
    F = flux
    n = 64
    sc = scenes.dam_break_system(n)
    ps, vel, diag = gpu_project(F, sc, 5, 1e-6)
    n_, h = sc["shape"], sc["spacing"]
    vin = [F.array_n("float", F.face_shape(n_, a), sc["vel"][a]) for a in range(3)]
    beta = [F.array_n("float", F.face_shape(n_, a), sc["beta"][a]) for a in range(3)]
    act = F.array_n("u8", n_, sc["active"])
    tgt = F.array_n("float", n_, sc["target"])
    pm = F.create_grid_pressure_solver3(resolution=n_, grid_spacing=h, dirichlet=(act, tgt), face_weights=beta,
                                        multigrid={"levels": 5, "storage": storage}, tolerance=1e-6)
    dm = pm.project(vin)
    print("plain", diag, storage, dm)
    assert dm["converged"] and dm["residual"] <= 1e-6 * dm["bNorm"]
    if storage == "f16":
        # FP16 with the h^2-scaled coarse levels preconditions as well as FP32
        assert dm["iterations"] <= diag["iterations"] + 2
    else:
        # BF16's 8-bit mantissa costs iterations
        assert dm["iterations"] <= int(1.5 * diag["iterations"])
    assert rel(pm.pressure.to_array(), ps.pressure.to_array()) <= 1e-4


@pytest.mark.parametrize("shape", [(12, 12, 8), (12, 16, 8), (16, 12, 8), (10, 6, 6), (16, 16, 8)])
def test_projection_on_planes_that_do_not_fill_wavefronts(flux, shape):
    """The PCG's r.z is fused into the V-cycle's last relaxation as a block
    reduction; a plane of (n0/2) x n1 cell pairs that is not a multiple of
    the 64-lane wavefront leaves part of a wavefront idle, and every thread
    must still pass the reduction once (a free-surface system, 2 levels)."""
    F = flux
    n = int(np.prod(shape))
    j = (np.arange(n) // shape[0]) % shape[1]
    active = (j >= shape[1] // 2).astype(np.uint8)
    target = (active * (0.05 * 9.81 * (j + 0.5))).astype(np.float32)
    rng = np.random.default_rng(2)
    vel = []
    for a in range(3):
        fs = list(shape)
        fs[a] += 1
        v = rng.standard_normal(fs[::-1]).astype(np.float32)
        lo, hi = [slice(None)] * 3, [slice(None)] * 3
        lo[2 - a], hi[2 - a] = 0, -1
        v[tuple(lo)] = 0
        v[tuple(hi)] = 0
        vel.append(v.ravel())
    ps = F.create_grid_pressure_solver3(resolution=shape, grid_spacing=(1.0, 1.0, 1.0),
                                        dirichlet=(F.array_n("u8", shape, active), F.array_n("float", shape, target)),
                                        multigrid={"levels": 2}, tolerance=1e-6)
    d = ps.project([F.array_n("float", F.face_shape(shape, a), vel[a]) for a in range(3)])
    A = O.pressure3_assemble(shape, (1.0, 1.0, 1.0), vel, active=active, target=target, suffix="f32")
    mg = O.Multigrid(shape, (1.0, 1.0, 1.0), levels=2, mask=A["mask"], coarse_mask=A["vent"], suffix="f32")
    x, _, od = O.pcg_solve_mixed(shape, (1.0, 1.0, 1.0), A["b"], x0=A["pressure"], mask=A["mask"], mg=mg, tol=1e-6,
                                 maxiter=100)
    assert d["converged"] and od["converged"], (d, od)
    assert abs(d["iterations"] - od["iterations"]) <= 1
    p = ps.pressure.to_array()
    assert np.linalg.norm(p - x) <= 1e-5 * np.linalg.norm(x)
