"""The 2-D Eulerian step on the GPU against the oracle: polygon SDF,
order-1 face advection stopped at the collider, constrainVelocity with a
collider and an inflow, the outflow velocity condition, the max-|u|
reduction, the 2-D projection, and whole frames of the Karman scene
(adaptive CFL sub-steps) at reduced and full size."""
import numpy as np
import pytest

from oracle import pyoracle as O
from tsl_array_n_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G(flux):
    from tsl_array_n_amd import grid2
    return grid2


def smooth_velocity(n, seed, amp=1.5):
    rng = np.random.default_rng(seed)
    su, sv = (n[0] + 1, n[1]), (n[0], n[1] + 1)
    ax, ay = rng.uniform(0.1, 0.4, 2)
    ph = rng.uniform(0, 6.28, 4)
    X, Y = np.meshgrid(np.arange(su[0]), np.arange(su[1]) + 0.5)
    u = amp * np.sin(ax * X + ph[0]) * np.cos(ay * Y + ph[1]) + 0.6
    X, Y = np.meshgrid(np.arange(sv[0]) + 0.5, np.arange(sv[1]))
    v = amp * np.cos(ax * X + ph[2]) * np.sin(ay * Y + ph[3])
    return u.astype(np.float32).ravel(), v.astype(np.float32).ravel()


def sdf_field(F, G, n, rings, invert=False):
    c = G.SDFStaticCollider2(n[0], n[1])
    c.add_polygons(rings, invert)
    return c


def test_polygon_sdf_matches_oracle(flux, G):
    n = (64, 40)
    rings = [scenes.circle_polygon(20.3, 17.9, 7.5, 48), [(40, 5), (55, 8), (50, 30), (42, 25)]]
    c = sdf_field(flux, G, n, rings)
    got = c.sdf.to_array()
    want = O.polygon_sdf2(rings, n, (1.0, 1.0), (0.0, 0.0))
    np.testing.assert_array_max_ulp(got, want, maxulp=1)
    c2 = sdf_field(flux, G, n, rings, invert=True)
    np.testing.assert_array_max_ulp(c2.sdf.to_array(), -want, maxulp=1)


def test_advect_faces2_matches_oracle(flux, G):
    F = flux
    n = (48, 32)
    u, v = smooth_velocity(n, 1)
    col = sdf_field(F, G, n, [scenes.circle_polygon(20.0, 16.0, 5.0, 48)])
    sdf = col.sdf.to_array()
    g = G.grid2(n)
    for collider, dt in ((None, 0.9), (col, 0.9), (col, 3.7)):
        fin = (F.array_n("float", G.face_shape2(n, 0), u), F.array_n("float", G.face_shape2(n, 1), v))
        fout = (F.array_n("float", G.face_shape2(n, 0), u), F.array_n("float", G.face_shape2(n, 1), v))
        G.advect_face_centered2(g, fin, fout, dt, collider)
        uo, vo = O.advect_face2(u, v, n, (1.0, 1.0), (0.0, 0.0), dt, sdf if collider is not None else None)
        np.testing.assert_array_equal(fout[0].to_array(), uo)
        np.testing.assert_array_equal(fout[1].to_array(), vo)


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("with_collider", [False, True])
def test_advection_solver2_matches_oracle(flux, G, order, with_collider):
    """createSemiLagrangianAdvectionSolver2 with a separate advected grid,
    order 1 and MacCormack, with and without the collider; faces inside the
    collider keep the output's previous value"""
    F = flux
    n = (48, 32)
    u, v = smooth_velocity(n, 7)
    rng = np.random.default_rng(8)
    du = (np.sin(0.4 * np.arange(u.size)) + 2.0 * (np.arange(u.size) % 49 > 24) + rng.uniform(-0.05, 0.05, u.size))
    dv = (np.cos(0.3 * np.arange(v.size)) + rng.uniform(-0.05, 0.05, v.size))
    du, dv = du.astype(np.float32), dv.astype(np.float32)
    col = sdf_field(F, G, n, [scenes.circle_polygon(20.0, 16.0, 5.0, 48)]) if with_collider else None
    vel = (F.array_n("float", G.face_shape2(n, 0), u), F.array_n("float", G.face_shape2(n, 1), v))
    din = (F.array_n("float", G.face_shape2(n, 0), du), F.array_n("float", G.face_shape2(n, 1), dv))
    prev = [np.full(u.size, 7.25, np.float32), np.full(v.size, -3.5, np.float32)]
    out = (F.array_n("float", G.face_shape2(n, 0), prev[0]), F.array_n("float", G.face_shape2(n, 1), prev[1]))
    adv = G.SemiLagrangianAdvectionSolver2(G.grid2(n), vel, 0.9, collider=col, order=order)
    for dt in (0.9, 2.6):
        adv.set_dt(dt)
        adv.advect_face_centered2(din, out)
        want = O.advect_faces2_vd([u, v], [du, dv], n, (1.0, 1.0), (0.0, 0.0), dt,
                                  sdf=col.sdf.to_array() if col is not None else None, order=order, out=prev)
        np.testing.assert_array_equal(out[0].to_array(), want[0])
        np.testing.assert_array_equal(out[1].to_array(), want[1])
        prev = want


def test_constrain2_matches_oracle(flux, G):
    F = flux
    n = (48, 32)
    u, v = smooth_velocity(n, 2, amp=3.0)
    u[17] = 400.0
    v[33] = np.inf
    col = sdf_field(F, G, n, [scenes.circle_polygon(22.0, 15.0, 6.0, 48)])
    inflow = G.SDFInflow2(n[0], n[1], velocity=(1.25, -0.5), mode="set")
    inflow.add_polygons([scenes.wall_strip_polygon(2.0, -1000.0, n[1])])
    add = G.SDFInflow2(n[0], n[1], velocity=(0.25, 0.125), mode="add")
    add.add_polygons([[(30, 20), (40, 20), (40, 30), (30, 30)]])
    vel = (F.array_n("float", G.face_shape2(n, 0), u), F.array_n("float", G.face_shape2(n, 1), v))
    b = G.GridBlockedBoundaryConditionSolver2(vel, n[0], n[1], collider=col, inflows=[inflow, add])
    b.closed_domain_boundary_flag = 4 | 8
    b.constrain_velocity(5)
    want = O.constrain_velocity2([u, v], n, (1.0, 1.0), (0.0, 0.0), 4 | 8, col.sdf.to_array(), 0.5, 5,
                                 inflows=[(inflow.sdf.to_array(), (1.25, -0.5), "set"),
                                          (add.sdf.to_array(), (0.25, 0.125), "add")])
    np.testing.assert_array_equal(vel[0].to_array(), want[0])
    np.testing.assert_array_equal(vel[1].to_array(), want[1])


def test_outflow2_matches_oracle(flux, G):
    F = flux
    n = (40, 24)
    u, v = smooth_velocity(n, 3)
    out = G.SDFOutflow2(n[0], n[1])
    out.add_polygons([scenes.wall_strip_polygon(n[0] - 3.0, n[0] + 1000.0, n[1])])
    g = G.grid2(n)
    o = G.GridOutflowSolver2(g, out)
    vel = (F.array_n("float", G.face_shape2(n, 0), u), F.array_n("float", G.face_shape2(n, 1), v))
    o.apply_outflow_velocity_bc(vel, 0.6)
    want = O.outflow_bc2([u, v], n, (1.0, 1.0), (0.0, 0.0), 0.6, out.sdf.to_array())
    np.testing.assert_array_equal(vel[0].to_array(), want[0])
    np.testing.assert_array_equal(vel[1].to_array(), want[1])
    active, target = G.outflow_dirichlet2(g, [out])
    np.testing.assert_array_equal(active.to_array(), O.outflow_dirichlet2(out.sdf.to_array(), n, (1.0, 1.0), (0.0, 0.0)))
    assert not target.to_array().any()


def test_max_abs_and_substeps(flux, G):
    F = flux
    rng = np.random.default_rng(6)
    a = rng.uniform(-2, 2, 100003).astype(np.float32)
    b = rng.uniform(-1, 1, 777).astype(np.float32)
    b[400] = -7.5
    fa, fb = F.array_n("float", (a.size,), a), F.array_n("float", (b.size,), b)
    assert G.max_abs([fa, fb]) == O.max_abs([a, b]) == 7.5
    for args in ((0.1, 1, 1 / 30, 5, 32), (100, 1, 1 / 30, 1, 32), (1e6, 1, 1 / 30, 1, 32), (2.3, 1, 1.0, 0.5, 32)):
        assert G.compute_adaptive_substeps(*args) == O.adaptive_substeps(*args)


def test_pressure2_matches_oracle(flux, G):
    F = flux
    n, h = (64, 32), (1.0, 1.0)
    u, v = smooth_velocity(n, 4)
    out = G.SDFOutflow2(n[0], n[1])
    out.add_polygons([scenes.wall_strip_polygon(n[0] - 2.0, n[0] + 1000.0, n[1])])
    active, target = G.outflow_dirichlet2(G.grid2(n), [out])
    ps = G.GridPressureSolver2(n, h, dirichlet=(active, target), multigrid={"levels": 4}, tolerance=1e-6)
    vel = (F.array_n("float", G.face_shape2(n, 0), u), F.array_n("float", G.face_shape2(n, 1), v))
    d = ps.project(vel)
    assert d["converged"] and not d["rejected"]
    act = active.to_array()
    A = O.pressure2_assemble(n, h, [u, v], active=act, target=target.to_array())
    np.testing.assert_array_equal(ps.b.to_array(), A["b"])
    mg = O.Multigrid(n, h, levels=4, mask=act, suffix="f32")
    x, _, od = O.pcg_solve_mixed(n, h, A["b"], x0=A["pressure"], mask=act, mg=mg, tol=1e-6, maxiter=768)
    assert od["converged"]
    p = ps.pressure.to_array()
    assert np.linalg.norm(p - x) <= 1e-5 * np.linalg.norm(x)
    want = O.pressure2_correct(n, h, p, [u, v])
    np.testing.assert_array_equal(vel[0].to_array(), want[0])
    np.testing.assert_array_equal(vel[1].to_array(), want[1])


def run_karman(F, G, scene, frames, order=1):
    solver, stepper, keep = G.create_karman2d(scene, order=order)
    st = {"u": scene["u0"].copy(), "v": scene["v0"].copy(),
          "pressure": np.zeros(int(np.prod(scene["resolution"])), np.float32)}
    st["p_accepted"] = st["pressure"].copy()
    col, inflow, outflow = [k.sdf.to_array() for k in keep]
    n = scene["resolution"]
    for f in range(frames):
        diags = G.advance_frame(solver, stepper)
        k, dt, odiags = O.solver2_frame(st, n, scene["spacing"], scene["origin"], scene["frame_dt"],
                                        courant=scene["courant"], sdf=col,
                                        inflows=[(inflow, scene["inflow_velocity"], "set")], outflow_sdf=outflow,
                                        flags=scene["closed_flags"], levels=scene["multigrid"]["levels"],
                                        tol=scene["tolerance"], order=order)
        assert k == len(diags) and np.float32(dt) == np.float32(stepper.state["lastSubDt"])
        for d, od in zip(diags, odiags):
            assert d["converged"] and od["converged"] and not d["rejected"]
        u, v = [f_.to_array() for f_ in solver.velocity_grid]
        vel, ovel = np.concatenate([u, v]), np.concatenate([st["u"], st["v"]])
        err = np.linalg.norm(vel - ovel) / np.linalg.norm(ovel)
        print("frame", f, "substeps", k, "iterations", [d["iterations"] for d in diags],
              [d["iterations"] for d in odiags], "rel err %.2e" % err)
        assert err <= 2e-5, (f, err)
    return solver, st


def test_karman_reduced_frames_match_oracle(flux, G):
    scene = scenes.karman2d(nx=64, ny=32, radius=4.0, center=(16.0, 16.0), levels=3)
    run_karman(flux, G, scene, 4)


def test_karman_maccormack_frames_match_oracle(flux, G):
    scene = scenes.karman2d(nx=64, ny=32, radius=4.0, center=(16.0, 16.0), levels=3)
    run_karman(flux, G, scene, 3, order=2)


def test_karman_full_size_frames_match_oracle(flux, G):
    scene = scenes.karman2d()
    run_karman(flux, G, scene, 2)


def test_karman_long_run_stays_healthy(flux, G):
    """60 frames of the full 256x128 scene on the GPU alone: every solve
    converges and is accepted, the velocity stays finite and bounded, and
    the inflow holds u = 1 on its faces."""
    scene = scenes.karman2d()
    solver, stepper, keep = G.create_karman2d(scene)
    steps = 0
    for f in range(60):
        for d in G.advance_frame(solver, stepper):
            steps += 1
            assert d["converged"] and not d["rejected"], (f, d)
    u, v = [x.to_array() for x in solver.velocity_grid]
    assert np.isfinite(u).all() and np.isfinite(v).all()
    assert np.abs(u).max() < 10 and np.abs(v).max() < 10
    U = u.reshape(128, 257)
    assert np.all(U[:, :2] == 1.0)
    print("60 frames,", steps, "steps, max |u| %.3f" % stepper.state["maxVelocity"])
