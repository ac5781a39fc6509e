"""The Node.js binding (N-API addon + ESM module): the reference's own host
language driving the HIP library."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from oracle import pyoracle as O
from tsl_array_n_amd import scenes

NODE = shutil.which("node")
ADDON = os.path.join(ROOT, "tsl_array_n_amd", "lib", "fluxflow_napi.node")

pytestmark = pytest.mark.skipif(NODE is None, reason="node is not installed")


def run_node(script, *args, timeout=120):
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "node", script)] + [str(a) for a in args],
                       capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout + r.stderr
    return r


def test_api_surface_and_host_helpers(tmp_path):
    assert os.path.exists(ADDON), "build the addon with __graft_entry__.build()"
    out = tmp_path / "api.json"
    run_node("check_api.mjs", out, os.path.join(GOLDEN, "stalled_system", "frame-110"))
    d = json.load(open(out))
    for name in ["init", "arrayN", "array0", "createLaplacianOperator", "createMultigridPreconditioner",
                 "createJacobiPreconditioner", "createIdentityPreconditioner",
                 "createPreconditionedConjugateGradientSolver", "createGridPressureSolver3",
                 "createFaceCenteredGrid3", "computeFlipBoxSeed3", "createJitterRandom", "binParticles",
                 "particleToGrid", "gridToParticle", "advectParticles", "advectFaceCentered3",
                 "createExtrapolateToRegion3", "createSDFStaticCollider3", "sphere", "box",
                 "createGridBlockedBoundaryConditionSolver3", "createGridFlipSolver3", "createGridSmokeSolver3",
                 "createFaceCenteredGrid2", "createSDFStaticCollider2", "createSDFInflow2", "createSDFOutflow2",
                 "createGridSolver2", "createGridAdaptiveTimeStep2", "computeAdaptiveSubSteps",
                 "createMaxAbsReducer", "exportSystem", "loadSystem", "createSemiLagrangianAdvectionSolver3",
                 "createCellCenteredScalarGrid3"]:
        assert name in d["exports"], name
    assert d["strides"] == [1, 4, 12]  # TA/test/array.test.js:41-93
    assert "1D/2D/3D" in d["shapeError"]
    assert "flux_init" in d["initError"]
    fx = d["fixture"]
    assert fx["resolution"] == [48, 24, 24] and fx["bLength"] == 48 * 24 * 24 and fx["weightU"] == 49 * 24 * 24
    assert abs(fx["bNorm"] - 3.888) < 5e-3 and fx["iterations"] == 15
    assert d["substeps"][0] == 4  # cfl.test.js:31-39
    golden = json.load(open(os.path.join(GOLDEN, "mulberry32.json")))
    np.testing.assert_array_equal(d["random"], golden["vectors"][0]["values"][:16])
    ref = O.flip_box_seed3([0, 0, 0], [0.4, 0.6, 1.0], [1 / 16] * 3)
    assert d["seedCount"] == ref.size // 3
    np.testing.assert_array_equal(np.array(d["seedHead"], np.float32), ref[:24])
    assert abs(d["seedSum"] - ref.astype(np.float64).sum()) < 1e-6 * abs(d["seedSum"])


@pytest.mark.gpu
def test_projection_through_node_matches_oracle(tmp_path):
    from tsl_array_n_amd import scenes
    n = 16
    r = run_node("projection.mjs", n, tmp_path, timeout=300)
    print(r.stdout)
    meta = json.load(open(tmp_path / "diag.json"))
    assert meta["outside"] == 0 and meta["diag"]["converged"], meta
    sc = scenes.dam_break_system(n)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "fraction.bin", np.float32), sc["fraction"])
    A = O.pressure3_assemble(sc["shape"], sc["spacing"], sc["vel"], active=sc["active"], target=sc["target"],
                             beta=sc["beta"], suffix="f32")
    mg = O.Multigrid(sc["shape"], sc["spacing"], levels=3, mask=A["mask"], coarse_mask=A["vent"],
                     weights=A["weights"], suffix="f32")
    x, x64, st = O.pcg_solve_mixed(sc["shape"], sc["spacing"], A["b"], x0=A["pressure"], mask=A["mask"],
                                   weights=A["weights"], mg=mg, tol=1e-6, maxiter=100)
    p = np.fromfile(tmp_path / "pressure.bin", np.float32)
    assert np.linalg.norm(p - x) <= 1e-5 * np.linalg.norm(x)
    # the projected dam-break velocity is ~0 (hydrostatic balance): the
    # correction is checked against the oracle's correction of the same
    # pressure, relative to the size of the input velocity
    own = O.pressure3_correct(sc["shape"], sc["spacing"], p, sc["vel"], beta=sc["beta"], suffix="f32")
    out = O.pressure3_correct(sc["shape"], sc["spacing"], x, sc["vel"], beta=sc["beta"], suffix="f32")
    vin = np.linalg.norm(np.concatenate(sc["vel"]))
    for a, name in enumerate("uvw"):
        got = np.fromfile(tmp_path / (name + ".bin"), np.float32)
        assert np.linalg.norm(got - own[a]) <= 1e-6 * vin
        assert np.linalg.norm(got - out[a]) <= 1e-3 * vin


@pytest.mark.gpu
def test_headline_projection_through_node_matches_oracle(tmp_path):
    """The benchmark's headline form through the Node binding (multigrid
    options storage 'f16', smoother 'brick', 12 sweeps; setCoarseSpace(true))
    on the 32^3 dam break: accepted at the FP64 1e-6 true residual, pressure
    within 1e-5 of the oracle's plain solve (both within ~1e-6 of the exact
    solution), in fewer iterations."""
    from tsl_array_n_amd import scenes
    n = 32
    r = run_node("projection.mjs", n, tmp_path, "headline", timeout=300)
    print(r.stdout)
    meta = json.load(open(tmp_path / "diag.json"))
    assert meta["outside"] == 0 and meta["diag"]["converged"], meta
    sc = scenes.dam_break_system(n)
    A = O.pressure3_assemble(sc["shape"], sc["spacing"], sc["vel"], active=sc["active"], target=sc["target"],
                             beta=sc["beta"], suffix="f32")
    mg = O.Multigrid(sc["shape"], sc["spacing"], levels=3, mask=A["mask"], coarse_mask=A["vent"],
                     weights=A["weights"], suffix="f32")
    x, x64, st = O.pcg_solve_mixed(sc["shape"], sc["spacing"], A["b"], x0=A["pressure"], mask=A["mask"],
                                   weights=A["weights"], mg=mg, tol=1e-6, maxiter=100)
    p = np.fromfile(tmp_path / "pressure.bin", np.float32)
    assert np.linalg.norm(p - x) <= 1e-5 * np.linalg.norm(x)
    assert meta["diag"]["iterations"] < st["iterations"]


@pytest.mark.gpu
def test_flip_headline_form_and_set_dt_through_node(flux, tmp_path):
    """createGridFlipSolver3 through Node with the headline projection form
    (multigrid storage 'f16', smoother 'brick', setCoarseSpace) and setDt
    between steps gives the positions and velocities of the same run driven
    from Python bit for bit, every projection accepted."""
    from tsl_array_n_amd import flip as FL
    r = run_node("flip_forms.mjs", tmp_path, timeout=300)
    print(r.stdout)
    meta = json.load(open(tmp_path / "diag.json"))
    assert all(d["converged"] and not d["rejected"] for d in meta["diags"]), meta
    n, h = 32, 1.0 / 32
    pos = []
    for k in range(n):
        for j in range(19):
            for i in range(13):
                for sk in range(2):
                    for sj in range(2):
                        for si in range(2):
                            pos.append(((i + (si + 0.5) / 2) * h, (j + (sj + 0.5) / 2) * h, (k + (sk + 0.5) / 2) * h))
    pos = np.asarray(pos, np.float32)
    assert len(pos) == meta["count"]
    s = FL.GridFlipSolver3(resolution=(n, n, n), grid_spacing=(h, h, h), max_particles=len(pos), dt=1.0 / 60.0,
                           pressure={"multigrid": {"levels": 3, "storage": "f16", "smoother": "brick", "nu_down": 12,
                                                   "nu_up": 12, "n_final": 12}, "tolerance": 1e-6})
    s.set_coarse_space(True)
    s.positions.from_array(pos)
    its = []
    for _ in range(2):
        its.append(s.on_advance_time_step()["iterations"])
    s.set_dt(1.0 / 120.0)
    for _ in range(2):
        its.append(s.on_advance_time_step()["iterations"])
    assert its == [d["iterations"] for d in meta["diags"]]
    np.testing.assert_array_equal(np.fromfile(tmp_path / "positions.bin", np.float32), s.positions.to_array().ravel())
    np.testing.assert_array_equal(np.fromfile(tmp_path / "velocities.bin", np.float32),
                                  s.velocities.to_array().ravel())


@pytest.mark.gpu
def test_flip_and_collider_through_node_match_oracle(tmp_path):
    """createGridFlipSolver3, createSDFStaticCollider3.computeFaceWeights and
    constrainVelocity driven from Node, against the oracle."""
    run_node("flip_step.mjs", tmp_path, timeout=300)
    meta = json.load(open(tmp_path / "meta.json"))
    n, h, dt = (16, 16, 16), (1 / 16,) * 3, 1 / 60
    pos = np.fromfile(tmp_path / "positions.bin", np.float32)
    st = {"pos": None, "vel": None, "pressure": np.zeros(16 ** 3, np.float32), "p_accepted": np.zeros(16 ** 3, np.float32)}
    p0 = []
    for k in range(16):
        for j in range(10):
            for i in range(6):
                for sk in range(2):
                    for sj in range(2):
                        for si in range(2):
                            p0.append([(i + (si + 0.5) / 2) * h[0], (j + (sj + 0.5) / 2) * h[1], (k + (sk + 0.5) / 2) * h[2]])
    st["pos"] = np.array(p0, np.float32).ravel()
    st["vel"] = np.zeros_like(st["pos"])
    for a, c in enumerate("uvw"):
        st[c] = np.zeros(int(np.prod(O.face_shape(n, a))), np.float32)
    for s in range(2):
        d = O.flip3_step(st, n, h, (0, 0, 0), dt, levels=3)
        assert meta["diags"][s]["converged"] and d["converged"]
    np.testing.assert_array_equal(np.fromfile(tmp_path / "fluid.bin", np.float32), st["mask"])
    assert np.max(np.abs(pos - st["pos"])) <= 1e-5
    vel = np.fromfile(tmp_path / "velocities.bin", np.float32)
    assert np.linalg.norm(vel - st["vel"]) <= 1e-4 * np.linalg.norm(st["vel"])
    # collider
    sh = [{"type": "sphere", "center": (5.3, 6.1, 4.7), "radius": 2.6}]
    m = (12, 12, 12)
    sdf = O.collider_sdf3(sh, m, (1, 1, 1), (0, 0, 0))
    np.testing.assert_allclose(np.fromfile(tmp_path / "sdf.bin", np.float32), sdf, atol=1e-6)
    wts = O.face_weights3(sh, m, (1, 1, 1), (0, 0, 0))
    for a, c in enumerate("uvw"):
        assert np.max(np.abs(np.fromfile(tmp_path / ("weight_" + c + ".bin"), np.float32) - wts[a])) <= 1e-6
    vin = [np.fromfile(tmp_path / ("vin_%s.bin" % "x".join(str(x) for x in O.face_shape(m, a))), np.float32)
           for a in range(3)]
    want = O.constrain_velocity3(vin, m, (1, 1, 1), (0, 0, 0), closed_flags=1 | 4 | 32, sdf=sdf, weights=wts,
                                 friction=0.5, depth=5)
    for a, c in enumerate("uvw"):
        np.testing.assert_allclose(np.fromfile(tmp_path / ("vout_" + c + ".bin"), np.float32), want[a], rtol=1e-6,
                                   atol=1e-6)


@pytest.mark.gpu
def test_smoke_through_node_matches_oracle(tmp_path):
    run_node("smoke_step.mjs", tmp_path, timeout=300)
    meta = json.load(open(tmp_path / "meta.json"))
    n, h = (24, 24, 24), (1.0, 1.0, 1.0)
    cells = 24 ** 3
    st = {"dens": np.zeros(cells, np.float32), "temp": np.zeros(cells, np.float32),
          "pressure": np.zeros(cells, np.float32), "p_accepted": np.zeros(cells, np.float32)}
    for a, c in enumerate("uvw"):
        st[c] = np.zeros(int(np.prod(O.face_shape(n, a))), np.float32)
    for s in range(3):
        st["dens"] = O.source_sphere3(st["dens"], n, h, (0, 0, 0), (12, 2.5, 12), 3.0, 1.0)
        st["temp"] = O.source_sphere3(st["temp"], n, h, (0, 0, 0), (12, 2.5, 12), 3.0, 1.0)
        d = O.smoke3_step(st, n, h, (0, 0, 0), 0.25, a=-0.05, b=0.1, levels=3)
        assert d["converged"] and meta["diags"][s]["converged"]
    vel = np.concatenate([st[c] for c in "uvw"])
    got = np.concatenate([np.fromfile(tmp_path / (c + ".bin"), np.float32) for c in "uvw"])
    assert np.linalg.norm(got - vel) <= 1e-4 * np.linalg.norm(vel)
    for name, key in (("density", "dens"), ("temperature", "temp")):
        g = np.fromfile(tmp_path / (name + ".bin"), np.float32)
        assert np.linalg.norm(g - st[key]) <= 1e-4 * np.linalg.norm(st[key])


@pytest.mark.gpu
def test_flip_concentration_through_node_matches_oracle(tmp_path):
    """createGridFlipSolver3 with carried concentration, density coupling and
    the top-centre pin from Node: two steps of a stably stratified closed
    box, against the oracle's composed step"""
    n = (8, 8, 8)
    g = np.stack(np.meshgrid(*[(np.arange(2 * n[a]) + 0.5) / 2 for a in range(3)], indexing="ij"), -1).reshape(-1, 3)
    pos = (g + np.random.default_rng(3).uniform(-0.1, 0.1, g.shape)).astype(np.float32).ravel()
    conc = (pos.reshape(-1, 3)[:, 1] < 4).astype(np.float32)
    pos.tofile(tmp_path / "pos.bin")
    conc.tofile(tmp_path / "conc.bin")
    run_node("flip_conc_step.mjs", tmp_path, *n, timeout=300)
    meta = json.load(open(tmp_path / "meta.json"))
    cells = int(np.prod(n))
    st = {"pos": pos.copy(), "vel": np.zeros_like(pos), "conc": conc.copy(), "pressure": np.zeros(cells, np.float32),
          "p_accepted": np.zeros(cells, np.float32)}
    for a, c in enumerate("uvw"):
        st[c] = np.zeros(int(np.prod(O.face_shape(n, a))), np.float32)
    for s in range(2):
        d = O.flip3_step(st, n, (1.0, 1.0, 1.0), (0, 0, 0), 0.05, levels=2, carry_concentration=True, mixing=0.3,
                         fade=0.01, density=(1.0, 3.0), pressure_pin="topCenter", resample=False)
        assert d["converged"] and meta["diags"][s]["converged"]
    rd = lambda name: np.fromfile(tmp_path / (name + ".bin"), np.float32)
    assert np.max(np.abs(rd("out_pos") - st["pos"])) <= 1e-5
    assert np.max(np.abs(rd("out_vel") - st["vel"])) <= 1e-4 * max(np.abs(st["vel"]).max(), 1e-3)
    assert np.max(np.abs(rd("out_conc") - st["conc"])) <= 1e-5
    np.testing.assert_array_equal(rd("out_cell_density"), st["cell_density"])


@pytest.mark.gpu
def test_maccormack_advection_through_node_matches_oracle(tmp_path):
    """createSemiLagrangianAdvectionSolver3 (order 2) from Node, dt a live
    array0 field changed between the two dispatches: bit for bit the oracle."""
    run_node("advect_step.mjs", tmp_path, timeout=300)
    n, h, o = (20, 16, 12), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)
    rd = lambda name: np.fromfile(tmp_path / (name + ".bin"), np.float32)
    V = [rd("vel" + c) for c in "UVW"]
    D = [rd("data" + c) for c in "UVW"]
    want = O.advect_faces3_vd(V, D, n, h, o, 1.3, order=2)
    for a, c in enumerate("UVW"):
        np.testing.assert_array_equal(rd("out" + c), want[a])
    np.testing.assert_array_equal(rd("outScalar"), O.advect_scalar3(rd("scalar"), *V, n, h, o, 2.2, order=2))


@pytest.mark.gpu
def test_karman2d_through_node_matches_oracle(tmp_path):
    """createGridSolver2 + createGridAdaptiveTimeStep2 driven from Node over
    two frames of a reduced Karman scene, against the composed oracle fed
    the same SDFs and initial velocity."""
    run_node("karman_step.mjs", tmp_path, timeout=300)
    meta = json.load(open(tmp_path / "meta.json"))
    n, h, o = (64, 32), (1.0, 1.0), (0.0, 0.0)
    rd = lambda name: np.fromfile(tmp_path / (name + ".bin"), np.float32)
    col, inflow, outflow = rd("collider"), rd("inflow"), rd("outflow")
    np.testing.assert_array_max_ulp(col, O.polygon_sdf2([scenes.circle_polygon(16, 16, 4, 48)], n, h, o), maxulp=1)
    np.testing.assert_array_equal(rd("u0"), scenes.karman2d(nx=64, ny=32)["u0"])
    cells = n[0] * n[1]
    st = {"u": rd("u0"), "v": np.zeros(n[0] * (n[1] + 1), np.float32), "pressure": np.zeros(cells, np.float32),
          "p_accepted": np.zeros(cells, np.float32)}
    for f in range(2):
        k, dt, diags = O.solver2_frame(st, n, h, o, 1.0, courant=0.5, sdf=col, inflows=[(inflow, (1.0, 0.0), "set")],
                                       outflow_sdf=outflow, flags=12, levels=3, tol=1e-6)
        assert k == meta["frames"][f]["substeps"] and all(d["converged"] for d in diags)
        assert all(d["converged"] for d in meta["frames"][f]["diags"])
    vel = np.concatenate([rd("u"), rd("v")])
    want = np.concatenate([st["u"], st["v"]])
    assert np.linalg.norm(vel - want) <= 2e-5 * np.linalg.norm(want)


@pytest.mark.gpu
def test_async_projection_dot_reducer_and_cg_through_node(tmp_path):
    """project() resolves without blocking the event loop (napi_async_work),
    concurrent calls are serialised; createDotReducer's single-cell probes and
    ramp (examples/31-conjugate-gradient-3d/main.js:168-238); CG on
    A = diag(1..8) (examples/04-conjugate-gradient)."""
    out = tmp_path / "async.json"
    r = run_node("async_linalg.mjs", out, timeout=300)
    print(r.stdout)
    d = json.load(open(out))
    assert d["project"]["diag"]["converged"]
    assert d["project"]["ticks"] > 0  # the event loop ran while the GPU solved
    assert d["both"] == [True, True]
    print("cold", d["coldIterations"], "free-cell coarse space", d["coarseSpace"])
    assert d["coarseSpace"]["converged"] and d["coarseSpace"]["iterations"] < d["coldIterations"]
    assert d["lanes"] == 6
    for p in d["probes"]:
        assert p["total"] == 1 and p["lit"] == [[p["cell"][0], 1]], p
    assert d["ramp"] == 7140
    assert d["cgSucceeded"]
    np.testing.assert_allclose(d["cgX"], 1.0 / np.arange(1, 9), atol=1e-3)


@pytest.mark.gpu
def test_collider_inflow_outflow_through_node_match_oracle(tmp_path):
    """createSDFInflow3 in constrainVelocity, createOutflowPressureDirichlet3
    combined with a user pin by combineDirichlet3 (evaluated on the GPU), and
    MacCormack advection through a collider, from Node, against the oracle"""
    run_node("collider3.mjs", tmp_path, timeout=300)
    meta = json.load(open(tmp_path / "meta.json"))
    n, h, o = (12, 12, 12), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)
    rd = lambda name: np.fromfile(tmp_path / (name + ".bin"), np.float32)
    vin = [rd("vin_" + c) for c in "UVW"]
    want = O.constrain_velocity3(vin, n, h, o, closed_flags=63, depth=5,
                                 inflows=[(rd("inflow_sdf"), (1.5, -0.25, 0.5), "set")])
    for a, c in enumerate("UVW"):
        np.testing.assert_allclose(rd("vout_" + c), want[a], rtol=1e-6, atol=1e-6)
    assert np.count_nonzero(rd("vout_U") == np.float32(1.5)) > 0
    # outflow Dirichlet + pin
    cells = 12 ** 3
    ua = np.zeros(cells, np.uint8)
    ut = np.zeros(cells, np.float32)
    p = meta["pin"][0] + 12 * (meta["pin"][1] + 12 * meta["pin"][2])
    ua[p], ut[p] = 1, 0.75
    wa, wt = O.outflow_dirichlet3([rd("outflow_sdf")], n, h, o, ua, ut)
    mask = np.fromfile(tmp_path / "mask.bin", np.uint8)
    np.testing.assert_array_equal(mask, wa)
    assert meta["diag"]["converged"]
    pr = rd("pressure")
    np.testing.assert_allclose(pr[(wa != 0) & (np.arange(cells) != p)], 0.0, atol=1e-6)
    assert abs(pr[p] - 0.75) <= 1e-6
    assert meta["hostInside"]["active"] and meta["hostPin"]["active"] and meta["hostPin"]["target"] == 0.75
    assert not meta["hostFree"]["active"]
    # advection through the collider
    V = [rd("avel_" + c) for c in "UVW"]
    prev = [np.full(int(np.prod(O.face_shape(n, a))), -3.0, np.float32) for a in range(3)]
    want = O.advect_faces3_c(V, V, n, h, o, rd("collider_sdf"), 1.7, order=2, out=prev)
    for a, c in enumerate("UVW"):
        np.testing.assert_array_equal(rd("aout_" + c), want[a])
