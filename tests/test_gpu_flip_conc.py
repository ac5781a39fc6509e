"""createGridFlipSolver3's carried concentration, density coupling,
mass-weighted scatter, pressure pin and phase-aware resampling on the GPU
against the oracle's composed step (grid_flip_solver3.js:265-357, 493-531,
848-880, 967-1049; grid_two_phase_flip_solver2.js:133-169, 1131-1503).

The same bar as the plain FLIP step: the fluid mask exactly, positions to
1e-5, velocities to 1e-4 of their norm (the projections agree to the PCG
tolerance), the cell concentration exactly (integer sums) and the carried
concentration to 1e-5."""
import numpy as np
import pytest

import tsl_array_n_amd as F
from oracle import pyoracle as O
from tsl_array_n_amd.flip import GridFlipSolver3

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30)


def stratified_box(n, heavy_below, per_axis=2, seed=3, bump=0.0, fill=1.0):
    g = np.stack(np.meshgrid(*[(np.arange(n[a] * per_axis) + 0.5) / per_axis for a in range(3)], indexing="ij"),
                 -1).reshape(-1, 3)
    g = g[g[:, 1] < n[1] * fill]
    rng = np.random.default_rng(seed)
    pos = (g + rng.uniform(-0.1, 0.1, g.shape)).astype(np.float32)
    low = pos[:, 1] < n[1] * fill / 2 + bump * np.sin(2 * np.pi * pos[:, 0] / n[0])
    conc = (low if heavy_below else ~low).astype(np.float32)
    return pos.ravel(), conc


def run_pair(n, pos, conc, steps, gpu_kw, oracle_kw, dt=0.05, levels=2):
    count = conc.size
    s = GridFlipSolver3(resolution=n, grid_spacing=(1.0, 1.0, 1.0), max_particles=count, dt=dt,
                        pressure={"multigrid": {"levels": levels}}, carry_concentration=True, **gpu_kw)
    s.positions.from_array(pos)
    s.concentration.from_array(conc)
    cells = int(np.prod(n))
    st = {"pos": pos.copy(), "vel": np.zeros_like(pos), "conc": conc.copy(), "pressure": np.zeros(cells, np.float32),
          "p_accepted": np.zeros(cells, np.float32)}
    for a, c in enumerate("uvw"):
        st[c] = np.zeros(int(np.prod(F.face_shape(n, a))), np.float32)
    for step in range(steps):
        dg = s.on_advance_time_step()
        do = O.flip3_step(st, n, (1.0, 1.0, 1.0), (0, 0, 0), dt, levels=levels, carry_concentration=True,
                          **oracle_kw)
        print("step", step, "gpu", dg, "oracle", do)
        assert dg["converged"] and do["converged"] and not dg["rejected"]
        np.testing.assert_array_equal(s.fluid_mask.to_array(), st["mask"])
        assert np.max(np.abs(s.positions.to_array() - st["pos"])) <= 1e-5
        assert rel(s.velocities.to_array(), st["vel"]) <= 1e-4
        scale = np.linalg.norm(np.concatenate([st[c] for c in "uvw"]))
        for a, c in enumerate("uvw"):
            assert np.linalg.norm(s.velocity_grid[a].to_array() - st[c]) <= 1e-4 * max(scale, 1e-6)
        np.testing.assert_array_equal(s.cell_concentration.to_array(), st["cell_conc"])
        if st.get("cell_density") is not None:
            np.testing.assert_array_equal(s.cell_density.to_array(), st["cell_density"])
        assert np.max(np.abs(s.concentration.to_array() - st["conc"])) <= 1e-5
    return s, st


def test_pinned_stable_box_matches_oracle_and_rests(flux):
    n = (8, 8, 8)
    pos, conc = stratified_box(n, heavy_below=True)
    s, st = run_pair(n, pos, conc, 3, dict(ambient_density=1.0, component_density=3.0, pressure_pin="topCenter",
                                           resample={"enabled": False}),
                     dict(density=(1.0, 3.0), pressure_pin="topCenter", resample=False))
    assert np.abs(s.velocities.to_array()).max() < 1e-3


def test_pinned_unstable_box_matches_oracle_and_overturns(flux):
    n = (8, 8, 8)
    pos, conc = stratified_box(n, heavy_below=False, bump=1.0)
    s, st = run_pair(n, pos, conc, 3, dict(ambient_density=1.0, component_density=3.0, pressure_pin="topCenter",
                                           resample={"enabled": False}, mixing=0.3, fade=0.01),
                     dict(density=(1.0, 3.0), pressure_pin="topCenter", resample=False, mixing=0.3, fade=0.01))
    v = s.velocities.to_array().reshape(-1, 3)
    heavy = conc > 0.5
    assert v[heavy, 1].mean() < 0 < v[~heavy, 1].mean()


def test_free_surface_mass_weighted_phase_resample_matches_oracle(flux):
    """a half-filled box (air above), two liquids of density 1 and 2 with the
    heavier on top, mass-weighted scatter and phase-aware resampling"""
    n = (12, 12, 8)
    pos, conc = stratified_box(n, heavy_below=False, fill=0.6, seed=5)
    run_pair(n, pos, conc, 3, dict(ambient_density=1.0, component_density=2.0, mass_weighted_transfer=True,
                                   resample={"phaseAware": True, "minParticlesPerCell": 6, "maxParticlesPerCell": 9}),
             dict(density=(1.0, 2.0), mass_weighted=True, phase_resample=True, min_pc=6, max_pc=9), levels=2)


def test_concentration_without_coupling_matches_oracle(flux):
    """carried concentration only: the plain free-surface step plus the cell
    concentration and mixAndFade"""
    n = (12, 12, 8)
    pos, conc = stratified_box(n, heavy_below=True, fill=0.5, seed=7)
    run_pair(n, pos, conc, 2, dict(mixing=0.5, fade=0.05), dict(mixing=0.5, fade=0.05))


def test_options_are_checked(flux):
    with pytest.raises(F.FluxError, match="carryConcentration"):
        GridFlipSolver3(resolution=(4, 4, 4), max_particles=8, dt=0.1, ambient_density=1.0, component_density=2.0)
    with pytest.raises(F.FluxError, match="must be > 0"):
        GridFlipSolver3(resolution=(4, 4, 4), max_particles=8, dt=0.1, carry_concentration=True,
                        ambient_density=-1.0, component_density=2.0)
    with pytest.raises(F.FluxError, match="pressurePin"):
        GridFlipSolver3(resolution=(4, 4, 4), max_particles=8, dt=0.1, pressure_pin=(4, 0, 0))
