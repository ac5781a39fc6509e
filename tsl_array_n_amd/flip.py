"""createGridFlipSolver3 on the GPU (packages/fluxflow/src/grid/grid_flip_solver3.js:135-1155).

Free-surface FLIP with reduced pressure, int32 fixed-point P2G, extrapolation,
constrainVelocity (optional static collider), deterministic resampling and
the MGPCG pressure projection; optionally a carried concentration (mixing,
fade) coupled to the density (face weights and the reduced-pressure gravity
term), a mass-weighted scatter, a pressure pin for closed all-fluid domains
and phase-aware resampling.  ``on_advance_time_step()`` runs one step;
positions/velocities/fluid_mask/pressure/concentration and the velocity grid
are fields.
"""
import ctypes as C

import numpy as np

from . import (F32, PREC, U8, VEC3, Field, MGOptions, Pressure3Desc, PressureDiag, STOP_NAMES, _check, _lib,
               face_shape, mg_options)


class Flip3Desc(C.Structure):
    _fields_ = [("resolution", C.c_int * 3), ("spacing", C.c_double * 3), ("origin", C.c_double * 3),
                ("max_particles", C.c_int64), ("dt", C.c_float), ("gravity", C.c_float * 3),
                ("reduced_pressure", C.c_int), ("flip_ratio", C.c_float), ("velocity_damping", C.c_float),
                ("p2g_atomic_scale", C.c_float), ("weight_epsilon", C.c_float), ("closed_domain_flags", C.c_int),
                ("resample", C.c_int), ("min_particles_per_cell", C.c_int), ("max_particles_per_cell", C.c_int),
                ("extrapolation_depth", C.c_int), ("max_substeps", C.c_int), ("collider_sdf", C.c_void_p),
                ("collider_weights", C.c_void_p * 3), ("friction", C.c_float), ("pressure", Pressure3Desc),
                ("carry_concentration", C.c_int), ("mixing", C.c_float), ("fade", C.c_float),
                ("concentration_atomic_scale", C.c_float), ("ambient_density", C.c_float),
                ("component_density", C.c_float), ("mass_weighted_transfer", C.c_int), ("pressure_pin", C.c_int),
                ("pin_cell", C.c_int * 3), ("phase_resample", C.c_int), ("collider_push_thresh", C.c_float),
                ("collider_push_shift", C.c_float)]


_FIELDS = {"positions": 0, "velocities": 1, "fluid_mask": 2, "u": 3, "v": 4, "w": 5, "pressure": 6,
           "concentration": 7, "cell_concentration": 8, "cell_density": 9, "dirichlet_active": 10,
           "dirichlet_target": 11, "beta_u": 12, "beta_v": 13, "beta_w": 14, "input_u": 15, "input_v": 16,
           "input_w": 17}


class GridFlipSolver3:
    def __init__(self, resolution, grid_spacing=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0), max_particles=None,
                 dt=None, gravity=(0.0, -9.81, 0.0), reduced_pressure=True, flip_ratio=0.97, velocity_damping=0.02,
                 p2g_atomic_scale=65536.0, weight_epsilon=1e-4, closed_domain_boundary_flag=63, collider_sdf=None,
                 collider_weights=None, friction=0.5, resample=None, pressure=None, max_substeps=32,
                 extrapolation_depth=5, carry_concentration=False, mixing=0.0, fade=0.0,
                 concentration_atomic_scale=4096.0, ambient_density=None, component_density=None,
                 mass_weighted_transfer=False, pressure_pin=None, collider_push_thresh=0.0, collider_push_shift=0.0):
        if not max_particles:
            raise ValueError("createGridFlipSolver3: options.maxParticles is required.")
        if dt is None:
            raise ValueError("createGridFlipSolver3: dt is required.")
        d = Flip3Desc()
        _lib.flux_flip3_desc_default(C.byref(d))
        for a in range(3):
            d.resolution[a] = int(resolution[a])
            d.spacing[a] = float(grid_spacing[a])
            d.origin[a] = float(origin[a])
            d.gravity[a] = float(gravity[a])
        d.max_particles = int(max_particles)
        d.dt = float(dt)
        d.reduced_pressure = int(bool(reduced_pressure))
        d.flip_ratio = flip_ratio
        d.velocity_damping = velocity_damping
        d.p2g_atomic_scale = p2g_atomic_scale
        d.weight_epsilon = weight_epsilon
        d.closed_domain_flags = int(closed_domain_boundary_flag)
        rs = resample or {}
        d.resample = int(bool(rs.get("enabled", True)))
        d.min_particles_per_cell = int(rs.get("minParticlesPerCell", 4))
        d.max_particles_per_cell = int(rs.get("maxParticlesPerCell", 8))
        d.phase_resample = int(bool(rs.get("phaseAware", False)))
        d.carry_concentration = int(bool(carry_concentration))
        d.mixing, d.fade = float(mixing), float(fade)
        d.concentration_atomic_scale = float(concentration_atomic_scale)
        if (ambient_density is None) != (component_density is None):
            raise ValueError("createGridFlipSolver3: ambientDensity and componentDensity go together")
        if ambient_density is not None:
            d.ambient_density, d.component_density = float(ambient_density), float(component_density)
        d.mass_weighted_transfer = int(bool(mass_weighted_transfer))
        if pressure_pin == "topCenter":
            d.pressure_pin = 1
        elif pressure_pin is not None:
            d.pressure_pin = 2
            for a in range(3):
                d.pin_cell[a] = int(pressure_pin[a])
        d.extrapolation_depth = int(extrapolation_depth)
        d.max_substeps = int(max_substeps)
        if collider_sdf is not None:
            d.collider_sdf = collider_sdf.h
        if collider_weights is not None:
            for a in range(3):
                d.collider_weights[a] = collider_weights[a].h
        d.friction = friction
        d.collider_push_thresh = float(collider_push_thresh)
        d.collider_push_shift = float(collider_push_shift)
        p = pressure or {}
        if "multigrid" in p:
            d.pressure.multigrid = mg_options(**p["multigrid"])
        if "preconditioner" in p:
            d.pressure.preconditioner = PREC[p["preconditioner"]]
        for key, field in (("tolerance", "tolerance"), ("maxIterationsFactor", "max_iterations_factor"),
                           ("maxPlausiblePressure", "max_plausible_pressure")):
            if key in p:
                setattr(d.pressure, field, float(p[key]))
        if "maxIterations" in p:
            d.pressure.max_iterations = int(p["maxIterations"])
        if "relativeTolerance" in p:
            d.pressure.relative_tolerance = int(bool(p["relativeTolerance"]))
        if "colliderWeights" in p:
            for a in range(3):
                d.pressure.collider_weights[a] = p["colliderWeights"][a].h
        self._keep = (collider_sdf, collider_weights, p)
        h = C.c_void_p()
        _check(_lib.flux_flip3_create(C.byref(d), C.byref(h)), "createGridFlipSolver3")
        self.h = h.value
        n = tuple(int(x) for x in resolution)
        self.resolution = n
        self.spacing = tuple(float(x) for x in grid_spacing)
        self.positions = self._view("positions", VEC3, (int(max_particles),))
        self.velocities = self._view("velocities", VEC3, (int(max_particles),))
        self.fluid_mask = self._view("fluid_mask", F32, n)
        self.velocity_grid = [self._view(c, F32, face_shape(n, a)) for a, c in enumerate("uvw")]
        self.pressure = self._view("pressure", F32, n)
        self.concentration = self._view("concentration", F32, (int(max_particles),)) if carry_concentration else None
        self.cell_concentration = self._view("cell_concentration", F32, n) if carry_concentration else None
        self.cell_density = self._view("cell_density", F32, n) if ambient_density is not None else None
        self.diagnostics = {}

    def _view(self, name, dtype, shape):
        _lib.flux_flip3_field.restype = C.c_void_p
        ptr = _lib.flux_flip3_field(C.c_void_p(self.h), _FIELDS[name])
        return Field(dtype, shape, handle=ptr, owned=False)

    def capture_projection_input(self, on=True):
        """keep each step's projection input velocity (see captured_system)"""
        _check(_lib.flux_flip3_capture_projection_input(C.c_void_p(self.h), int(bool(on))), "captureProjectionInput")

    def captured_system(self):
        """the last step's pressure system on the host, in the scene form the
        projection tools take: {"shape", "spacing", "vel" (the projection's
        input), "beta" (None without density coupling), "active", "target"}"""
        n = self.resolution
        _lib.flux_flip3_field.restype = C.c_void_p
        if not _lib.flux_flip3_field(C.c_void_p(self.h), _FIELDS["input_u"]):
            raise RuntimeError("captured_system: call capture_projection_input() before the step")
        vel = [self._view("input_" + c, F32, face_shape(n, a)).to_array() for a, c in enumerate("uvw")]
        beta = None
        if _lib.flux_flip3_field(C.c_void_p(self.h), _FIELDS["beta_u"]):
            beta = [self._view("beta_" + c, F32, face_shape(n, a)).to_array() for a, c in enumerate("uvw")]
        return {"shape": n, "spacing": self.spacing, "vel": vel, "beta": beta,
                "active": self._view("dirichlet_active", U8, n).to_array(),
                "target": self._view("dirichlet_target", F32, n).to_array()}

    def set_coarse_space(self, free_indicator=True, fields=()):
        """the additive coarse-space correction of the projection's
        preconditioner (flux_flip3_pressure + flux_pressure3_set_coarse_space):
        the free-cell indicator and up to three caller fields"""
        fields = list(fields)
        self._cs_keep = fields
        _lib.flux_flip3_pressure.restype = C.c_void_p
        ps = _lib.flux_flip3_pressure(C.c_void_p(self.h))
        arr = (C.c_void_p * max(1, len(fields)))(*[f.h for f in fields]) if fields else None
        _check(_lib.flux_pressure3_set_coarse_space(C.c_void_p(ps), int(bool(free_indicator)), len(fields), arr),
               "coarseSpace")

    def set_dt(self, dt):
        """the step's dt from the next step on (the reference's live dt node)"""
        _check(_lib.flux_flip3_set_dt(C.c_void_p(self.h), C.c_float(dt)), "dt")

    def on_advance_time_step(self):
        dg = PressureDiag()
        _check(_lib.flux_flip3_step(C.c_void_p(self.h), C.byref(dg)), "onAdvanceTimeStep")
        self.diagnostics = {"converged": bool(dg.converged), "rejected": bool(dg.rejected),
                            "iterations": dg.iterations, "stoppedBy": STOP_NAMES.get(dg.stopped_by, "unknown"),
                            "residual": dg.residual, "bNorm": dg.b_norm, "badCells": dg.bad_cells}
        return self.diagnostics

    def __del__(self):
        try:
            _lib.flux_flip3_destroy(C.c_void_p(self.h))
        except Exception:
            pass


def create_grid_flip_solver3(**kw):
    return GridFlipSolver3(**kw)


class GridAdaptiveTimeStep3:
    """createGridAdaptiveTimeStep3({velocityGrid, gridSpacing, targetDt,
    courantNumber, maxSubSteps}) (grid_adaptive_timestep3.js:15-44): update()
    reads max |u| over the three face arrays on the GPU, computes the
    sub-step count and dt (computeAdaptiveSubSteps, time/cfl.js:38-48) and
    writes the sub-step dt into the solver (the reference's live dt node);
    returns (sub-step count, sub-step dt as f32)."""

    def __init__(self, solver, target_dt, courant_number=5.0, max_sub_steps=32):
        self.solver = solver
        self.min_spacing = float(min(solver.spacing))
        self.target_dt = float(target_dt)
        self.courant_number = float(courant_number)
        self.max_sub_steps = int(max_sub_steps)
        self.state = {"lastNumSubSteps": 1, "lastSubDt": self.target_dt}

    def update(self):
        from .grid2 import compute_adaptive_substeps, max_abs
        m = max_abs(list(self.solver.velocity_grid))
        k, sub = compute_adaptive_substeps(m, self.min_spacing, self.target_dt, self.courant_number,
                                           self.max_sub_steps)
        sub32 = float(np.float32(sub))
        self.solver.set_dt(sub32)
        self.state = {"lastNumSubSteps": k, "lastSubDt": sub32, "maxVelocity": m}
        return k, sub32


# projection forms of the dam break: the reference's method (red-black 2/2,
# FP32 V-cycle) and the benchmark's headline form (FP16 V-cycle, brick
# smoother with 12 sweeps per pass, free-cell coarse correction)
PROJECTION_FORMS = {
    "plain": {"storage": "f32", "smoother": "rbgs", "nu": 2, "coarse_space": False},
    "headline": {"storage": "f16", "smoother": "brick", "nu": 12, "coarse_space": True},
}


def two_phase_dam_break(n, dt=1.0 / 60.0, levels=None, flip_ratio=0.95, ratio=1000.0, storage="f32", seed=0,
                        particles=None, smoother="rbgs", nu=2, coarse_space=False):
    """createGridFlipSolver3 on the two-phase dam break (the liquid block
    x < 0.4, y < 0.6 of the unit domain, gas everywhere else, 2^3 particles
    per cell, density ratio `ratio`:1, reduced pressure, gravity (0, -9.8, 0),
    FLIP ratio 0.95, the fixed-point (deterministic) scatter weighted by
    mass, phase-aware resampling, the closed box pinned at its top-centre
    cell).  The projection's V-cycle takes `storage`, `smoother` and `nu`
    sweeps (down, up and final); `coarse_space` adds the free-cell coarse
    correction.  Returns (solver, particles)."""
    from . import scenes
    p = particles or scenes.flip_dam_break(n, seed=seed)
    levels = levels or max(2, int(np.log2(n)) - 2)
    s = GridFlipSolver3(resolution=(n, n, n), grid_spacing=p["spacing"], max_particles=p["count"], dt=dt,
                        gravity=(0.0, -9.8, 0.0), flip_ratio=flip_ratio, velocity_damping=0.0,
                        carry_concentration=True, ambient_density=1.0, component_density=float(ratio),
                        mass_weighted_transfer=True, pressure_pin="topCenter",
                        resample={"enabled": True, "phaseAware": True},
                        pressure={"multigrid": {"levels": levels, "storage": storage, "smoother": smoother,
                                                "nu_down": nu, "nu_up": nu, "n_final": nu},
                                  "tolerance": 1e-6})
    if coarse_space:
        s.set_coarse_space(True)
    s.positions.from_array(p["positions"])
    s.concentration.from_array(p["concentration"])
    return s, p


def main(argv=None):
    """python -m tsl_array_n_amd.flip --n 256 --steps 3: the two-phase FLIP
    dam break on one GPU; prints each step's projection diagnostics and
    property checks (particle count and concentration kept, every
    projection converged to the FP64 tolerance, velocities bounded) and the
    time per step."""
    import argparse
    import sys
    import time
    from . import init, sync
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--storage", default=None, help="V-cycle storage (overrides the form's)")
    ap.add_argument("--form", default="plain", choices=sorted(PROJECTION_FORMS),
                    help="projection form: plain (the reference's method) or headline (FP16 V-cycle, brick "
                         "smoother 12 sweeps, free-cell coarse correction)")
    ap.add_argument("--frames", type=int, default=0,
                    help="adaptive mode: this many frames of 1/60 s, each split into sub-steps by the CFL condition "
                         "(createGridAdaptiveTimeStep3) instead of --steps fixed steps of 1/60 s")
    ap.add_argument("--courant", type=float, default=1.0, help="Courant number of the adaptive mode")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy a 512 MiB field three times (the known-size kernel that scales FETCH_SIZE / "
                         "WRITE_SIZE counter readings to bytes)")
    a = ap.parse_args(argv)
    init(0)
    if a.copy_calibration:
        from . import array_n
        src = array_n("float", (128 << 20,)).fill(1.0)
        dst = array_n("float", (128 << 20,))
        for _ in range(3):
            dst.copy_from(src)
        sync()
        del src, dst
    t0 = time.perf_counter()
    form = dict(PROJECTION_FORMS[a.form])
    if a.storage:
        form["storage"] = a.storage
    s, p = two_phase_dam_break(a.n, **form)
    print("flip: %d^3, %d particles (%d liquid), setup %.1f s" % (a.n, p["count"], p["liquid"],
                                                                   time.perf_counter() - t0), flush=True)
    conc0 = float(p["concentration"].astype(np.float64).sum())
    ok = True
    times = []

    def check(label, d, ms):
        vel = s.velocities.to_array()
        conc = s.concentration.to_array()
        vmax = float(np.abs(vel).max())
        csum = float(conc.astype(np.float64).sum())
        good = (d["converged"] and not d["rejected"] and np.isfinite(vmax) and vmax < 50.0 and csum == conc0 and
                np.all(np.isfinite(s.positions.to_array())))
        print("flip %s: %.1f ms, %s, max|v| %.3f, conc sum %.1f (start %.1f), fluid cells %d" % (
            label, ms, d, vmax, csum, conc0, int((s.fluid_mask.to_array() > 0.5).sum())), flush=True)
        return good

    if a.frames > 0:
        ats = GridAdaptiveTimeStep3(s, 1.0 / 60.0, courant_number=a.courant)
        steps = 0
        t_all = time.perf_counter()
        for f in range(a.frames):
            nsub, sub = ats.update()
            for q in range(nsub):
                sync()
                t = time.perf_counter()
                d = s.on_advance_time_step()
                sync()
                ms = (time.perf_counter() - t) * 1e3
                times.append(ms)
                steps += 1
                if q == nsub - 1:
                    ok = check("frame %d (%d sub-steps of %.3g s, max|u| %.3f)" % (
                        f, nsub, sub, ats.state["maxVelocity"]), d, ms) and ok
                else:
                    ok = ok and d["converged"] and not d["rejected"]
        print("flip: %d frames in %d steps, %.1f s" % (a.frames, steps, time.perf_counter() - t_all), flush=True)
    for k in range(a.steps if a.frames <= 0 else 0):
        sync()
        t = time.perf_counter()
        d = s.on_advance_time_step()
        sync()
        ms = (time.perf_counter() - t) * 1e3
        times.append(ms)
        ok = check("step %d" % k, d, ms) and ok
    if len(times) > 1:
        print("flip: form %s, mean step time over steps 1..%d: %.1f ms" % (a.form, len(times) - 1,
                                                                        float(np.mean(times[1:]))), flush=True)
    print("flip: %s" % ("ok" if ok else "FAILED"), flush=True)
    sys.exit(0 if ok else 1)


if __name__ == "__main__":
    main()
