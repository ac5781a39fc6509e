"""The 2-D Eulerian solver around the projection, on the GPU.

Host mirror of packages/fluxflow/src/grid/: createSDFStaticCollider2 /
createSDFInflow2 / createSDFOutflow2 (sdf_collider2.js, sdf_inflow_outflow2.js),
createSemiLagrangianAdvectionSolver2 (order 1), createGridBlockedBoundary-
ConditionSolver2, createGridOutflowSolver2, createGridPressureSolver2,
createGridSolver2 and createGridAdaptiveTimeStep2 (time/cfl.js).  Every
stage runs in libfluxhip.so; polygons are rasterised to SDF fields on the
GPU in FP64.
"""
import ctypes as C

import numpy as np

from . import F32, PREC, Field, MGOptions, PressureDiag, STOP_NAMES, _check, _lib, array_n, mg_options

DIRECTION_LEFT, DIRECTION_RIGHT, DIRECTION_DOWN, DIRECTION_UP = 1, 2, 4, 8
DIRECTION_ALL = 15


class Grid2(C.Structure):
    _fields_ = [("n", C.c_int * 2), ("h", C.c_float * 2), ("origin", C.c_float * 2)]


def grid2(resolution, spacing=(1.0, 1.0), origin=(0.0, 0.0)):
    g = Grid2()
    for a in range(2):
        g.n[a] = int(resolution[a])
        g.h[a] = float(spacing[a])
        g.origin[a] = float(origin[a])
    return g


def face_shape2(resolution, axis):
    return (resolution[0] + 1, resolution[1]) if axis == 0 else (resolution[0], resolution[1] + 1)


class SDFStaticCollider2:
    """createSDFStaticCollider2(nx, ny, hx, hy, ox, oy): the SDF lives in a
    cell-centred float field; addPolygons rasterises it on the GPU."""

    def __init__(self, nx, ny, hx=1.0, hy=1.0, ox=0.0, oy=0.0):
        self.grid = grid2((nx, ny), (hx, hy), (ox, oy))
        self.resolution = (int(nx), int(ny))
        # a collider with no polygon reads 0 everywhere, as the reference's zeroed grid
        self.sdf = array_n("float", self.resolution, np.zeros(nx * ny, np.float32))
        self.friction_coefficient = 0.5

    def add_polygons(self, polygons, invert=False):
        pts, starts = [], [0]
        for ring in polygons:
            for x, y in ring:
                pts += [float(x), float(y)]
            starts.append(len(pts) // 2)
        xy = np.array(pts, np.float64)
        st = np.array(starts, np.int32)
        _check(_lib.flux_polygon_sdf2(C.c_void_p(self.sdf.h), C.byref(self.grid), xy.ctypes.data_as(C.c_void_p),
                                      st.ctypes.data_as(C.c_void_p), len(polygons), int(bool(invert))),
               "addPolygons")

    def add_polygon(self, points, invert=False):
        self.add_polygons([points], invert)


class SDFInflow2(SDFStaticCollider2):
    """createSDFInflow2(..., {velocity, mode})"""

    def __init__(self, nx, ny, hx=1.0, hy=1.0, ox=0.0, oy=0.0, velocity=(0.0, 0.0), mode="set"):
        super().__init__(nx, ny, hx, hy, ox, oy)
        if mode not in ("set", "add"):
            raise ValueError("createSDFInflow2: mode must be 'set' or 'add'")
        self.velocity = (float(velocity[0]), float(velocity[1]))
        self.mode = mode


class SDFOutflow2(SDFStaticCollider2):
    """createSDFOutflow2(...)"""


def advect_face_centered2(g, vin, vout, dt, collider=None, max_substeps=32):
    """advectFaceCentered2 (order 1) from vin = (u, v) into vout = (u, v)"""
    _check(_lib.flux_advect_faces2(C.byref(g), C.c_void_p(vin[0].h), C.c_void_p(vin[1].h),
                                   C.c_void_p(collider.sdf.h if collider is not None else None), C.c_float(dt),
                                   int(max_substeps), C.c_void_p(vout[0].h), C.c_void_p(vout[1].h)),
           "advectFaceCentered2")


class SemiLagrangianAdvectionSolver2:
    """createSemiLagrangianAdvectionSolver2({velocityGrid, collider, dt,
    maxSubsteps, order}) (advection_solver2.js:174-526) for face grids: the
    velocity pair and the collider are bound here; order 1 or 2 (MacCormack)."""

    def __init__(self, g, velocity_grid, dt, collider=None, max_substeps=32, order=1):
        if order not in (1, 2):
            raise ValueError("createSemiLagrangianAdvectionSolver2: order must be 1 or 2")
        self.velocity_grid = tuple(velocity_grid)
        self.collider = collider
        h = C.c_void_p()
        _check(_lib.flux_advection2_create(C.byref(g), C.c_void_p(velocity_grid[0].h), C.c_void_p(velocity_grid[1].h),
                                           C.c_void_p(collider.sdf.h if collider is not None else None),
                                           C.c_float(dt), int(max_substeps), int(order), C.byref(h)),
               "createSemiLagrangianAdvectionSolver2")
        self.h = h.value

    def set_dt(self, dt):
        _check(_lib.flux_advection2_set_dt(C.c_void_p(self.h), C.c_float(dt)), "advection dt")

    def advect_face_centered2(self, vin, vout):
        _check(_lib.flux_advection2_faces(C.c_void_p(self.h), C.c_void_p(vin[0].h), C.c_void_p(vin[1].h),
                                          C.c_void_p(vout[0].h), C.c_void_p(vout[1].h)), "advectFaceCentered2")

    def __del__(self):
        try:
            _lib.flux_advection2_destroy(C.c_void_p(self.h))
        except Exception:
            pass


def max_abs(fields):
    arr = (C.c_void_p * len(fields))(*[f.h for f in fields])
    out = C.c_float()
    _check(_lib.flux_max_abs(arr, len(fields), C.byref(out)), "createMaxAbsReducer")
    return out.value


def compute_adaptive_substeps(max_velocity, min_spacing, frame_dt, courant_number=5.0, max_sub_steps=32):
    steps = C.c_int()
    sub = C.c_double()
    _check(_lib.flux_adaptive_substeps(C.c_double(max_velocity), C.c_double(min_spacing), C.c_double(frame_dt),
                                       C.c_double(courant_number), int(max_sub_steps), C.byref(steps),
                                       C.byref(sub)), "computeAdaptiveSubSteps")
    return steps.value, sub.value


class GridBlockedBoundaryConditionSolver2:
    """createGridBlockedBoundaryConditionSolver2(velocity, nx, ny, hx, hy, ox, oy, collider, inflows)"""

    def __init__(self, velocity, nx, ny, hx=1.0, hy=1.0, ox=0.0, oy=0.0, collider=None, inflows=None):
        self.grid = grid2((nx, ny), (hx, hy), (ox, oy))
        vel = (C.c_void_p * 2)(velocity[0].h, velocity[1].h)
        h = C.c_void_p()
        _check(_lib.flux_boundary2_create(C.byref(self.grid), vel,
                                          C.c_void_p(collider.sdf.h if collider is not None else None),
                                          C.c_float(collider.friction_coefficient if collider is not None else 0.5),
                                          C.byref(h)), "createGridBlockedBoundaryConditionSolver2")
        self.h = h.value
        self._keep = (velocity, collider, inflows)
        lst = [] if inflows is None else (inflows if isinstance(inflows, (list, tuple)) else [inflows])
        for q in lst:
            _check(_lib.flux_boundary2_add_inflow(C.c_void_p(self.h), C.c_void_p(q.sdf.h), C.c_float(q.velocity[0]),
                                                  C.c_float(q.velocity[1]), 1 if q.mode == "add" else 0),
                   "setInflows")
        self._flag = DIRECTION_ALL

    @property
    def closed_domain_boundary_flag(self):
        return self._flag

    @closed_domain_boundary_flag.setter
    def closed_domain_boundary_flag(self, flags):
        _check(_lib.flux_boundary2_set_closed(C.c_void_p(self.h), int(flags)))
        self._flag = int(flags)

    def constrain_velocity(self, extrapolation_depth=5):
        _check(_lib.flux_boundary2_constrain(C.c_void_p(self.h), int(extrapolation_depth)), "constrainVelocity")

    def __del__(self):
        try:
            from . import _initialised
            if _initialised and self.h:
                _lib.flux_boundary2_destroy(C.c_void_p(self.h))
        except Exception:
            pass


class GridOutflowSolver2:
    """createGridOutflowSolver2({velocityGrid, outflows, dt}): one outflow per object"""

    def __init__(self, g, outflow):
        self.grid = g
        h = C.c_void_p()
        _check(_lib.flux_outflow2_create(C.byref(g), C.c_void_p(outflow.sdf.h), C.byref(h)),
               "createGridOutflowSolver2")
        self.h = h.value
        self._keep = outflow

    def apply_outflow_velocity_bc(self, velocity, dt):
        _check(_lib.flux_outflow2_apply(C.c_void_p(self.h), C.c_void_p(velocity[0].h), C.c_void_p(velocity[1].h),
                                        C.c_float(dt)), "applyOutflowVelocityBC")

    def __del__(self):
        try:
            from . import _initialised
            if _initialised and self.h:
                _lib.flux_outflow2_destroy(C.c_void_p(self.h))
        except Exception:
            pass


def outflow_dirichlet2(g, outflows, user=None):
    """createOutflowPressureDirichlet2 + combineDirichlet, evaluated at the
    cell centres: (active u8 field, target f32 field)"""
    n = (g.n[0], g.n[1])
    active = array_n("u8", n)
    target = array_n("float", n)
    arr = (C.c_void_p * max(len(outflows), 1))(*[o.sdf.h for o in outflows])
    _check(_lib.flux_outflow2_dirichlet(C.byref(g), arr, len(outflows),
                                        C.c_void_p(user[0].h if user else None),
                                        C.c_void_p(user[1].h if user else None), C.c_void_p(active.h),
                                        C.c_void_p(target.h)), "createOutflowPressureDirichlet2")
    return active, target


class Pressure2Desc(C.Structure):
    _fields_ = [("resolution", C.c_int * 2), ("spacing", C.c_double * 2), ("origin", C.c_double * 2),
                ("dirichlet_active", C.c_void_p), ("dirichlet_target", C.c_void_p),
                ("face_weights", C.c_void_p * 2), ("multigrid", MGOptions), ("preconditioner", C.c_int),
                ("tolerance", C.c_double), ("max_iterations", C.c_int), ("max_iterations_factor", C.c_double),
                ("relative_tolerance", C.c_int), ("max_plausible_pressure", C.c_double)]


def _diag(dg):
    return {"converged": bool(dg.converged), "rejected": bool(dg.rejected), "iterations": dg.iterations,
            "stoppedBy": STOP_NAMES.get(dg.stopped_by, "unknown"), "residual": dg.residual, "bNorm": dg.b_norm,
            "badCells": dg.bad_cells}


class GridPressureSolver2:
    """createGridPressureSolver2({resolution, gridSpacing, dirichlet, faceWeights, multigrid, ...});
    dirichlet is (active u8 field, target f32 field)."""

    def __init__(self, resolution, grid_spacing=(1.0, 1.0), origin=(0.0, 0.0), dirichlet=None, face_weights=None,
                 multigrid=None, preconditioner="multigrid", tolerance=1e-6, max_iterations=None,
                 max_iterations_factor=1.5, relative_tolerance=True, max_plausible_pressure=1e6):
        d = Pressure2Desc()
        _lib.flux_pressure2_desc_default(C.byref(d))
        for a in range(2):
            d.resolution[a] = int(resolution[a])
            d.spacing[a] = float(grid_spacing[a])
            d.origin[a] = float(origin[a])
        if dirichlet is not None:
            d.dirichlet_active, d.dirichlet_target = dirichlet[0].h, dirichlet[1].h
        if face_weights is not None:
            d.face_weights[0], d.face_weights[1] = face_weights[0].h, face_weights[1].h
        d.multigrid = mg_options(**(multigrid or {}))
        if preconditioner not in PREC:
            raise ValueError("grid_pressure_solver2: unknown preconditioner '%s'" % preconditioner)
        d.preconditioner = PREC[preconditioner]
        d.tolerance = tolerance
        d.max_iterations = int(max_iterations or 0)
        d.max_iterations_factor = max_iterations_factor
        d.relative_tolerance = int(bool(relative_tolerance))
        d.max_plausible_pressure = max_plausible_pressure
        h = C.c_void_p()
        _check(_lib.flux_pressure2_create(C.byref(d), C.byref(h)), "createGridPressureSolver2")
        self.h = h.value
        self.resolution = (int(resolution[0]), int(resolution[1]))
        self._keep = (dirichlet, face_weights)
        _lib.flux_pressure2_field.restype = C.c_void_p
        self.pressure = Field(F32, self.resolution, handle=_lib.flux_pressure2_field(C.c_void_p(self.h), 0),
                              owned=False)
        self.b = Field(F32, self.resolution, handle=_lib.flux_pressure2_field(C.c_void_p(self.h), 1), owned=False)
        self.diagnostics = {}

    def project(self, vin, vout=None):
        vout = vin if vout is None else vout
        dg = PressureDiag()
        _check(_lib.flux_pressure2_project(C.c_void_p(self.h), C.c_void_p(vin[0].h), C.c_void_p(vin[1].h),
                                           C.c_void_p(vout[0].h), C.c_void_p(vout[1].h), C.byref(dg)), "project")
        self.diagnostics = _diag(dg)
        return self.diagnostics

    def __del__(self):
        try:
            from . import _initialised
            if _initialised and self.h:
                _lib.flux_pressure2_destroy(C.c_void_p(self.h))
        except Exception:
            pass


class Solver2Desc(C.Structure):
    _fields_ = [("grid", Grid2), ("collider_sdf", C.c_void_p), ("friction", C.c_float),
                ("advect_collider", C.c_int), ("max_substeps", C.c_int), ("n_inflows", C.c_int),
                ("inflow_sdf", C.c_void_p * 4), ("inflow_velocity", (C.c_float * 2) * 4), ("inflow_mode", C.c_int * 4),
                ("n_outflows", C.c_int), ("outflow_sdf", C.c_void_p * 4), ("outflow_velocity_bc", C.c_int),
                ("dirichlet_active", C.c_void_p), ("dirichlet_target", C.c_void_p), ("closed_flags", C.c_int),
                ("extrapolation_depth", C.c_int), ("multigrid", MGOptions), ("preconditioner", C.c_int),
                ("tolerance", C.c_double), ("max_iterations", C.c_int), ("max_iterations_factor", C.c_double),
                ("relative_tolerance", C.c_int), ("max_plausible_pressure", C.c_double),
                ("advection_order", C.c_int)]


_S2 = {"u": 0, "v": 1, "u_prev": 2, "v_prev": 3, "pressure": 4, "dirichlet": 5}


class GridSolver2:
    """createGridSolver2({velocityGrid resolution, gridSpacing, origin, collider, inflows, outflows,
    closedDomainBoundaryFlag, advection: {collider, maxSubsteps, order}, pressure: {...}}) with no external
    force.  The solver owns its velocity (``velocity_grid`` = (U, V) fields, zero at creation);
    ``on_advance_time_step(dt)`` runs one step and returns the pressure diagnostics."""

    def __init__(self, resolution, grid_spacing=(1.0, 1.0), origin=(0.0, 0.0), collider=None, inflows=None,
                 outflows=None, outflow_velocity_bc=True, closed_domain_boundary_flag=None, dirichlet=None,
                 advection=None, pressure=None, extrapolation_depth=5):
        d = Solver2Desc()
        _lib.flux_solver2_desc_default(C.byref(d))
        d.grid = grid2(resolution, grid_spacing, origin)
        self.resolution = (int(resolution[0]), int(resolution[1]))
        self.grid = d.grid
        if collider is not None:
            d.collider_sdf = collider.sdf.h
            d.friction = collider.friction_coefficient
        adv = advection or {}
        if adv.get("order", 1) not in (1, 2):
            raise ValueError("createGridSolver2: advection.order must be 1 or 2")
        d.advection_order = int(adv.get("order", 1))
        d.advect_collider = int(adv.get("collider") is not None)
        if adv.get("collider") is not None and adv["collider"] is not collider:
            raise ValueError("createGridSolver2: advection.collider must be the solver's collider")
        d.max_substeps = int(adv.get("maxSubsteps", 32))
        inl = [] if inflows is None else (inflows if isinstance(inflows, (list, tuple)) else [inflows])
        outl = [] if outflows is None else (outflows if isinstance(outflows, (list, tuple)) else [outflows])
        if len(inl) > 4 or len(outl) > 4:
            raise ValueError("createGridSolver2: at most 4 inflows and 4 outflows")
        d.n_inflows = len(inl)
        for q, f in enumerate(inl):
            d.inflow_sdf[q] = f.sdf.h
            d.inflow_velocity[q][0], d.inflow_velocity[q][1] = f.velocity
            d.inflow_mode[q] = 1 if f.mode == "add" else 0
        d.n_outflows = len(outl)
        for q, f in enumerate(outl):
            d.outflow_sdf[q] = f.sdf.h
        d.outflow_velocity_bc = int(bool(outflow_velocity_bc))
        if dirichlet is not None:
            d.dirichlet_active, d.dirichlet_target = dirichlet[0].h, dirichlet[1].h
        if closed_domain_boundary_flag is not None:
            d.closed_flags = int(closed_domain_boundary_flag)
        d.extrapolation_depth = int(extrapolation_depth)
        p = pressure or {}
        if "multigrid" in p:
            d.multigrid = mg_options(**p["multigrid"])
        if "preconditioner" in p:
            d.preconditioner = PREC[p["preconditioner"]]
        if "tolerance" in p:
            d.tolerance = float(p["tolerance"])
        if "maxIterations" in p:
            d.max_iterations = int(p["maxIterations"])
        h = C.c_void_p()
        _check(_lib.flux_solver2_create(C.byref(d), C.byref(h)), "createGridSolver2")
        self.h = h.value
        self._keep = (collider, inl, outl, dirichlet)
        self.diagnostics = {}

    def field(self, name):
        _lib.flux_solver2_field.restype = C.c_void_p
        ptr = _lib.flux_solver2_field(C.c_void_p(self.h), _S2[name])
        if not ptr:
            return None
        n = self.resolution
        if name in ("u", "u_prev"):
            shape = face_shape2(n, 0)
        elif name in ("v", "v_prev"):
            shape = face_shape2(n, 1)
        else:
            shape = n
        return Field("u8" if name == "dirichlet" else F32, shape, handle=ptr, owned=False)

    @property
    def velocity_grid(self):
        return (self.field("u"), self.field("v"))

    @property
    def pressure(self):
        return self.field("pressure")

    def on_advance_time_step(self, dt):
        dg = PressureDiag()
        _check(_lib.flux_solver2_step(C.c_void_p(self.h), C.c_float(dt), C.byref(dg)), "onAdvanceTimeStep")
        self.diagnostics = _diag(dg)
        return self.diagnostics

    def __del__(self):
        try:
            from . import _initialised
            if _initialised and self.h:
                _lib.flux_solver2_destroy(C.c_void_p(self.h))
        except Exception:
            pass


class GridAdaptiveTimeStep2:
    """createGridAdaptiveTimeStep2({velocityGrid, gridSpacing, targetDt, courantNumber, maxSubSteps}):
    update() reads max |u| over the face arrays and returns (sub-step count, sub-step dt as f32)."""

    def __init__(self, velocity, grid_spacing, target_dt, courant_number=5.0, max_sub_steps=32):
        self.velocity = velocity
        self.min_spacing = float(min(grid_spacing))
        self.target_dt = float(target_dt)
        self.courant_number = float(courant_number)
        self.max_sub_steps = int(max_sub_steps)
        self.state = {"lastNumSubSteps": 1, "lastSubDt": self.target_dt}

    def update(self):
        m = max_abs(list(self.velocity))
        k, sub = compute_adaptive_substeps(m, self.min_spacing, self.target_dt, self.courant_number,
                                           self.max_sub_steps)
        sub32 = float(np.float32(sub))
        self.state = {"lastNumSubSteps": k, "lastSubDt": sub32, "maxVelocity": m}
        return k, sub32


def create_karman2d(scene, order=1):
    """The 2-D Karman scene (scenes.karman2d) as a GridSolver2 with its
    collider, inflow strip and outflow slab, the initial velocity uploaded,
    and its adaptive time step (order: the advection's, 1 or MacCormack 2);
    returns (solver, stepper, keep)."""
    nx, ny = scene["resolution"]
    h, o = scene["spacing"], scene["origin"]
    collider = SDFStaticCollider2(nx, ny, h[0], h[1], o[0], o[1])
    collider.add_polygons(scene["collider"])
    inflow = SDFInflow2(nx, ny, h[0], h[1], o[0], o[1], velocity=scene["inflow_velocity"], mode="set")
    inflow.add_polygons(scene["inflow"])
    outflow = SDFOutflow2(nx, ny, h[0], h[1], o[0], o[1])
    outflow.add_polygons(scene["outflow"])
    solver = GridSolver2((nx, ny), h, o, collider=collider, inflows=inflow, outflows=outflow,
                         closed_domain_boundary_flag=scene["closed_flags"],
                         advection={"collider": collider if scene["advect_collider"] else None, "order": order},
                         pressure={"multigrid": scene["multigrid"], "tolerance": scene["tolerance"]})
    u, v = solver.velocity_grid
    u.from_array(scene["u0"])
    v.from_array(scene["v0"])
    stepper = GridAdaptiveTimeStep2((u, v), h, scene["frame_dt"], scene["courant"], scene["max_sub_steps"])
    return solver, stepper, (collider, inflow, outflow)


def advance_frame(solver, stepper):
    """one rendered frame: update() then that many sub-steps; returns the diagnostics of each"""
    k, dt = stepper.update()
    return [solver.on_advance_time_step(dt) for _ in range(k)]
