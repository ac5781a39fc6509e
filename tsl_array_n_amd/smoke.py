"""createGridSmokeSolver3 on the GPU (packages/fluxflow/src/grid/grid_smoke_solver3.js:25-154).

Buoyancy, constrainVelocity, projection, semi-Lagrangian advection (order 1,
or MacCormack with ``advection={"order": 2}``), then density and temperature
advected and decayed.  ``on_advance_time_step()``
runs one step; ``source_sphere`` is the example scenes' re-injected ball.
"""
import ctypes as C

from . import F32, PREC, Field, Pressure3Desc, PressureDiag, STOP_NAMES, _check, _lib, face_shape, mg_options


class Smoke3Desc(C.Structure):
    _fields_ = [("resolution", C.c_int * 3), ("spacing", C.c_double * 3), ("origin", C.c_double * 3),
                ("dt", C.c_float), ("buoyancy_smoke_density_factor", C.c_float),
                ("buoyancy_temperature_factor", C.c_float), ("ambient_temperature", C.c_float),
                ("smoke_decay", C.c_float), ("temperature_decay", C.c_float), ("up", C.c_float * 3),
                ("closed_domain_flags", C.c_int), ("max_substeps", C.c_int), ("extrapolation_depth", C.c_int),
                ("pressure", Pressure3Desc), ("advection_order", C.c_int)]


_FIELDS = {"u": 0, "v": 1, "w": 2, "density": 3, "temperature": 4, "pressure": 5}


class GridSmokeSolver3:
    def __init__(self, resolution, grid_spacing=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0), dt=None,
                 buoyancy_smoke_density_factor=-0.000625, buoyancy_temperature_factor=5.0, ambient_temperature=0.0,
                 smoke_decay=0.001, temperature_decay=0.001, up=(0.0, 1.0, 0.0), closed_domain_boundary_flag=63,
                 pressure=None, advection=None):
        if dt is None:
            raise ValueError("createGridSmokeSolver3: dt is required.")
        d = Smoke3Desc()
        _lib.flux_smoke3_desc_default(C.byref(d))
        for a in range(3):
            d.resolution[a] = int(resolution[a])
            d.spacing[a] = float(grid_spacing[a])
            d.origin[a] = float(origin[a])
            d.up[a] = float(up[a])
        d.dt = float(dt)
        d.buoyancy_smoke_density_factor = buoyancy_smoke_density_factor
        d.buoyancy_temperature_factor = buoyancy_temperature_factor
        d.ambient_temperature = ambient_temperature
        d.smoke_decay = smoke_decay
        d.temperature_decay = temperature_decay
        d.closed_domain_flags = int(closed_domain_boundary_flag)
        adv = advection or {}
        d.advection_order = int(adv.get("order", 1))
        if "maxSubsteps" in adv:
            d.max_substeps = int(adv["maxSubsteps"])
        p = pressure or {}
        if "multigrid" in p:
            d.pressure.multigrid = mg_options(**p["multigrid"])
        if "preconditioner" in p:
            d.pressure.preconditioner = PREC[p["preconditioner"]]
        if "tolerance" in p:
            d.pressure.tolerance = float(p["tolerance"])
        if "maxIterations" in p:
            d.pressure.max_iterations = int(p["maxIterations"])
        h = C.c_void_p()
        _check(_lib.flux_smoke3_create(C.byref(d), C.byref(h)), "createGridSmokeSolver3")
        self.h = h.value
        self.resolution = tuple(int(x) for x in resolution)
        self.diagnostics = {}

    def field(self, name):
        _lib.flux_smoke3_field.restype = C.c_void_p
        ptr = _lib.flux_smoke3_field(C.c_void_p(self.h), _FIELDS[name])
        n = self.resolution
        shape = face_shape(n, _FIELDS[name]) if name in ("u", "v", "w") else n
        return Field(F32, shape, handle=ptr, owned=False)

    @property
    def velocity_grid(self):
        return [self.field(c) for c in "uvw"]

    @property
    def density(self):
        """density.current"""
        return self.field("density")

    @property
    def temperature(self):
        return self.field("temperature")

    def source_sphere(self, center, radius, density=1.0, temperature=1.0):
        c = (C.c_float * 3)(*[float(x) for x in center])
        _check(_lib.flux_smoke3_source_sphere(C.c_void_p(self.h), c, C.c_float(radius), C.c_float(density),
                                              C.c_float(temperature)), "source")

    def source_field(self, density=None, temperature=None):
        """state = max(state, source) per cell for the current density and/or
        temperature (float Fields of the grid's cells, None leaves one alone)"""
        _check(_lib.flux_smoke3_source_field(C.c_void_p(self.h), C.c_void_p(density.h if density else None),
                                             C.c_void_p(temperature.h if temperature else None)), "source")

    def on_advance_time_step(self):
        dg = PressureDiag()
        _check(_lib.flux_smoke3_step(C.c_void_p(self.h), C.byref(dg)), "onAdvanceTimeStep")
        self.diagnostics = {"converged": bool(dg.converged), "rejected": bool(dg.rejected),
                            "iterations": dg.iterations, "stoppedBy": STOP_NAMES.get(dg.stopped_by, "unknown"),
                            "residual": dg.residual, "bNorm": dg.b_norm, "badCells": dg.bad_cells}
        return self.diagnostics

    def __del__(self):
        try:
            _lib.flux_smoke3_destroy(C.c_void_p(self.h))
        except Exception:
            pass


def create_grid_smoke_solver3(**kw):
    return GridSmokeSolver3(**kw)
