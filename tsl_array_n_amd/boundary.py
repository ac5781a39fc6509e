"""FLIP grid stages around the projection on the GPU: extrapolation
(array_utils3.js), fluid marking and fluid-adjacent faces
(grid_flip_solver3.js), static-collider geometry (sdf_collider3.js,
level_set_utils.js) and constrainVelocity
(grid_blocked_boundary_condition_solver3.js).
"""
import ctypes as C

from . import _check, _lib, array_n
from .transfer import _mac

# closedDomainBoundaryFlag bits (constant.js:24-37)
DIRECTION_LEFT, DIRECTION_RIGHT, DIRECTION_DOWN, DIRECTION_UP, DIRECTION_BACK, DIRECTION_FRONT = 1, 2, 4, 8, 16, 32
DIRECTION_ALL_3D = 63


class Primitive3(C.Structure):
    _fields_ = [("type", C.c_int), ("rotated", C.c_int), ("center", C.c_double * 3), ("radius", C.c_double),
                ("half", C.c_double * 3), ("axis", C.c_double * 3), ("angle", C.c_double)]


class Grid3D(C.Structure):
    _fields_ = [("n", C.c_int * 3), ("h", C.c_double * 3), ("origin", C.c_double * 3)]


def sphere(center, radius):
    """primitive_sdf3.js sphere(center, radius)"""
    return {"type": "sphere", "center": tuple(center), "radius": float(radius)}


def box(center, half_extents, rotation=None):
    """primitive_sdf3.js box(center, halfExtents, rotation={axis, angle})"""
    return {"type": "box", "center": tuple(center), "halfExtents": tuple(half_extents), "rotation": rotation}


def _shapes(shapes):
    arr = (Primitive3 * max(len(shapes), 1))()
    for q, s in enumerate(shapes):
        p = arr[q]
        if s["type"] == "sphere":
            p.type = 0
            p.radius = s["radius"]
        elif s["type"] == "box":
            p.type = 1
            for a in range(3):
                p.half[a] = s["halfExtents"][a]
            rot = s.get("rotation")
            if rot:
                p.rotated = 1
                for a in range(3):
                    p.axis[a] = rot["axis"][a]
                p.angle = rot["angle"]
        else:
            raise ValueError("primitive_sdf3: unknown shape type '%s'." % s["type"])
        for a in range(3):
            p.center[a] = s["center"][a]
    return arr, len(shapes)


def _grid3d(n, h, origin):
    g = Grid3D()
    for a in range(3):
        g.n[a] = int(n[a])
        g.h[a] = float(h[a])
        g.origin[a] = float(origin[a])
    return g


class Extrapolator:
    """createExtrapolateToRegion3(inputField, validField, outputField): run(iterations=5)."""

    def __init__(self, value, valid, out=None):
        out = out if out is not None else value
        self._keep = (value, valid, out)
        h = C.c_void_p()
        _check(_lib.flux_extrapolator3_create(C.c_void_p(value.h), C.c_void_p(valid.h), C.c_void_p(out.h), C.byref(h)),
               "createExtrapolateToRegion3")
        self.h = h.value

    def __call__(self, iterations=5):
        _check(_lib.flux_extrapolator3_run(C.c_void_p(self.h), int(iterations)), "extrapolate")

    run = __call__

    def __del__(self):
        try:
            _lib.flux_extrapolator3_destroy(C.c_void_p(self.h))
        except Exception:
            pass


def mark_fluid_cells(grid, positions, count, fluid_mask):
    _check(_lib.flux_mark_fluid_cells(C.byref(grid.c), C.c_void_p(positions.h), C.c_int64(count),
                                      C.c_void_p(fluid_mask.h)), "markFluidCells")


def fluid_adjacent(grid, fluid_mask, axis, out):
    _check(_lib.flux_fluid_adjacent(C.byref(grid.c), C.c_void_p(fluid_mask.h), int(axis), C.c_void_p(out.h)),
           "fluidAdjacent")


def collider_face_weights(n, h, origin, shapes, weights, invert=False):
    """computeFaceWeights of a static collider into three face fields."""
    arr, cnt = _shapes(shapes)
    g = _grid3d(n, h, origin)
    _check(_lib.flux_collider_face_weights(C.byref(g), arr, cnt, int(bool(invert)), _mac(weights)),
           "computeFaceWeights")


def collider_sdf(n, h, origin, shapes, sdf, invert=False):
    """addShapes: the collider's signed distance at cell centres."""
    arr, cnt = _shapes(shapes)
    g = _grid3d(n, h, origin)
    _check(_lib.flux_collider_sdf(C.byref(g), arr, cnt, int(bool(invert)), C.c_void_p(sdf.h)), "addShapes")


def outflow_dirichlet3(grid, outflows, user_active=None, user_target=None, active=None, target=None):
    """createOutflowPressureDirichlet3(outflows) combined with the caller's
    Dirichlet fields by combineDirichlet3 (sdf_inflow_outflow3.js:44-94), at
    the cell centres of `grid`: returns (active u8, target float) fields for
    the pressure solver's `dirichlet`.  Each outflow is its SDF field,
    cell-centred on the same grid."""
    n = grid.n
    active = active if active is not None else array_n("u8", n)
    target = target if target is not None else array_n("float", n)
    arr = (C.c_void_p * max(1, len(outflows)))(*[o.h for o in outflows])
    _check(_lib.flux_outflow3_dirichlet(C.byref(grid.c), arr, len(outflows),
                                        C.c_void_p(user_active.h) if user_active is not None else None,
                                        C.c_void_p(user_target.h) if user_target is not None else None,
                                        C.c_void_p(active.h), C.c_void_p(target.h)), "createOutflowPressureDirichlet3")
    return active, target


class BoundarySolver3:
    """createGridBlockedBoundaryConditionSolver3 for a static collider:
    constrain_velocity(extrapolation_depth=5)."""

    def __init__(self, grid, vel, sdf=None, face_weights=None, friction=0.5, inflows=()):
        """inflows: (sdf field, (vx, vy, vz), mode 'set' | 'add') triples
        (createSDFInflow3 objects, their SDF on the velocity grid's cells)"""
        self._keep = (vel, sdf, face_weights, tuple(q[0] for q in inflows))
        h = C.c_void_p()
        fw = _mac(face_weights) if face_weights is not None else None
        _check(_lib.flux_boundary3_create(C.byref(grid.c), _mac(vel), C.c_void_p(sdf.h) if sdf is not None else None,
                                          fw, C.c_float(friction), C.byref(h)), "boundary")
        self.h = h.value
        self._flags = DIRECTION_ALL_3D
        for isdf, v, mode in inflows:
            if mode not in ("set", "add"):
                raise ValueError("setInflows: mode must be 'set' or 'add', got %r" % (mode,))
            _check(_lib.flux_boundary3_add_inflow(C.c_void_p(self.h), C.c_void_p(isdf.h), C.c_float(v[0]),
                                                  C.c_float(v[1]), C.c_float(v[2]), 1 if mode == "add" else 0),
                   "setInflows")

    @property
    def closed_domain_boundary_flag(self):
        return self._flags

    @closed_domain_boundary_flag.setter
    def closed_domain_boundary_flag(self, flags):
        _check(_lib.flux_boundary3_set_closed(C.c_void_p(self.h), int(flags)), "boundary")
        self._flags = int(flags)

    def constrain_velocity(self, extrapolation_depth=5):
        _check(_lib.flux_boundary3_constrain(C.c_void_p(self.h), int(extrapolation_depth)), "constrainVelocity")

    def __del__(self):
        try:
            _lib.flux_boundary3_destroy(C.c_void_p(self.h))
        except Exception:
            pass
