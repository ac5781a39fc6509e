"""FLIP grid stages around the projection on the GPU: extrapolation
(array_utils3.js), fluid marking and fluid-adjacent faces
(grid_flip_solver3.js), static-collider geometry (sdf_collider3.js,
level_set_utils.js) and constrainVelocity
(grid_blocked_boundary_condition_solver3.js).
"""
import ctypes as C

from . import _check, _lib
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


class BoundarySolver3:
    """createGridBlockedBoundaryConditionSolver3 for a static collider:
    constrain_velocity(extrapolation_depth=5)."""

    def __init__(self, grid, vel, sdf=None, face_weights=None, friction=0.5):
        self._keep = (vel, sdf, face_weights)
        h = C.c_void_p()
        fw = _mac(face_weights) if face_weights is not None else None
        _check(_lib.flux_boundary3_create(C.byref(grid.c), _mac(vel), C.c_void_p(sdf.h) if sdf is not None else None,
                                          fw, C.c_float(friction), C.byref(h)), "boundary")
        self.h = h.value
        self._flags = DIRECTION_ALL_3D

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
