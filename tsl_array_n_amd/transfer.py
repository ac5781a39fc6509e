"""Particle/grid transfers of the FLIP solver on the GPU
(packages/fluxflow/src/grid/grid_flip_solver3.js, advection_solver3.js).

Particles are ``vec3`` fields (3 x f32 per particle); MAC velocity is a triple
of face fields.  ``Grid`` carries cells, spacing and origin.
"""
import ctypes as C

import numpy as np

from . import Field, Grid3, _check, _lib, array_n, face_shape


class Grid:
    def __init__(self, n, h, origin=(0.0, 0.0, 0.0)):
        self.n = tuple(int(x) for x in n)
        self.h = tuple(float(x) for x in h)
        self.origin = tuple(float(x) for x in origin)
        g = Grid3()
        for a in range(3):
            g.n[a] = self.n[a]
            g.h[a] = self.h[a]
            g.origin[a] = self.origin[a]
        self.c = g

    @property
    def cells(self):
        return int(np.prod(self.n))

    def face_shape(self, axis):
        return face_shape(self.n, axis)


def _mac(fields):
    return (C.c_void_p * 3)(*[f.h for f in fields])


class Bins:
    """Result of bin_particles: per-particle cell, stable order, cell starts."""

    def __init__(self, grid, count):
        self.count = count
        self.cell = array_n("int", max(count, 1))
        self.perm = array_n("int", max(count, 1))
        self.cell_start = array_n("int", grid.cells + 1)
        self.outside = 0


def bin_particles(grid, positions, count, bins=None):
    """cellIndexOf + stable sort by cell (deterministic binning)."""
    bins = bins or Bins(grid, count)
    out = C.c_int64()
    _check(_lib.flux_bin_particles(C.byref(grid.c), C.c_void_p(positions.h), C.c_int64(count), C.c_void_p(bins.cell.h),
                                   C.c_void_p(bins.perm.h), C.c_void_p(bins.cell_start.h), C.byref(out)),
           "bin_particles")
    bins.outside = out.value
    return bins


def p2g_gather(grid, positions, velocities, bins, component, out, valid, eps=1e-4):
    _check(_lib.flux_p2g_gather(C.byref(grid.c), C.c_void_p(positions.h), C.c_void_p(velocities.h),
                                C.c_void_p(bins.perm.h), C.c_void_p(bins.cell_start.h), int(component), C.c_float(eps),
                                C.c_void_p(out.h), C.c_void_p(valid.h)), "p2g")


def p2g_fixed(grid, positions, velocities, count, component, numer, denom, out, valid, scale=65536.0, eps=1e-4):
    _check(_lib.flux_p2g_fixed(C.byref(grid.c), C.c_void_p(positions.h), C.c_void_p(velocities.h), C.c_int64(count),
                               int(component), C.c_float(scale), C.c_float(eps), C.c_void_p(numer.h),
                               C.c_void_p(denom.h), C.c_void_p(out.h), C.c_void_p(valid.h)), "p2g_fixed")


def advect_particles_collider(grid, positions, count, vel, collider_sdf, dt, max_substeps=32):
    """advectParticles with the FLIP solver's static collider (the trace stops
    at the collider; collider_sdf cell-centred on the grid, or None)"""
    _check(_lib.flux_advect_particles_collider(C.byref(grid.c), C.c_void_p(positions.h), C.c_int64(count), _mac(vel),
                                               C.c_void_p(collider_sdf.h if collider_sdf is not None else None),
                                               C.c_float(dt), int(max_substeps)), "advectParticles")


def g2p(grid, positions, velocities, count, vel_new, vel_old, flip_ratio=0.97, damping=0.02):
    _check(_lib.flux_g2p(C.byref(grid.c), C.c_void_p(positions.h), C.c_void_p(velocities.h), C.c_int64(count),
                         _mac(vel_new), _mac(vel_old), C.c_float(flip_ratio), C.c_float(damping)), "g2p")


def advect_particles(grid, positions, count, vel, dt, max_substeps=32):
    _check(_lib.flux_advect_particles(C.byref(grid.c), C.c_void_p(positions.h), C.c_int64(count), _mac(vel),
                                      C.c_float(dt), int(max_substeps)), "advectParticles")


def advect_faces(grid, vel_in, vel_out, dt, max_substeps=32):
    _check(_lib.flux_advect_faces(C.byref(grid.c), _mac(vel_in), _mac(vel_out), C.c_float(dt), int(max_substeps)),
           "advectFaceCentered3")


class SemiLagrangianAdvectionSolver3:
    """createSemiLagrangianAdvectionSolver3({velocityGrid, collider, dt,
    maxSubsteps, order}) (advection_solver3.js:67-362): the velocity fields
    are bound here and read at each call; order 1 or 2 (MacCormack);
    collider: the static collider's SDF field, cell-centred on the velocity
    grid, or None."""

    def __init__(self, grid, velocity_grid, dt, max_substeps=32, order=1, collider=None):
        if order not in (1, 2):
            raise ValueError("createSemiLagrangianAdvectionSolver3: order must be 1 or 2")
        self.grid, self.velocity_grid, self.order, self.collider = grid, list(velocity_grid), order, collider
        h = C.c_void_p()
        _check(_lib.flux_advection3_create(C.byref(grid.c), _mac(self.velocity_grid), C.c_float(dt),
                                           int(max_substeps), int(order), C.byref(h)),
               "createSemiLagrangianAdvectionSolver3")
        self.h = h.value
        if collider is not None:
            _check(_lib.flux_advection3_set_collider(C.c_void_p(self.h), C.c_void_p(collider.h)),
                   "createSemiLagrangianAdvectionSolver3")

    def set_dt(self, dt):
        _check(_lib.flux_advection3_set_dt(C.c_void_p(self.h), C.c_float(dt)), "advection dt")

    def advect_face_centered3(self, input, output):
        _check(_lib.flux_advection3_faces(C.c_void_p(self.h), _mac(input), _mac(output)), "advectFaceCentered3")

    def advect_scalar3(self, input, output):
        _check(_lib.flux_advection3_scalar(C.c_void_p(self.h), C.c_void_p(input.h), C.c_void_p(output.h)),
               "advectScalar3")

    def __del__(self):
        try:
            _lib.flux_advection3_destroy(C.c_void_p(self.h))
        except Exception:
            pass


def create_semi_lagrangian_advection_solver3(grid, velocity_grid, dt, max_substeps=32, order=1):
    return SemiLagrangianAdvectionSolver3(grid, velocity_grid, dt, max_substeps, order)


def two_phase_beta(grid, bins, rho_gas, rho_liquid, fraction, beta):
    _check(_lib.flux_two_phase_beta(C.byref(grid.c), C.c_void_p(bins.cell_start.h if bins is not None else None),
                                    C.c_float(rho_gas), C.c_float(rho_liquid), C.c_void_p(fraction.h), _mac(beta)),
           "two_phase_beta")


def divergence(grid, vel, out):
    _check(_lib.flux_divergence(C.byref(grid.c), _mac(vel), C.c_void_p(out.h)), "divergence")


def mac_fields(grid, arrays=None):
    return [array_n("float", grid.face_shape(a), None if arrays is None else arrays[a]) for a in range(3)]


def main(argv=None):
    """python -m tsl_array_n_amd.transfer --n 256: the two transfers of P2G
    on the dam-break particles (8 per cell, random velocities), each face
    component once warm and then timed: the deterministic gather over
    cell-sorted particles (flux_p2g_gather: binning + stable sort, then one
    thread per face summing its reaching particles in sorted order) and the
    fixed-point scatter (flux_p2g_fixed: int32 atomics, block-aggregated in
    LDS); their velocities agree to the fixed-point resolution."""
    import argparse
    import time
    from . import init, sync
    from .scenes import flip_dam_break
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=128)
    a = ap.parse_args(argv)
    init(0)
    n = a.n
    p = flip_dam_break(n)
    count = p["count"]
    g = Grid((n, n, n), (1.0 / n,) * 3)
    pos = array_n("vec3", count, p["positions"])
    vel = array_n("vec3", count, np.random.default_rng(1).uniform(-1.0, 1.0, (count, 3)).astype(np.float32))
    sync()
    t0 = time.perf_counter()
    bins = bin_particles(g, pos, count)
    sync()
    t_bin = (time.perf_counter() - t0) * 1e3
    print("p2g: %d^3, %d particles, binning + stable sort %.1f ms" % (n, count, t_bin), flush=True)
    worst = 0.0
    for comp in range(3):
        fs = g.face_shape(comp)
        out_g, val_g = array_n("float", fs), array_n("int", fs)
        out_f, val_f = array_n("float", fs), array_n("int", fs)
        num, den = array_n("int", fs), array_n("int", fs)
        times = {}
        for name, run in (("gather", lambda: p2g_gather(g, pos, vel, bins, comp, out_g, val_g)),
                          ("fixed", lambda: p2g_fixed(g, pos, vel, count, comp, num, den, out_f, val_f))):
            run()
            sync()
            t0 = time.perf_counter()
            run()
            sync()
            times[name] = (time.perf_counter() - t0) * 1e3
        vg, vf = out_g.to_array(), out_f.to_array()
        both = (val_g.to_array() > 0) & (val_f.to_array() > 0)
        err = float(np.max(np.abs(vg[both] - vf[both]))) if both.any() else 0.0
        worst = max(worst, err)
        print("p2g component %d: gather %.2f ms, fixed-point scatter %.2f ms, max |difference| %.2e" % (
            comp, times["gather"], times["fixed"], err), flush=True)
    print("p2g: %s" % ("ok" if worst < 1e-3 else "MISMATCH"), flush=True)


if __name__ == "__main__":
    main()
