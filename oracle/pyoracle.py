"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the CPU-baseline leg of ``bench.py``, as the checker.  The product path
(``tsl_array_n_amd``) never imports this module.

Arrays are numpy, first index fastest: a cell field of shape (nx, ny, nz) is
stored as ``arr.ravel(order="F")``; helpers below take and return flat
contiguous arrays plus the shape.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

STOP_NAMES = {
    0: "none",
    1: "degenerate-pAp",
    2: "pAp-growth",
    3: "alpha-magnitude",
    4: "degenerate-oldRZ",
    5: "converged",
    6: "residual-growing",
}
PREC = {"multigrid": 0, "jacobi": 1, "none": 2}


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


class Grid(C.Structure):
    _fields_ = [("ndim", C.c_int), ("n", C.c_int * 3), ("h", C.c_double * 3)]


class MGOptions(C.Structure):
    _fields_ = [
        ("levels", C.c_int),
        ("nu_down", C.c_int),
        ("nu_up", C.c_int),
        ("n_coarsest", C.c_int),
        ("n_final", C.c_int),
        ("omega", C.c_double),
        ("smoother", C.c_int),
    ]


SMOOTHERS = {"rbgs": 0, "jacobi": 1, "brick": 2}


class PCGState(C.Structure):
    _fields_ = [
        ("iterations", C.c_int),
        ("stopped_by", C.c_int),
        ("converged", C.c_int),
        ("residual_sq", C.c_double),
        ("threshold", C.c_double),
        ("b_norm", C.c_double),
        ("restarts", C.c_int),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
        _lib.orc_mg_create_f64.restype = C.c_void_p
        _lib.orc_mg_create_f32.restype = C.c_void_p
        _lib.orc_true_residual_f64.restype = C.c_double
        _lib.orc_true_residual_f32.restype = C.c_double
        _lib.orc_count_bad_f64.restype = C.c_int64
        _lib.orc_count_bad_f32.restype = C.c_int64
        _lib.orc_flip_box_seed3.restype = C.c_int64
        _lib.orc_rng_next_u32.restype = C.c_uint32
        _lib.orc_sample_trilinear3.restype = C.c_float
        for s in ("f64", "f32"):
            getattr(_lib, "orc_mg_destroy_" + s).argtypes = [C.c_void_p]
            getattr(_lib, "orc_mg_apply_" + s).argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
            getattr(_lib, "orc_mg_level_info_" + s).argtypes = [
                C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    return _lib


def set_threads(n):
    lib().orc_set_threads(int(n))


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _dt(suffix):
    return np.float64 if suffix == "f64" else np.float32


def make_grid(shape, spacing):
    g = Grid()
    g.ndim = len(shape)
    n = list(shape) + [1] * (3 - len(shape))
    h = list(spacing) + [1.0] * (3 - len(spacing))
    for a in range(3):
        g.n[a] = int(n[a])
        g.h[a] = float(h[a])
    return g


def face_shape(shape, axis):
    s = list(shape) + [1] * (3 - len(shape))
    s[axis] += 1
    return tuple(s[: len(shape)])


class _PtrArray:
    """A (const real* const*) built from up to three numpy arrays."""

    def __init__(self, arrays):
        self.arrays = arrays
        self.c = (C.c_void_p * 3)(*[(_p(a).value if a is not None else None) for a in list(arrays) + [None] * (3 - len(arrays))])


def _w(weights):
    if weights is None:
        return None, None
    pa = _PtrArray(weights)
    return pa, C.cast(pa.c, C.c_void_p)


def _u8(mask):
    if mask is None:
        return None
    return np.ascontiguousarray(mask, dtype=np.uint8)


def _real(a, suffix):
    return np.ascontiguousarray(a, dtype=_dt(suffix))


def apply(shape, spacing, x, mask=None, weights=None, suffix="f64"):
    g = make_grid(shape, spacing)
    x = _real(x, suffix)
    out = np.zeros_like(x)
    mask = _u8(mask)
    weights = None if weights is None else [_real(w, suffix) for w in weights]
    keep, wp = _w(weights)
    getattr(lib(), "orc_apply_" + suffix)(C.byref(g), _p(x), _p(out), _p(mask), wp)
    return out


def diag(shape, spacing, mask=None, weights=None, suffix="f64"):
    g = make_grid(shape, spacing)
    out = np.zeros(int(np.prod(shape)), dtype=_dt(suffix))
    mask = _u8(mask)
    weights = None if weights is None else [_real(w, suffix) for w in weights]
    keep, wp = _w(weights)
    getattr(lib(), "orc_diag_" + suffix)(C.byref(g), _p(out), _p(mask), wp)
    return out


def relax(shape, spacing, x, b, color, mask=None, weights=None, omega=1.0, suffix="f64"):
    g = make_grid(shape, spacing)
    x = _real(x, suffix).copy()
    b = _real(b, suffix)
    mask = _u8(mask)
    weights = None if weights is None else [_real(w, suffix) for w in weights]
    keep, wp = _w(weights)
    getattr(lib(), "orc_relax_" + suffix)(C.byref(g), _p(x), _p(b), _p(mask), wp, C.c_double(omega), int(color))
    return x


def residual(shape, spacing, x, b, mask=None, weights=None, suffix="f64"):
    g = make_grid(shape, spacing)
    x = _real(x, suffix)
    b = _real(b, suffix)
    r = np.zeros_like(x)
    mask = _u8(mask)
    weights = None if weights is None else [_real(w, suffix) for w in weights]
    keep, wp = _w(weights)
    getattr(lib(), "orc_residual_" + suffix)(C.byref(g), _p(x), _p(b), _p(r), _p(mask), wp)
    return r


def coarse_shape(shape):
    return tuple(n // 2 for n in shape)


def restrict(shape, r_fine, fine_mask=None, coarse_mask=None, suffix="f64"):
    g = make_grid(shape, [1.0] * len(shape))
    r_fine = _real(r_fine, suffix)
    out = np.zeros(int(np.prod(coarse_shape(shape))), dtype=_dt(suffix))
    getattr(lib(), "orc_restrict_" + suffix)(C.byref(g), _p(r_fine), _p(_u8(fine_mask)), _p(out), _p(_u8(coarse_mask)))
    return out


def correct(shape, x_coarse, x_fine, fine_mask=None, suffix="f64"):
    g = make_grid(shape, [1.0] * len(shape))
    xc = _real(x_coarse, suffix)
    xf = _real(x_fine, suffix).copy()
    getattr(lib(), "orc_correct_" + suffix)(C.byref(g), _p(xc), _p(xf), _p(_u8(fine_mask)))
    return xf


class Multigrid:
    """createMultigridPreconditioner restated (multigrid.js:1213-1735)."""

    def __init__(self, shape, spacing, levels=1, nu_down=2, nu_up=2, n_coarsest=20, n_final=2,
                 omega=None, mask=None, coarse_mask=None, weights=None, suffix="f64", smoother="rbgs"):
        if omega is None:
            omega = 2.0 / 3.0 if smoother == "jacobi" else 1.0
        self.suffix = suffix
        self.shape = tuple(shape)
        self.g = make_grid(shape, spacing)
        o = MGOptions()
        o.levels, o.nu_down, o.nu_up, o.n_coarsest, o.n_final, o.omega = levels, nu_down, nu_up, n_coarsest, n_final, omega
        o.smoother = SMOOTHERS[smoother]
        self.o = o
        self.mask = _u8(mask)
        self.coarse_mask = _u8(coarse_mask)
        self.weights = None if weights is None else [_real(w, suffix) for w in weights]
        self._wkeep, wp = _w(self.weights)
        self.h = getattr(lib(), "orc_mg_create_" + suffix)(C.byref(self.g), C.byref(o), _p(self.mask), _p(self.coarse_mask), wp)
        if not self.h:
            if smoother == "jacobi" and any(v % 2 for v in (nu_down, nu_up, n_coarsest, n_final)):
                raise ValueError("multigrid (jacobi): sweep counts must be even so the ping-pong lands back in x")
            raise ValueError("multigrid: shape is not divisible by 2^(numberOfLevels-1)")

    def apply(self, r):
        r = _real(r, self.suffix)
        z = np.zeros_like(r)
        getattr(lib(), "orc_mg_apply_" + self.suffix)(C.c_void_p(self.h), _p(r), _p(z))
        return z

    def level(self, l):
        n = (C.c_int * 3)()
        mp = C.c_void_p()
        wp = (C.c_void_p * 3)()
        rc = getattr(lib(), "orc_mg_level_info_" + self.suffix)(C.c_void_p(self.h), l, n, C.byref(mp), wp)
        if rc:
            raise IndexError(l)
        dims = [n[0], n[1], n[2]][: len(self.shape)]
        cells = int(np.prod(dims))
        mask = None
        if mp.value:
            mask = np.ctypeslib.as_array(C.cast(mp, C.POINTER(C.c_uint8)), shape=(cells,)).copy()
        ws = []
        for a in range(len(self.shape)):
            if wp[a]:
                ctype = C.c_double if self.suffix == "f64" else C.c_float
                nf = int(np.prod(face_shape(dims, a)))
                ws.append(np.ctypeslib.as_array(C.cast(wp[a], C.POINTER(ctype)), shape=(nf,)).copy())
        return tuple(dims), mask, (ws or None)

    def __del__(self):
        try:
            if self.h:
                getattr(lib(), "orc_mg_destroy_" + self.suffix)(C.c_void_p(self.h))
        except Exception:
            pass


def pcg_solve(shape, spacing, b, x0=None, mask=None, weights=None, precond="multigrid", mg=None,
              tol=1e-6, maxiter=100, relative=True, suffix="f64"):
    g = make_grid(shape, spacing)
    b = _real(b, suffix)
    x = np.zeros_like(b) if x0 is None else _real(x0, suffix).copy()
    mask = _u8(mask)
    weights = None if weights is None else [_real(w, suffix) for w in weights]
    keep, wp = _w(weights)
    st = PCGState()
    getattr(lib(), "orc_pcg_solve_" + suffix)(
        C.byref(g), _p(mask), wp, PREC[precond], C.c_void_p(mg.h if mg is not None else None),
        _p(b), _p(x), C.c_double(tol), int(maxiter), int(bool(relative)), C.byref(st))
    return x, {
        "iterations": st.iterations,
        "stoppedBy": STOP_NAMES[st.stopped_by],
        "converged": bool(st.converged),
        "residual": float(np.sqrt(abs(st.residual_sq))),
        "threshold": st.threshold,
        "bNorm": st.b_norm,
        "restarts": st.restarts,
    }


def true_residual(shape, spacing, b, x, mask=None, weights=None, suffix="f64"):
    g = make_grid(shape, spacing)
    b = _real(b, suffix)
    x = _real(x, suffix)
    weights = None if weights is None else [_real(w, suffix) for w in weights]
    keep, wp = _w(weights)
    return getattr(lib(), "orc_true_residual_" + suffix)(C.byref(g), _p(_u8(mask)), wp, _p(b), _p(x))


def pressure3_assemble(shape, spacing, vel, active=None, target=None, beta=None, cw=None,
                       pressure=None, suffix="f64"):
    """updateDirichletFields + combineAxisWeight + dispatchBuildSystem."""
    g = make_grid(shape, spacing)
    n = int(np.prod(shape))
    dt = _dt(suffix)
    vel = [_real(v, suffix) for v in vel]
    active = _u8(active)
    target = None if target is None else _real(target, suffix)
    beta = None if beta is None else [_real(v, suffix) for v in beta]
    cw = None if cw is None else [_real(v, suffix) for v in cw]
    mask_out = np.zeros(n, np.uint8)
    vent_out = np.zeros(n, np.uint8)
    target_out = np.zeros(n, dt)
    wsys = [np.zeros(int(np.prod(face_shape(shape, a))), dt) for a in range(3)]
    b = np.zeros(n, dt)
    p = np.zeros(n, dt) if pressure is None else _real(pressure, suffix).copy()
    kv, vp = _w(vel)
    kb, bp = _w(beta)
    kc, cp = _w(cw)
    kw, wp = _w(wsys)
    getattr(lib(), "orc_pressure3_assemble_" + suffix)(
        C.byref(g), _p(active), _p(target), bp, cp, vp, _p(mask_out), _p(vent_out), _p(target_out), wp, _p(b), _p(p))
    has_mask = active is not None or cw is not None
    has_w = beta is not None or cw is not None
    return {
        "mask": mask_out if has_mask else None,
        "vent": vent_out if has_mask else None,
        "target": target_out,
        "weights": wsys if has_w else None,
        "b": b,
        "pressure": p,
    }


def pressure3_correct(shape, spacing, pressure, vel_in, beta=None, cw=None, vel_out=None, suffix="f64"):
    g = make_grid(shape, spacing)
    pressure = _real(pressure, suffix)
    vel_in = [_real(v, suffix) for v in vel_in]
    out = [v.copy() for v in (vel_in if vel_out is None else vel_out)]
    out = [_real(v, suffix) for v in out]
    kb, bp = _w(None if beta is None else [_real(v, suffix) for v in beta])
    kc, cp = _w(None if cw is None else [_real(v, suffix) for v in cw])
    ki, ip = _w(vel_in)
    ko, op = _w(out)
    getattr(lib(), "orc_pressure3_correct_" + suffix)(C.byref(g), _p(pressure), bp, cp, ip, op)
    return out


def count_bad(p, max_plausible=1e6, suffix="f64"):
    p = _real(p, suffix)
    return int(getattr(lib(), "orc_count_bad_" + suffix)(_p(p), C.c_int64(p.size), C.c_double(max_plausible)))


def divergence3(shape, spacing, vel, suffix="f64"):
    g = make_grid(shape, spacing)
    vel = [_real(v, suffix) for v in vel]
    out = np.zeros(int(np.prod(shape)), _dt(suffix))
    kv, vp = _w(vel)
    getattr(lib(), "orc_divergence3_" + suffix)(C.byref(g), vp, _p(out))
    return out


# ---------------- transfers (f32) ----------------

def rng(seed, n):
    out = np.zeros(n, np.float64)
    lib().orc_rng_fill(C.c_uint32(seed), _p(out), C.c_int64(n))
    return out


def rng_u32(seed, n):
    class R(C.Structure):
        _fields_ = [("state", C.c_uint32)]
    r = R()
    lib().orc_rng_seed(C.byref(r), C.c_uint32(seed))
    return np.array([lib().orc_rng_next_u32(C.byref(r)) for _ in range(n)], np.uint32)


def flip_box_seed3(box_min, box_max, h, ppa=2, jitter=0.2, seed=0x9E3779B9):
    bmin = (C.c_double * 3)(*box_min)
    bmax = (C.c_double * 3)(*box_max)
    hh = (C.c_double * 3)(*h)
    count = lib().orc_flip_box_seed3(bmin, bmax, hh, int(ppa), C.c_double(jitter), C.c_uint32(seed), None)
    out = np.zeros(count * 3, np.float32)
    lib().orc_flip_box_seed3(bmin, bmax, hh, int(ppa), C.c_double(jitter), C.c_uint32(seed), _p(out))
    return out


def _f3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


def _i3(v):
    return (C.c_int * 3)(*[int(x) for x in v])


def cell_index3(pos, origin, h, n):
    pos = np.ascontiguousarray(pos, np.float32)
    count = pos.size // 3
    out = np.zeros(count, np.int32)
    lib().orc_cell_index3(_p(pos), C.c_int64(count), _f3(origin), _f3(h), _i3(n), _p(out))
    return out


def sort_by_cell(cell, n_cells):
    cell = np.ascontiguousarray(cell, np.int32)
    perm = np.zeros(cell.size, np.int32)
    start = np.zeros(n_cells + 1, np.int32)
    lib().orc_sort_by_cell(_p(cell), C.c_int64(cell.size), C.c_int64(n_cells), _p(perm), _p(start))
    return perm, start


def mac_origin(origin, h, comp):
    o = [np.float32(x) for x in origin]
    hh = [np.float32(x) for x in h]
    for a in range(3):
        if a != comp:
            o[a] = np.float32(o[a] + np.float32(hh[a] * np.float32(0.5)))
    return o


def p2g_fixed3(pos, vel, n, h, origin, comp, scale=65536.0, eps=1e-4, mass=None):
    """mass: None, or (concentration, d0, d1) for the mass-weighted scatter"""
    pos = np.ascontiguousarray(pos, np.float32)
    vel = np.ascontiguousarray(vel, np.float32)
    fn = list(n)
    fn[comp] += 1
    nf = int(np.prod(fn))
    numer = np.zeros(nf, np.int32)
    denom = np.zeros(nf, np.int32)
    if mass is None:
        lib().orc_p2g_fixed3(_p(pos), _p(vel), C.c_int64(pos.size // 3), comp, _i3(fn),
                             _f3(mac_origin(origin, h, comp)), _f3(h), C.c_float(scale), _p(numer), _p(denom))
    else:
        conc = np.ascontiguousarray(mass[0], np.float32)
        lib().orc_p2g_fixed_mass3(_p(pos), _p(vel), _p(conc), C.c_int64(pos.size // 3), comp, _i3(fn),
                                  _f3(mac_origin(origin, h, comp)), _f3(h), C.c_float(scale), C.c_float(mass[1]),
                                  C.c_float(mass[2]), _p(numer), _p(denom))
    out = np.zeros(nf, np.float32)
    valid = np.zeros(nf, np.int32)
    lib().orc_p2g_finalize_fixed(_p(numer), _p(denom), C.c_int64(nf), C.c_float(scale), C.c_float(eps), _p(out), _p(valid))
    return out, valid, numer, denom


def p2g_gather3(pos, vel, perm, start, n, h, origin, comp, eps=1e-4):
    pos = np.ascontiguousarray(pos, np.float32)
    vel = np.ascontiguousarray(vel, np.float32)
    fn = list(n)
    fn[comp] += 1
    nf = int(np.prod(fn))
    out = np.zeros(nf, np.float32)
    valid = np.zeros(nf, np.int32)
    lib().orc_p2g_gather3(_p(pos), _p(vel), _p(np.ascontiguousarray(perm, np.int32)), _p(np.ascontiguousarray(start, np.int32)),
                          comp, _i3(n), _f3(mac_origin(origin, h, comp)), _f3(h), C.c_float(eps), _p(out), _p(valid))
    return out, valid


def p2g_float3(pos, vel, n, h, origin, comp, eps=1e-4):
    pos = np.ascontiguousarray(pos, np.float32)
    vel = np.ascontiguousarray(vel, np.float32)
    fn = list(n)
    fn[comp] += 1
    nf = int(np.prod(fn))
    numer = np.zeros(nf, np.float64)
    denom = np.zeros(nf, np.float64)
    lib().orc_p2g_float3(_p(pos), _p(vel), C.c_int64(pos.size // 3), comp, _i3(fn), _f3(mac_origin(origin, h, comp)),
                         _f3(h), _p(numer), _p(denom))
    out = np.zeros(nf, np.float32)
    valid = np.zeros(nf, np.int32)
    lib().orc_p2g_finalize_float(_p(numer), _p(denom), C.c_int64(nf), C.c_float(eps), _p(out), _p(valid))
    return out, valid


def g2p3(pos, vel, u, v, w, uo, vo, wo, n, h, origin, flip_ratio=0.95, damping=0.0):
    pos = np.ascontiguousarray(pos, np.float32)
    vel = np.ascontiguousarray(vel, np.float32).copy()
    arrs = [np.ascontiguousarray(a, np.float32) for a in (u, v, w, uo, vo, wo)]
    lib().orc_g2p3(_p(pos), _p(vel), C.c_int64(pos.size // 3), *[_p(a) for a in arrs], _i3(n), _f3(h), _f3(origin),
                   C.c_float(flip_ratio), C.c_float(damping))
    return vel


def advect_particles3(pos, u, v, w, n, h, origin, dt, max_substeps=32):
    pos = np.ascontiguousarray(pos, np.float32).copy()
    arrs = [np.ascontiguousarray(a, np.float32) for a in (u, v, w)]
    lib().orc_advect_particles3(_p(pos), C.c_int64(pos.size // 3), *[_p(a) for a in arrs], _i3(n), _f3(h), _f3(origin),
                                C.c_float(dt), int(max_substeps))
    return pos


def advect_face3(u, v, w, n, h, origin, dt, max_substeps=32, order=1):
    arrs = [np.ascontiguousarray(a, np.float32) for a in (u, v, w)]
    outs = [np.zeros_like(a) for a in arrs]
    fn = lib().orc_advect_face3 if order == 1 else lib().orc_advect_face3_mc
    fn(*[_p(a) for a in arrs], *[_p(a) for a in outs], _i3(n), _f3(h), _f3(origin), C.c_float(dt), int(max_substeps))
    return outs


def advect_faces3_vd(vel, data, n, h, origin, dt, max_substeps=32, order=1):
    """advectFaceCentered3 of the face grid `data` through the velocity grid
    `vel` (advection_solver3.js:130-282), order 1 or 2."""
    keep = [np.ascontiguousarray(a, np.float32) for a in vel] + [np.ascontiguousarray(a, np.float32) for a in data]
    V = (C.c_void_p * 3)(*[_p(a) for a in keep[:3]])
    D = (C.c_void_p * 3)(*[_p(a) for a in keep[3:]])
    outs = [np.zeros_like(a) for a in keep[3:]]
    Oa = (C.c_void_p * 3)(*[_p(a) for a in outs])
    fn = lib().orc_advect_faces3_vd if order == 1 else lib().orc_advect_faces3_mc_vd
    fn(V, D, Oa, _i3(n), _f3(h), _f3(origin), C.c_float(dt), int(max_substeps))
    return outs


def trace3(pos, u, v, w, n, h, origin, dt, direction=1.0, max_substeps=32):
    pos = np.ascontiguousarray(pos, np.float32).copy()
    arrs = [np.ascontiguousarray(a, np.float32) for a in (u, v, w)]
    lib().orc_trace3(_p(pos), C.c_int64(pos.size // 3), *[_p(a) for a in arrs], _i3(n), _f3(h), _f3(origin),
                     C.c_float(dt), int(max_substeps), C.c_float(direction))
    return pos


def advect_faces3_c(vel, data, n, h, origin, csdf, dt, max_substeps=32, order=1, out=None):
    """advectFaceCentered3 through a static collider (csdf: its SDF
    cell-centred on the grid, or None), advection_solver3.js:20-282; faces
    inside the collider keep `out`'s values (zeros when out is None)"""
    keep = [np.ascontiguousarray(a, np.float32) for a in vel] + [np.ascontiguousarray(a, np.float32) for a in data]
    csdf = None if csdf is None else np.ascontiguousarray(csdf, np.float32)
    V = (C.c_void_p * 3)(*[_p(a) for a in keep[:3]])
    D = (C.c_void_p * 3)(*[_p(a) for a in keep[3:]])
    outs = [np.zeros_like(a) for a in keep[3:]] if out is None else \
        [np.ascontiguousarray(a, np.float32).copy() for a in out]
    Oa = (C.c_void_p * 3)(*[_p(a) for a in outs])
    lib().orc_advect_faces3_c(V, D, Oa, _i3(n), _f3(h), _f3(origin), _p(csdf), C.c_float(dt), int(max_substeps),
                              int(order))
    return outs


def advect_scalar3_c(src, u, v, w, n, h, origin, csdf, dt, max_substeps=32, order=1, out=None):
    """advectScalar3 through a static collider (advection_solver3.js:285-357);
    cells inside keep `out`'s values"""
    src = np.ascontiguousarray(src, np.float32)
    dst = np.zeros_like(src) if out is None else np.ascontiguousarray(out, np.float32).copy()
    csdf = None if csdf is None else np.ascontiguousarray(csdf, np.float32)
    arrs = [np.ascontiguousarray(a, np.float32) for a in (u, v, w)]
    lib().orc_advect_scalar3_c(_p(src), _p(dst), *[_p(a) for a in arrs], _i3(n), _f3(h), _f3(origin), _p(csdf),
                               C.c_float(dt), int(max_substeps), int(order))
    return dst


def push_out3(pos, sdf, n, h, origin, thresh=0.0, shift=0.0):
    pos = np.ascontiguousarray(pos, np.float32).copy()
    sdf = np.ascontiguousarray(sdf, np.float32)
    lib().orc_push_out3(_p(pos), C.c_int64(pos.size // 3), _p(sdf), _i3(n), _f3(h), _f3(origin), C.c_float(thresh),
                        C.c_float(shift))
    return pos


def advect_particles3_c(pos, u, v, w, n, h, origin, csdf, dt, max_substeps=32):
    pos = np.ascontiguousarray(pos, np.float32).copy()
    csdf = None if csdf is None else np.ascontiguousarray(csdf, np.float32)
    arrs = [np.ascontiguousarray(a, np.float32) for a in (u, v, w)]
    lib().orc_advect_particles3_c(_p(pos), C.c_int64(pos.size // 3), *[_p(a) for a in arrs], _i3(n), _f3(h),
                                  _f3(origin), _p(csdf), C.c_float(dt), int(max_substeps))
    return pos


def trace3_c(pos, u, v, w, n, h, origin, csdf, dt, direction=1.0, max_substeps=32):
    pos = np.ascontiguousarray(pos, np.float32).copy()
    csdf = None if csdf is None else np.ascontiguousarray(csdf, np.float32)
    arrs = [np.ascontiguousarray(a, np.float32) for a in (u, v, w)]
    lib().orc_trace3_c(_p(pos), C.c_int64(pos.size // 3), *[_p(a) for a in arrs], _i3(n), _f3(h), _f3(origin),
                       _p(csdf), C.c_float(dt), int(max_substeps), C.c_float(direction))
    return pos


def two_phase_beta3(frac, n, rho_gas, rho_liquid):
    frac = np.ascontiguousarray(frac, np.float32)
    outs = [np.zeros(int(np.prod(face_shape(n, a))), np.float32) for a in range(3)]
    lib().orc_two_phase_beta3(_p(frac), _i3(n), C.c_float(rho_gas), C.c_float(rho_liquid), *[_p(o) for o in outs])
    return outs


def set_post_forward(on):
    """post-smoothing in the pre-smoothing colour order (a non-symmetric
    V-cycle), for the stop-code tests"""
    lib().orc_set_post_forward(int(bool(on)))


def set_flexible(on):
    """the flexible (Polak-Ribiere) beta in pcg_solve_mixed"""
    lib().orc_set_flexible(int(bool(on)))


def pcg_solve_mixed(shape, spacing, b, x0=None, mask=None, weights=None, precond="multigrid", mg=None,
                    tol=1e-6, maxiter=100, relative=True):
    """The HIP plain path's arithmetic: f32 vectors, f64 accumulator and residual."""
    g = make_grid(shape, spacing)
    b = _real(b, "f32")
    x = np.zeros_like(b) if x0 is None else _real(x0, "f32").copy()
    x64 = np.zeros(x.size, np.float64)
    mask = _u8(mask)
    weights = None if weights is None else [_real(w, "f32") for w in weights]
    keep, wp = _w(weights)
    st = PCGState()
    if mg is not None and mg.suffix != "f32":
        raise ValueError("pcg_solve_mixed needs an f32 multigrid")
    lib().orc_pcg_solve_mixed(C.byref(g), _p(mask), wp, PREC[precond], C.c_void_p(mg.h if mg is not None else None),
                              _p(b), _p(x), _p(x64), C.c_double(tol), int(maxiter), int(bool(relative)), C.byref(st))
    return x, x64, {
        "iterations": st.iterations,
        "stoppedBy": STOP_NAMES[st.stopped_by],
        "converged": bool(st.converged),
        "residual": float(np.sqrt(abs(st.residual_sq))),
        "threshold": st.threshold,
        "bNorm": st.b_norm,
        "restarts": st.restarts,
    }


# ---------------------------------------------------------------- FLIP grid stages
class Primitive3(C.Structure):
    _fields_ = [("type", C.c_int), ("rotated", C.c_int), ("center", C.c_double * 3), ("radius", C.c_double),
                ("half", C.c_double * 3), ("axis", C.c_double * 3), ("angle", C.c_double)]


def _prims(shapes):
    arr = (Primitive3 * max(len(shapes), 1))()
    for q, s in enumerate(shapes):
        p = arr[q]
        p.type = 0 if s["type"] == "sphere" else 1
        for a in range(3):
            p.center[a] = s["center"][a]
        if p.type == 0:
            p.radius = s["radius"]
        else:
            for a in range(3):
                p.half[a] = s["halfExtents"][a]
            if s.get("rotation"):
                p.rotated = 1
                for a in range(3):
                    p.axis[a] = s["rotation"]["axis"][a]
                p.angle = s["rotation"]["angle"]
    return arr, len(shapes)


def _d3(v):
    return (C.c_double * 3)(*[float(x) for x in v])


def extrapolate3(n, value, valid, iterations=5):
    value = np.ascontiguousarray(value, np.float32)
    valid = np.ascontiguousarray(valid, np.int32)
    out = np.empty_like(value)
    lib().orc_extrapolate3(_i3(n), _p(value), _p(valid), _p(out), int(iterations))
    return out


def mark_fluid3(pos, n, h, origin):
    pos = np.ascontiguousarray(pos, np.float32)
    mask = np.zeros(int(np.prod(n)), np.float32)
    lib().orc_mark_fluid3(_p(pos), C.c_int64(pos.size // 3), _f3(origin), _f3(h), _i3(n), _p(mask))
    return mask


def fluid_adjacent3(mask, n, axis):
    mask = np.ascontiguousarray(mask, np.float32)
    out = np.zeros(int(np.prod(face_shape(n, axis))), np.int32)
    lib().orc_fluid_adjacent3(_p(mask), _i3(n), int(axis), _p(out))
    return out


def primitives_distance3(shapes, x, y, z):
    arr, cnt = _prims(shapes)
    lib().orc_primitives_distance3.restype = C.c_double
    return lib().orc_primitives_distance3(arr, cnt, C.c_double(x), C.c_double(y), C.c_double(z))


def fraction_inside(bl, br, tl, tr):
    lib().orc_fraction_inside.restype = C.c_double
    return lib().orc_fraction_inside(C.c_double(bl), C.c_double(br), C.c_double(tl), C.c_double(tr))


def face_weights3(shapes, n, h, origin, invert=False):
    arr, cnt = _prims(shapes)
    outs = [np.zeros(int(np.prod(face_shape(n, a))), np.float32) for a in range(3)]
    lib().orc_face_weights3(arr, cnt, int(bool(invert)), _i3(n), _d3(h), _d3(origin), *[_p(o) for o in outs])
    return outs


def collider_sdf3(shapes, n, h, origin, invert=False):
    arr, cnt = _prims(shapes)
    out = np.zeros(int(np.prod(n)), np.float32)
    lib().orc_collider_sdf3(arr, cnt, int(bool(invert)), _i3(n), _d3(h), _d3(origin), _p(out))
    return out


def constrain_velocity3(vel, n, h, origin, closed_flags=63, sdf=None, weights=None, friction=0.5, depth=5,
                        inflows=()):
    """constrainVelocity (grid_blocked_boundary_condition_solver3.js:389-477);
    inflows: (sdf, (vx, vy, vz), mode) triples, mode 'set' or 'add', each SDF
    cell-centred on the velocity grid"""
    u, v, w = [np.ascontiguousarray(a, np.float32).copy() for a in vel]
    sdf = None if sdf is None else np.ascontiguousarray(sdf, np.float32)
    ws = [None, None, None] if weights is None else [np.ascontiguousarray(a, np.float32) for a in weights]
    isdf = [np.ascontiguousarray(q[0], np.float32) for q in inflows]
    ptrs = (C.c_void_p * max(1, len(isdf)))(*[a.ctypes.data for a in isdf])
    ivel = np.array([list(q[1]) for q in inflows] or [[0, 0, 0]], np.float32).ravel()
    imode = np.array([1 if q[2] == "add" else 0 for q in inflows] or [0], np.int32)
    lib().orc_constrain_velocity3_inflows(_p(u), _p(v), _p(w), _i3(n), _f3(h), _f3(origin), int(closed_flags), _p(sdf),
                                          _p(ws[0]), _p(ws[1]), _p(ws[2]), C.c_float(friction), int(depth),
                                          len(isdf), ptrs, _p(ivel), _p(imode))
    return [u, v, w]


def resample3(pos, vel, u, v, w, n, h, origin, min_pc=4, max_pc=8):
    pos = np.ascontiguousarray(pos, np.float32).copy()
    vel = np.ascontiguousarray(vel, np.float32).copy()
    arrs = [np.ascontiguousarray(a, np.float32) for a in (u, v, w)]
    lib().orc_resample3(_p(pos), _p(vel), C.c_int64(pos.size // 3), *[_p(a) for a in arrs], _i3(n), _f3(h),
                        _f3(origin), int(min_pc), int(max_pc))
    return pos, vel


# ------------------------------------------------ concentration and density
def conc_accumulate3(pos, conc, n, h, origin, scale=4096.0):
    pos = np.ascontiguousarray(pos, np.float32)
    conc = np.ascontiguousarray(conc, np.float32)
    cells = int(np.prod(n))
    accum, cnt = np.zeros(cells, np.int32), np.zeros(cells, np.int32)
    lib().orc_conc_accumulate3(_p(pos), _p(conc), C.c_int64(conc.size), _i3(n), _f3(h), _f3(origin),
                               C.c_float(scale), _p(accum), _p(cnt))
    return accum, cnt


def conc_resolve3(accum, cnt, scale=4096.0, density=None):
    """density: None or (d0, d1); returns (cell concentration, cell density or None)"""
    accum = np.ascontiguousarray(accum, np.int32)
    cnt = np.ascontiguousarray(cnt, np.int32)
    cc = np.zeros(accum.size, np.float32)
    rho = np.zeros(accum.size, np.float32)
    d0, d1 = density if density is not None else (0.0, 0.0)
    lib().orc_conc_resolve3(_p(accum), _p(cnt), C.c_int64(accum.size), C.c_float(scale), int(density is not None),
                            C.c_float(d0), C.c_float(d1), _p(cc), _p(rho))
    return cc, (rho if density is not None else None)


def density_beta3(density, fluid_mask, n, d0, d1):
    density = np.ascontiguousarray(density, np.float32)
    fluid_mask = np.ascontiguousarray(fluid_mask, np.float32)
    outs = [np.zeros(int(np.prod(face_shape(n, a))), np.float32) for a in range(3)]
    lib().orc_density_beta3(_p(density), _p(fluid_mask), _i3(n), C.c_float(d0), C.c_float(d1), *[_p(o) for o in outs])
    return outs


def reduced_gravity3(vel, beta, density, n, h, origin, gravity, dt, d0, d1):
    u, v, w = [np.ascontiguousarray(x, np.float32).copy() for x in vel]
    bk = [np.ascontiguousarray(b, np.float32) for b in beta]
    B = (C.c_void_p * 3)(*[_p(b) for b in bk])
    density = np.ascontiguousarray(density, np.float32)
    lib().orc_reduced_gravity3(_p(u), _p(v), _p(w), B, _p(density), _i3(n), _f3(h), _f3(origin), _f3(gravity),
                               C.c_float(dt), C.c_float(d0), C.c_float(d1))
    return [u, v, w]


def mix_fade3(pos, conc, cell_conc, n, h, origin, mixing, fade):
    pos = np.ascontiguousarray(pos, np.float32)
    conc = np.ascontiguousarray(conc, np.float32).copy()
    cell_conc = np.ascontiguousarray(cell_conc, np.float32)
    lib().orc_mix_fade3(_p(pos), _p(conc), C.c_int64(conc.size), _p(cell_conc), _i3(n), _f3(h), _f3(origin),
                        C.c_float(mixing), C.c_float(fade))
    return conc


def resample_phase3(pos, vel, conc, u, v, w, n, h, origin, min_pc, max_pc, cnt, accum, scale=4096.0):
    pos = np.ascontiguousarray(pos, np.float32).copy()
    vel = np.ascontiguousarray(vel, np.float32).copy()
    conc = np.ascontiguousarray(conc, np.float32)
    arrs = [np.ascontiguousarray(a, np.float32) for a in (u, v, w)]
    cnt = np.ascontiguousarray(cnt, np.int32)
    accum = np.ascontiguousarray(accum, np.int32)
    lib().orc_resample_phase3(_p(pos), _p(vel), _p(conc), C.c_int64(conc.size), *[_p(a) for a in arrs], _i3(n),
                              _f3(h), _f3(origin), int(min_pc), int(max_pc), _p(cnt), _p(accum), C.c_float(scale))
    return pos, vel


def free_surface_fields(mask, n, h, origin, gravity, dt, reduced=True):
    """dirichlet(pos) of grid_flip_solver3.js:252-259 at cell centres (f32)."""
    f = np.float32
    k, j, i = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    px = (f(origin[0]) + f(h[0]) * f(0.5)) + i.astype(f) * f(h[0])
    py = (f(origin[1]) + f(h[1]) * f(0.5)) + j.astype(f) * f(h[1])
    pz = (f(origin[2]) + f(h[2]) * f(0.5)) + k.astype(f) * f(h[2])
    g = [f(x) for x in gravity]
    dot = ((g[0] * px + g[1] * py) + g[2] * pz).astype(f)
    target = (-(dot * f(dt))).astype(f).ravel() if reduced else np.zeros(int(np.prod(n)), f)
    active = (np.asarray(mask, f) < f(0.5)).astype(np.uint8)
    return active, target


def pin_cell3(n, pin):
    """pressurePin: 'topCenter' (floor(nx/2), ny-1, floor(nz/2)), an (i, j, k) cell, or None"""
    if pin is None:
        return None
    if pin == "topCenter":
        return (n[0] // 2, n[1] - 1, n[2] // 2)
    return tuple(int(x) for x in pin)


def flip3_step(st, n, h, origin, dt, gravity=(0.0, -9.81, 0.0), levels=1, tol=1e-6, flip_ratio=0.97,
               damping=0.02, scale=65536.0, eps=1e-4, flags=63, resample=True, min_pc=4, max_pc=8, depth=5,
               max_substeps=32, reduced=True, max_plausible=None, carry_concentration=False, mixing=0.0, fade=0.0,
               conc_scale=4096.0, density=None, mass_weighted=False, pressure_pin=None, phase_resample=False,
               collider_sdf=None, collider_weights=None, friction=0.5, push_thresh=0.0, push_shift=0.0):
    """One createGridFlipSolver3 time step (grid_flip_solver3.js:1083-1132)
    composed from the oracle's stages; st holds pos, vel, u, v, w, pressure
    and the last accepted pressure (and conc with carry_concentration), and is
    updated in place.  density = (ambient, component) couples the
    concentration to the pressure's face weights (265-357) and, with reduced
    pressure, to the gravity term (848-880); mass_weighted weights the
    scatter by rho / rho_ref (493-531); pressure_pin pins one cell to 0 when
    the free surface leaves no Dirichlet cell; phase_resample splits the
    resampling's donors by phase.  collider_sdf (cell-centred on the grid)
    is the static collider: the particle trace stops at it, particles inside
    are pushed out (push_thresh, push_shift; 945-964) and constrainVelocity
    projects against it (collider_weights: its face weights, friction)."""
    pos, vel = st["pos"], st["vel"]
    u, v, w = st["u"], st["v"], st["w"]
    conc = st.get("conc")
    pos = advect_particles3_c(pos, u, v, w, n, h, origin, collider_sdf, dt, max_substeps)
    if collider_sdf is not None:
        pos = push_out3(pos, collider_sdf, n, h, origin, push_thresh, push_shift)
    cons = dict(sdf=collider_sdf, weights=collider_weights, friction=friction)
    mass = (conc, density[0], density[1]) if (mass_weighted and density is not None and carry_concentration) else None
    vel_grid, valid = [], []
    for a in range(3):
        o, vd, _, _ = p2g_fixed3(pos, vel, n, h, origin, a, scale, eps, mass=mass)
        vel_grid.append(o)
        valid.append(vd)
    old = [x.copy() for x in vel_grid]
    vel_grid = [extrapolate3(face_shape(n, a), vel_grid[a], valid[a], depth) for a in range(3)]
    vel_grid = constrain_velocity3(vel_grid, n, h, origin, closed_flags=flags, depth=depth, **cons)
    mask = mark_fluid3(pos, n, h, origin)
    if resample:
        if phase_resample and carry_concentration:
            accum, cnt = conc_accumulate3(pos, conc, n, h, origin, conc_scale)
            pos, vel = resample_phase3(pos, vel, conc, *vel_grid, n, h, origin, min_pc, max_pc, cnt, accum, conc_scale)
        else:
            pos, vel = resample3(pos, vel, *vel_grid, n, h, origin, min_pc, max_pc)
        mask = mark_fluid3(pos, n, h, origin)
    adj = [fluid_adjacent3(mask, n, a) for a in range(3)]
    beta, cell_conc, cell_density = None, None, None
    if carry_concentration:
        accum, cnt = conc_accumulate3(pos, conc, n, h, origin, conc_scale)
        cell_conc, cell_density = conc_resolve3(accum, cnt, conc_scale, density)
        if density is not None:
            beta = density_beta3(cell_density, mask, n, density[0], density[1])
    if not reduced:
        vel_grid = [(vel_grid[a] + np.float32(gravity[a]) * np.float32(dt)).astype(np.float32) for a in range(3)]
    elif density is not None:
        vel_grid = reduced_gravity3(vel_grid, beta, cell_density, n, h, origin, gravity, dt, density[0], density[1])
    vel_grid = constrain_velocity3(vel_grid, n, h, origin, closed_flags=flags, depth=depth, **cons)
    active, target = free_surface_fields(mask, n, h, origin, gravity, dt, reduced)
    pin = pin_cell3(n, pressure_pin)
    if pin is not None and not active.any():
        c = pin[0] + n[0] * (pin[1] + n[1] * pin[2])
        active[c] = 1
        target[c] = 0.0
    if max_plausible is None:
        extent = sum(n[a] * h[a] for a in range(3))
        gm = float(np.sqrt(sum(x * x for x in gravity)))
        max_plausible = 8.0 * dt * gm * extent if (reduced and gm > 0) else 1e6
    A = pressure3_assemble(n, h, vel_grid, active=active, target=target, beta=beta, pressure=st["pressure"],
                           suffix="f32")
    mg = Multigrid(n, h, levels=levels, mask=A["mask"], coarse_mask=A["vent"], weights=A["weights"], suffix="f32")
    x, _, diag = pcg_solve_mixed(n, h, A["b"], x0=A["pressure"], mask=A["mask"], weights=A["weights"], mg=mg, tol=tol,
                                 maxiter=int(np.ceil(1.5 * max(n))))
    bad = count_bad(x, max_plausible, suffix="f32")
    if bad > 0:
        x = st["p_accepted"].copy()
    else:
        st["p_accepted"] = x.copy()
    diag["rejected"] = bad > 0
    vel_grid = pressure3_correct(n, h, x, vel_grid, beta=beta, suffix="f32")
    vel_grid = constrain_velocity3(vel_grid, n, h, origin, closed_flags=flags, depth=depth, **cons)
    vel_grid = [extrapolate3(face_shape(n, a), vel_grid[a], adj[a], depth) for a in range(3)]
    vel = g2p3(pos, vel, *vel_grid, *old, n, h, origin, flip_ratio, damping)
    if carry_concentration:
        conc = mix_fade3(pos, conc, cell_conc, n, h, origin, mixing, fade)
        st.update(conc=conc, cell_conc=cell_conc, cell_density=cell_density)
    st.update(pos=pos, vel=vel, u=vel_grid[0], v=vel_grid[1], w=vel_grid[2], pressure=x, mask=mask)
    return diag


# ---------------------------------------------------------------- smoke
def advect_scalar3(src, u, v, w, n, h, origin, dt, max_substeps=32, order=1):
    src = np.ascontiguousarray(src, np.float32)
    out = np.empty_like(src)
    arrs = [np.ascontiguousarray(a, np.float32) for a in (u, v, w)]
    fn = lib().orc_advect_scalar3 if order == 1 else lib().orc_advect_scalar3_mc
    fn(_p(src), _p(out), *[_p(a) for a in arrs], _i3(n), _f3(h), _f3(origin), C.c_float(dt), int(max_substeps))
    return out


def buoyancy3(vel, dens, temp, n, h, origin, a, b, tamb, up, dt):
    u, v, w = [np.ascontiguousarray(x, np.float32).copy() for x in vel]
    dens = np.ascontiguousarray(dens, np.float32)
    temp = np.ascontiguousarray(temp, np.float32)
    lib().orc_buoyancy3(_p(u), _p(v), _p(w), _p(dens), _p(temp), _i3(n), _f3(h), _f3(origin), C.c_float(a),
                        C.c_float(b), C.c_float(tamb), _f3(up), C.c_float(dt))
    return [u, v, w]


def source_sphere3(d, n, h, origin, center, radius, value):
    d = np.ascontiguousarray(d, np.float32).copy()
    lib().orc_source_sphere3(_p(d), _i3(n), _f3(h), _f3(origin), _f3(center), C.c_float(radius), C.c_float(value))
    return d


def smoke3_step(st, n, h, origin, dt, a=-0.000625, b=5.0, tamb=0.0, up=(0.0, 1.0, 0.0), smoke_decay=0.001,
                temperature_decay=0.001, flags=63, levels=1, tol=1e-6, max_substeps=32, max_plausible=1e6, order=1):
    """One createGridSmokeSolver3 step (grid_smoke_solver3.js:110-130 over
    grid_solver3.js:128-134) composed from the oracle's stages; st holds u, v,
    w, dens, temp, pressure and the last accepted pressure."""
    vel = buoyancy3([st["u"], st["v"], st["w"]], st["dens"], st["temp"], n, h, origin, a, b, tamb, up, dt)
    vel = constrain_velocity3(vel, n, h, origin, closed_flags=flags)
    A = pressure3_assemble(n, h, vel, pressure=st["pressure"], suffix="f32")
    mg = Multigrid(n, h, levels=levels, mask=A["mask"], coarse_mask=A["vent"], suffix="f32")
    x, _, diag = pcg_solve_mixed(n, h, A["b"], x0=A["pressure"], mask=A["mask"], mg=mg, tol=tol,
                                 maxiter=int(np.ceil(1.5 * max(n))))
    if count_bad(x, max_plausible, suffix="f32") > 0:
        x = st["p_accepted"].copy()
        diag["rejected"] = True
    else:
        st["p_accepted"] = x.copy()
        diag["rejected"] = False
    vel = pressure3_correct(n, h, x, vel, suffix="f32")
    vel = constrain_velocity3(vel, n, h, origin, closed_flags=flags)
    vel = advect_face3(*vel, n, h, origin, dt, max_substeps, order=order)
    vel = constrain_velocity3(vel, n, h, origin, closed_flags=flags)
    raw = advect_scalar3(st["dens"], *vel, n, h, origin, dt, max_substeps, order=order)
    dens = (raw * np.float32(np.float32(1.0) - np.float32(smoke_decay))).astype(np.float32)
    raw = advect_scalar3(st["temp"], *vel, n, h, origin, dt, max_substeps, order=order)
    temp = (raw * np.float32(np.float32(1.0) - np.float32(temperature_decay))).astype(np.float32)
    st.update(u=vel[0], v=vel[1], w=vel[2], dens=dens, temp=temp, pressure=x)
    return diag


# ---------------------------------------------------------------- 2-D Eulerian step
def _i2(v):
    return (C.c_int * 2)(*[int(x) for x in v])


def _f2(v):
    return (C.c_float * 2)(*[float(x) for x in v])


def _d2(v):
    return (C.c_double * 2)(*[float(x) for x in v])


def _rings(rings):
    starts = [0]
    pts = []
    for r in rings:
        for x, y in r:
            pts += [float(x), float(y)]
        starts.append(len(pts) // 2)
    xy = np.array(pts, np.float64)
    st = np.array(starts, np.int32)
    return xy, st, len(rings)


def face_shape2(n, axis):
    return (n[0] + 1, n[1]) if axis == 0 else (n[0], n[1] + 1)


def polygons_distance2(rings, x, y):
    xy, st, nr = _rings(rings)
    lib().orc_polygons_distance2.restype = C.c_double
    return lib().orc_polygons_distance2(_p(xy), _p(st), nr, C.c_double(x), C.c_double(y))


def polygon_sdf2(rings, n, h, origin, invert=False):
    xy, st, nr = _rings(rings)
    out = np.zeros(int(n[0] * n[1]), np.float32)
    lib().orc_polygon_sdf2(_p(xy), _p(st), nr, int(bool(invert)), _i2(n), _d2(h), _d2(origin), _p(out))
    return out


def sample_bilinear2(d, shape, data_origin, h, x, y):
    d = np.ascontiguousarray(d, np.float32)
    lib().orc_sample_bilinear2.restype = C.c_float
    return lib().orc_sample_bilinear2(_p(d), _i2(shape), _f2(data_origin), _f2(h), C.c_float(x), C.c_float(y))


def sample_cubic2(d, shape, data_origin, h, x, y):
    d = np.ascontiguousarray(d, np.float32)
    lib().orc_sample_cubic2.restype = C.c_float
    return lib().orc_sample_cubic2(_p(d), _i2(shape), _f2(data_origin), _f2(h), C.c_float(x), C.c_float(y))


def advect_face2(u, v, n, h, origin, dt, sdf=None, max_substeps=32, order=1, out=None):
    """out: the output's previous contents (kept on faces inside the
    collider); defaults to the input."""
    u = np.ascontiguousarray(u, np.float32)
    v = np.ascontiguousarray(v, np.float32)
    uo, vo = (u.copy(), v.copy()) if out is None else [np.ascontiguousarray(a, np.float32).copy() for a in out]
    sdf = None if sdf is None else np.ascontiguousarray(sdf, np.float32)
    fn = lib().orc_advect_face2 if order == 1 else lib().orc_advect_face2_mc
    fn(_p(u), _p(v), _i2(n), _f2(h), _f2(origin), _p(sdf), C.c_float(dt), int(max_substeps), _p(uo), _p(vo))
    return [uo, vo]


def advect_faces2_vd(vel, data, n, h, origin, dt, sdf=None, max_substeps=32, order=1, out=None):
    """advectFaceCentered2 of the face grid `data` through `vel` with the
    collider; out: the output's previous contents (default: data)."""
    u, v = [np.ascontiguousarray(a, np.float32) for a in vel]
    du, dv = [np.ascontiguousarray(a, np.float32) for a in data]
    uo, vo = (du.copy(), dv.copy()) if out is None else [np.ascontiguousarray(a, np.float32).copy() for a in out]
    sdf = None if sdf is None else np.ascontiguousarray(sdf, np.float32)
    lib().orc_advect_faces2_vd(_p(u), _p(v), _p(du), _p(dv), _i2(n), _f2(h), _f2(origin), _p(sdf), C.c_float(dt),
                               int(max_substeps), int(order), _p(uo), _p(vo))
    return [uo, vo]


def trace2(pos, u, v, n, h, origin, dt, sdf=None, direction=1.0, max_substeps=32):
    pos = np.ascontiguousarray(pos, np.float32).copy()
    u, v = [np.ascontiguousarray(a, np.float32) for a in (u, v)]
    sdf = None if sdf is None else np.ascontiguousarray(sdf, np.float32)
    lib().orc_trace2(_p(pos), C.c_int64(pos.size // 2), _p(u), _p(v), _i2(n), _f2(h), _f2(origin), _p(sdf),
                     C.c_float(dt), int(max_substeps), C.c_float(direction))
    return pos


def constrain_velocity2(vel, n, h, origin, closed_flags=15, sdf=None, friction=0.5, depth=5, inflows=()):
    """inflows: sequence of (sdf, (vx, vy), mode) with mode 'set' or 'add'"""
    u, v = [np.ascontiguousarray(a, np.float32).copy() for a in vel]
    sdf = None if sdf is None else np.ascontiguousarray(sdf, np.float32)
    keep = [np.ascontiguousarray(q[0], np.float32) for q in inflows]
    arr = (C.c_void_p * max(len(keep), 1))(*[a.ctypes.data for a in keep])
    vels = np.array([c for q in inflows for c in q[1]] or [0.0], np.float32)
    modes = np.array([1 if q[2] == "add" else 0 for q in inflows] or [0], np.int32)
    lib().orc_constrain_velocity2(_p(u), _p(v), _i2(n), _f2(h), _f2(origin), int(closed_flags), _p(sdf),
                                  C.c_float(friction), int(depth), len(keep), arr, _p(vels), _p(modes))
    return [u, v]


def outflow_bc2(vel, n, h, origin, dt, osdf):
    u, v = [np.ascontiguousarray(a, np.float32).copy() for a in vel]
    osdf = np.ascontiguousarray(osdf, np.float32)
    lib().orc_outflow_bc2(_p(u), _p(v), _i2(n), _f2(h), _f2(origin), C.c_float(dt), _p(osdf))
    return [u, v]


def pressure2_assemble(n, h, vel, active=None, target=None, pressure=None):
    cells = int(n[0] * n[1])
    u, v = [np.ascontiguousarray(a, np.float32) for a in vel]
    active = _u8(active)
    target = np.zeros(cells, np.float32) if target is None else np.ascontiguousarray(target, np.float32)
    b = np.zeros(cells, np.float32)
    p = np.zeros(cells, np.float32) if pressure is None else np.ascontiguousarray(pressure, np.float32).copy()
    lib().orc_pressure2_assemble(_i2(n), _f2(h), _p(u), _p(v), _p(active), _p(target), _p(b), _p(p))
    return {"b": b, "pressure": p, "mask": active}


def pressure2_correct(n, h, p, vel):
    u, v = [np.ascontiguousarray(a, np.float32).copy() for a in vel]
    p = np.ascontiguousarray(p, np.float32)
    lib().orc_pressure2_correct(_i2(n), _f2(h), _p(p), _p(u), _p(v))
    return [u, v]


def outflow_dirichlet2(osdf, n, h, origin):
    osdf = np.ascontiguousarray(osdf, np.float32)
    active = np.zeros(int(n[0] * n[1]), np.uint8)
    lib().orc_outflow_dirichlet2(_p(osdf), _i2(n), _f2(h), _f2(origin), _p(active))
    return active


def outflow_dirichlet3(outflows, n, h, origin, user_active=None, user_target=None):
    """createOutflowPressureDirichlet3 over a list of outflow SDFs (active =
    inside any, target 0), combined by combineDirichlet3(outflow, user)
    (sdf_inflow_outflow3.js:44-94: active = either; the user's target where
    the user's mask is set, the outflow's 0 elsewhere)"""
    cells = int(np.prod(n))
    act = np.zeros(cells, bool)
    for osdf in outflows:
        a = np.zeros(cells, np.uint8)
        lib().orc_outflow_dirichlet3(_p(np.ascontiguousarray(osdf, np.float32)), _i3(n), _f3(h), _f3(origin), _p(a))
        act |= a != 0
    target = np.zeros(cells, np.float32)
    if user_active is not None:
        ua = np.asarray(user_active).ravel() != 0
        act |= ua
        if user_target is not None:
            target = np.where(ua, np.asarray(user_target, np.float32).ravel(), np.float32(0)).astype(np.float32)
    return act.astype(np.uint8), target


def max_abs(arrays):
    keep = [np.ascontiguousarray(a, np.float32) for a in arrays]
    ptrs = (C.c_void_p * len(keep))(*[a.ctypes.data for a in keep])
    cnt = (C.c_int64 * len(keep))(*[a.size for a in keep])
    lib().orc_max_abs.restype = C.c_float
    return lib().orc_max_abs(ptrs, cnt, len(keep))


def adaptive_substeps(max_velocity, min_spacing, frame_dt, courant=5.0, max_sub=32):
    sub = C.c_double()
    k = lib().orc_adaptive_substeps(C.c_double(max_velocity), C.c_double(min_spacing), C.c_double(frame_dt),
                                    C.c_double(courant), int(max_sub), C.byref(sub))
    return k, sub.value


def solver2_step(st, n, h, origin, dt, sdf=None, inflows=(), outflow_sdf=None, flags=15, levels=1, nu=(2, 2),
                 n_coarsest=20, tol=1e-6, max_plausible=1e6, advect_collider=True, max_substeps=32, depth=5,
                 order=1):
    """One createGridSolver2 step (grid_solver2.js:235-255, no external
    force): projection, constrain, advection of a copy, outflow velocity
    condition, constrain.  st holds u, v, pressure and the last accepted
    pressure; the outflow SDF (cells inside it are Dirichlet p = 0,
    sdf_inflow_outflow2.js:152-166) may be None."""
    vel = [st["u"], st["v"]]
    active = None if outflow_sdf is None else outflow_dirichlet2(outflow_sdf, n, h, origin)
    A = pressure2_assemble(n, h, vel, active=active, pressure=st["pressure"])
    mg = Multigrid(n, h, levels=levels, nu_down=nu[0], nu_up=nu[1], n_coarsest=n_coarsest, mask=A["mask"],
                   suffix="f32")
    x, _, diag = pcg_solve_mixed(n, h, A["b"], x0=A["pressure"], mask=A["mask"], mg=mg, tol=tol,
                                 maxiter=int(np.ceil(1.5 * max(n) * 4)))
    if count_bad(x, max_plausible, suffix="f32") > 0:
        x = st["p_accepted"].copy()
        diag["rejected"] = True
    else:
        st["p_accepted"] = x.copy()
        diag["rejected"] = False
    vel = pressure2_correct(n, h, x, vel)
    vel = constrain_velocity2(vel, n, h, origin, flags, sdf, depth=depth, inflows=inflows)
    vel = advect_face2(vel[0], vel[1], n, h, origin, dt, sdf if advect_collider else None, max_substeps,
                       order=order)
    if outflow_sdf is not None:
        vel = outflow_bc2(vel, n, h, origin, dt, outflow_sdf)
    vel = constrain_velocity2(vel, n, h, origin, flags, sdf, depth=depth, inflows=inflows)
    st.update(u=vel[0], v=vel[1], pressure=x)
    return diag


def solver2_frame(st, n, h, origin, frame_dt, courant=5.0, max_sub=32, **kw):
    """createGridAdaptiveTimeStep2.update() (grid_adaptive_timestep2.js) and
    the frame's sub-steps: max |u| over both face arrays, the sub-step count
    from it, then that many steps of sub_dt (rounded to f32 like the
    reference's dt field)."""
    m = max_abs([st["u"], st["v"]])
    k, sub = adaptive_substeps(float(m), min(h), frame_dt, courant, max_sub)
    dt = float(np.float32(sub))
    diags = [solver2_step(st, n, h, origin, dt, **kw) for _ in range(k)]
    return k, dt, diags


def pcg_coarse_space(shape, spacing, b, mask, weights, mg, vectors, free_indicator=True, tol=1e-6, maxiter=200):
    """PCG with the additive coarse-space correction of the preconditioner,
    M^-1 r = V(r) + Phi (Phi^T A Phi)^-1 Phi^T r (Phi = [free-cell indicator]
    + vectors, each restricted to the free cells), in the GPU plain path's
    arithmetic: f32 r, z, p; f64 solution accumulator and true residual;
    f64 dots; the reference's guards (linalg.js:1147-1311) and beta.  A
    restatement in numpy around the oracle's f32 V-cycle and operator."""
    n = int(np.prod(shape))
    free = (np.asarray(mask) == 0) if mask is not None else np.ones(n, bool)
    cols = ([np.where(free, 1.0, 0.0)] if free_indicator else []) + [np.where(free, np.asarray(v, np.float64), 0.0)
                                                                        for v in vectors]
    Phi = np.stack(cols, 1) if cols else np.zeros((n, 0))
    AP = np.stack([apply(shape, spacing, Phi[:, q], mask, weights, suffix="f64") for q in range(Phi.shape[1])], 1) \
        if cols else np.zeros((n, 0))
    E = Phi.T @ AP
    if cols:
        lam, V = np.linalg.eigh(-0.5 * (E + E.T))
        keep = lam > 1e-12 * lam.max()
        Einv = -(V[:, keep] / lam[keep]) @ V[:, keep].T
    phi32 = Phi.astype(np.float32)
    b32 = np.asarray(b, np.float32)
    x = np.zeros(n)
    r64 = b32.astype(np.float64) - apply(shape, spacing, x, mask, weights, suffix="f64")
    r = r64.astype(np.float32)
    bb = float(np.dot(b32.astype(np.float64), b32.astype(np.float64)))
    thr = max(tol * np.sqrt(bb), tol)
    rr0 = float(np.dot(r64, r64))
    st = {"iterations": 0, "stoppedBy": "none", "converged": False}
    if np.sqrt(rr0) < thr:
        st.update(stoppedBy="converged", converged=True)
        return x, st

    def prec(r):
        z = mg.apply(r).astype(np.float32)
        if cols:
            c = Phi.T @ r.astype(np.float64)
            a = (Einv @ c).astype(np.float32)
            z = (z + (phi32 @ a)).astype(np.float32)
        return z

    z = prec(r)
    rz_old = float(np.dot(r.astype(np.float64), z.astype(np.float64)))
    base = max(abs(rr0), 1e-12)
    best, beta = -1.0, 0.0
    p = np.zeros(n, np.float32)
    for it in range(maxiter):
        st["iterations"] = it + 1
        p = (z + p * np.float32(beta)).astype(np.float32)
        Ap = apply(shape, spacing, p, mask, weights, suffix="f64")
        pAp = float(np.dot(p.astype(np.float64), Ap))
        alpha = rz_old / pAp
        if not np.isfinite(pAp) or pAp == 0:
            st["stoppedBy"] = "degenerate-pAp"
            break
        if abs(pAp) > base * 1e8:
            st["stoppedBy"] = "pAp-growth"
            break
        if not abs(alpha) <= 1e6:
            st["stoppedBy"] = "alpha-magnitude"
            break
        x = x + alpha * p.astype(np.float64)
        r64 = b32.astype(np.float64) - apply(shape, spacing, x, mask, weights, suffix="f64")
        r = r64.astype(np.float32)
        rr = float(np.dot(r64, r64))
        if np.sqrt(rr) <= thr:
            st.update(stoppedBy="converged", converged=True)
            break
        if best < 0 or rr < best:
            best = rr
        elif rr > best * 100:
            st["stoppedBy"] = "residual-growing"
            break
        z = prec(r)
        rz = float(np.dot(r.astype(np.float64), z.astype(np.float64)))
        flip = (rz_old > 0) != (rz > 0)
        braw = rz / rz_old
        beta = 0.0 if (flip or not abs(braw) <= 10) else braw
        rz_old = rz
    st["residual"] = float(np.sqrt(rr)) if st["iterations"] else float(np.sqrt(rr0))
    return x, st


def set_brick(bx=64, by=8, bz=8, passes=2):
    """Brick shape and passes per smoothing of the brick smoother."""
    lib().orc_set_brick(int(bx), int(by), int(bz), int(passes))
