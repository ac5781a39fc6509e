"""tsl_array_n_amd: MI355X-native pressure-projection path of fluxflow.

Python host mirror of the reference's JavaScript API on top of the C-ABI in
``include/fluxflow_hip.h`` (libfluxhip.so, HIP kernels for gfx950):

* ``init`` / ``array_n`` / ``Field``: tsl_array_n's ``init`` and ``arrayN``
  (packages/tsl_array_n/src/context.js, array.js)
* ``create_laplacian_operator``, ``create_multigrid_preconditioner``,
  ``create_pcg_solver``: fluxflow linalg (multigrid.js, linalg.js)
* ``create_grid_pressure_solver3``: grid_pressure_solver3.js
* transfers in ``tsl_array_n_amd.transfer``

Fields are first-index-fastest: a numpy array of shape (nx, ny, nz) is
uploaded as ``a.ravel(order="F")``.  Importing the package requires the built
library; there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libfluxhip.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        "tsl_array_n_amd: %s is missing; build it with `python tsl_array_n_amd/build.py` "
        "(or __graft_entry__.build())" % LIB_PATH)

_lib = C.CDLL(LIB_PATH)

F32, I32, U8, F64, F16, BF16, VEC3 = 0, 1, 2, 3, 4, 5, 6
_NP = {F32: np.float32, I32: np.int32, U8: np.uint8, F64: np.float64, F16: np.float16, VEC3: np.float32}
_TYPE_NAMES = {"float": F32, "int": I32, "u8": U8, "uint8": U8, "f64": F64, "vec3": VEC3, "f16": F16, "bf16": BF16}

PREC = {"multigrid": 0, "jacobi": 1, "none": 2}
STORAGE = {"f32": 0, "f16": 1, "bf16": 2}
STOP_NAMES = {0: "none", 1: "degenerate-pAp", 2: "pAp-growth", 3: "alpha-magnitude", 4: "degenerate-oldRZ",
              5: "converged", 6: "residual-growing"}


class FluxError(RuntimeError):
    pass


_lib.flux_last_error.restype = C.c_char_p
_lib.flux_field_data.restype = C.c_void_p
_lib.flux_pressure3_field.restype = C.c_void_p
_lib.flux_pressure3_field.argtypes = [C.c_void_p, C.c_int]


def _check(rc, what=""):
    if rc != 0:
        msg = _lib.flux_last_error().decode()
        raise FluxError(msg or ("%s failed with code %d" % (what, rc)))


def lib():
    return _lib


class MGOptions(C.Structure):
    _fields_ = [("levels", C.c_int), ("nu_down", C.c_int), ("nu_up", C.c_int), ("n_coarsest", C.c_int),
                ("n_final", C.c_int), ("omega", C.c_double), ("storage", C.c_int), ("smoother", C.c_int),
                ("jacobi_omega", C.c_double)]


class PCGState(C.Structure):
    _fields_ = [("iterations", C.c_int), ("stopped_by", C.c_int), ("converged", C.c_int),
                ("residual", C.c_double), ("threshold", C.c_double), ("b_norm", C.c_double),
                ("restarts", C.c_int)]


class Pressure3Desc(C.Structure):
    _fields_ = [("resolution", C.c_int * 3), ("spacing", C.c_double * 3), ("origin", C.c_double * 3),
                ("dirichlet_active", C.c_void_p), ("dirichlet_target", C.c_void_p),
                ("face_weights", C.c_void_p * 3), ("collider_weights", C.c_void_p * 3),
                ("multigrid", MGOptions), ("preconditioner", C.c_int), ("tolerance", C.c_double),
                ("max_iterations", C.c_int), ("max_iterations_factor", C.c_double),
                ("relative_tolerance", C.c_int), ("max_plausible_pressure", C.c_double)]


class PressureDiag(C.Structure):
    _fields_ = [("converged", C.c_int), ("rejected", C.c_int), ("iterations", C.c_int), ("stopped_by", C.c_int),
                ("residual", C.c_double), ("b_norm", C.c_double), ("bad_cells", C.c_int64)]


class Grid3(C.Structure):
    _fields_ = [("n", C.c_int * 3), ("h", C.c_float * 3), ("origin", C.c_float * 3)]


_initialised = False


def init(device=0):
    """tsl_array_n init(): selects the device, creates the library stream."""
    global _initialised
    _check(_lib.flux_init(int(device)), "init")
    _initialised = True


def shutdown():
    global _initialised
    _check(_lib.flux_shutdown())
    _initialised = False


def sync():
    _check(_lib.flux_sync())


def device_info():
    buf = C.create_string_buffer(256)
    cus = C.c_int()
    _check(_lib.flux_device_info(buf, 256, C.byref(cus)))
    return buf.value.decode(), cus.value


class Field:
    """arrayN(type, shape): a device field owned by the library."""

    def __init__(self, dtype, shape, handle=None, owned=True):
        if isinstance(dtype, str):
            if dtype not in _TYPE_NAMES:
                raise FluxError("arrayN: unknown type %r" % dtype)
            dtype = _TYPE_NAMES[dtype]
        if isinstance(shape, int):
            shape = (shape,)
        self.dtype = dtype
        self.shape = tuple(int(s) for s in shape)
        self.count = int(np.prod(self.shape))
        self._owned = owned
        if handle is None:
            h = C.c_void_p()
            arr = (C.c_int64 * len(self.shape))(*self.shape)
            _check(_lib.flux_field_create(int(dtype), len(self.shape), arr, C.byref(h)), "arrayN")
            handle = h.value
        self.h = handle

    @property
    def np_dtype(self):
        return _NP[self.dtype]

    @property
    def nbytes(self):
        return self.count * (12 if self.dtype == VEC3 else np.dtype(self.np_dtype).itemsize)

    def from_array(self, a):
        """fromArray(typed): a flat (first-index-fastest) array."""
        a = np.ascontiguousarray(np.asarray(a).ravel(), dtype=self.np_dtype)
        if a.nbytes != self.nbytes:
            raise FluxError("fromArray: %d bytes given, field holds %d" % (a.nbytes, self.nbytes))
        _check(_lib.flux_field_upload(C.c_void_p(self.h), a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes)))
        return self

    def to_array(self):
        """toArray(): flat copy in host memory."""
        n = self.count * (3 if self.dtype == VEC3 else 1)
        out = np.empty(n, dtype=self.np_dtype)
        _check(_lib.flux_field_download(C.c_void_p(self.h), out.ctypes.data_as(C.c_void_p), C.c_size_t(out.nbytes)))
        return out

    def fill(self, value):
        _check(_lib.flux_field_fill(C.c_void_p(self.h), C.c_double(value)))
        return self

    def copy_from(self, other):
        _check(_lib.flux_field_copy(C.c_void_p(self.h), C.c_void_p(other.h)))
        return self

    def data_ptr(self):
        return _lib.flux_field_data(C.c_void_p(self.h))

    def free(self):
        if self.h and self._owned:
            _lib.flux_field_destroy(C.c_void_p(self.h))
        self.h = None

    def __del__(self):
        try:
            if _initialised:
                self.free()
        except Exception:
            pass


def array_n(dtype, shape, data=None):
    f = Field(dtype, shape)
    if data is not None:
        f.from_array(data)
    return f


def _fields3(fields, n):
    if fields is None:
        return None
    arr = (C.c_void_p * 3)(*([f.h for f in fields] + [None] * (3 - len(fields))))
    return arr


SMOOTHER = {"rbgs": 0, "jacobi": 1, "brick": 2}


def mg_options(levels=1, nu_down=2, nu_up=2, n_coarsest=20, n_final=2, omega=1.0, storage="f32", smoother="rbgs",
               jacobi_omega=2.0 / 3.0, **_):
    o = MGOptions()
    o.levels, o.nu_down, o.nu_up, o.n_coarsest, o.n_final, o.omega = levels, nu_down, nu_up, n_coarsest, n_final, omega
    o.storage = STORAGE[storage] if isinstance(storage, str) else int(storage)
    o.smoother = SMOOTHER[smoother] if isinstance(smoother, str) else int(smoother)
    o.jacobi_omega = jacobi_omega
    return o


class LaplacianOperator:
    """createLaplacianOperator(shape, gridSpacing, {dirichletMask, faceWeights})."""

    def __init__(self, shape, spacing, dirichlet_mask=None, face_weights=None):
        if len(spacing) != len(shape):
            raise FluxError("createLaplacianOperator: gridSpacing length (%d) must match shape length (%d)."
                            % (len(spacing), len(shape)))
        self.shape = tuple(shape)
        self._keep = (dirichlet_mask, face_weights)
        h = C.c_void_p()
        sh = (C.c_int * len(shape))(*shape)
        sp = (C.c_double * len(spacing))(*spacing)
        w = _fields3(face_weights, len(shape))
        _check(_lib.flux_operator_create(len(shape), sh, sp, C.c_void_p(dirichlet_mask.h if dirichlet_mask else None),
                                         w, C.byref(h)), "createLaplacianOperator")
        self.h = h.value

    def __call__(self, inp, out):
        _check(_lib.flux_operator_apply(C.c_void_p(self.h), C.c_void_p(inp.h), C.c_void_p(out.h)))

    def diagonal(self, out):
        _check(_lib.flux_operator_diagonal(C.c_void_p(self.h), C.c_void_p(out.h)))

    def __del__(self):
        try:
            if _initialised and self.h:
                _lib.flux_operator_destroy(C.c_void_p(self.h))
        except Exception:
            pass


def create_laplacian_operator(shape, spacing, dirichlet_mask=None, face_weights=None):
    return LaplacianOperator(shape, spacing, dirichlet_mask, face_weights)


class MultigridPreconditioner:
    """createMultigridPreconditioner(shape, gridSpacing, options)."""

    def __init__(self, shape, spacing, dirichlet_mask=None, coarse_dirichlet_mask=None, face_weights=None, **opts):
        if len(shape) < 1 or len(shape) > 3:
            raise FluxError("multigrid: only 1D/2D/3D shapes are supported, got %dD." % len(shape))
        if len(spacing) != len(shape):
            raise FluxError("multigrid: gridSpacing length (%d) must match shape length (%d)." % (len(spacing), len(shape)))
        self.shape = tuple(shape)
        self.options = mg_options(**opts)
        self._keep = (dirichlet_mask, coarse_dirichlet_mask, face_weights)
        h = C.c_void_p()
        sh = (C.c_int * len(shape))(*shape)
        sp = (C.c_double * len(spacing))(*spacing)
        _check(_lib.flux_multigrid_create(len(shape), sh, sp, C.byref(self.options),
                                          C.c_void_p(dirichlet_mask.h if dirichlet_mask else None),
                                          C.c_void_p(coarse_dirichlet_mask.h if coarse_dirichlet_mask else None),
                                          _fields3(face_weights, len(shape)), C.byref(h)), "multigrid")
        self.h = h.value

    def __call__(self, inp, out):
        _check(_lib.flux_multigrid_apply(C.c_void_p(self.h), C.c_void_p(inp.h), C.c_void_p(out.h)))

    def refresh_coarse_levels(self):
        _check(_lib.flux_multigrid_refresh(C.c_void_p(self.h)))

    def level(self, l):
        shape = (C.c_int * 3)()
        _check(_lib.flux_multigrid_level(C.c_void_p(self.h), int(l), shape, None, None))
        dims = tuple(shape[a] for a in range(len(self.shape)))
        full = [shape[0], shape[1], shape[2]]
        mask = np.zeros(int(np.prod(full)), np.uint8)
        ws = []
        for a in range(len(self.shape)):
            fs = list(full)
            fs[a] += 1
            ws.append(np.zeros(int(np.prod(fs)), np.float32))
        wp = (C.c_void_p * 3)(*([w.ctypes.data for w in ws] + [None] * (3 - len(ws))))
        _check(_lib.flux_multigrid_level(C.c_void_p(self.h), int(l), shape, mask.ctypes.data_as(C.c_void_p), wp))
        return dims, mask, ws

    def __del__(self):
        try:
            if _initialised and self.h:
                _lib.flux_multigrid_destroy(C.c_void_p(self.h))
        except Exception:
            pass


def create_multigrid_preconditioner(shape, spacing, **options):
    return MultigridPreconditioner(shape, spacing, **options)


class PCGSolver:
    """createPreconditionedConjugateGradientSolver(applyOperator, applyPreconditioner, b, x)."""

    def __init__(self, operator, b, x, preconditioner="multigrid", multigrid=None):
        if b.dtype != x.dtype:
            raise FluxError("conjugateGradient: element type mismatch, b.type(%s) != x.type(%s)." % (b.dtype, x.dtype))
        if b.shape != x.shape:
            raise FluxError("conjugateGradient: shape mismatch, b.shape(%s) != x.shape(%s)." % (b.shape, x.shape))
        self._keep = (operator, b, x, multigrid)
        self.b, self.x = b, x
        h = C.c_void_p()
        _check(_lib.flux_pcg_create(C.c_void_p(operator.h), PREC[preconditioner],
                                    C.c_void_p(multigrid.h if multigrid else None), C.c_void_p(b.h), C.c_void_p(x.h),
                                    C.byref(h)), "conjugateGradient")
        self.h = h.value
        self.state = {}

    def set_coarse_space(self, free_indicator=True, fields=()):
        fields = list(fields)
        self._cs_keep = fields
        arr = (C.c_void_p * max(1, len(fields)))(*[f.h for f in fields]) if fields else None
        _check(_lib.flux_pcg_set_coarse_space(C.c_void_p(self.h), int(bool(free_indicator)), len(fields), arr),
               "coarseSpace")

    def set_flexible(self, on=True):
        _check(_lib.flux_pcg_set_flexible(C.c_void_p(self.h), int(bool(on))), "flexible")

    def solve(self, tol, maxiter, relative_tolerance=False):
        st = PCGState()
        _check(_lib.flux_pcg_solve(C.c_void_p(self.h), C.c_double(tol), int(maxiter), int(bool(relative_tolerance)),
                                   C.byref(st)), "solve")
        self.state = {"iterations": st.iterations, "stoppedBy": STOP_NAMES.get(st.stopped_by, "unknown"),
                      "converged": bool(st.converged), "residual": st.residual, "threshold": st.threshold,
                      "bNorm": st.b_norm, "restarts": st.restarts}
        return bool(st.converged)

    def __del__(self):
        try:
            if _initialised and self.h:
                _lib.flux_pcg_destroy(C.c_void_p(self.h))
        except Exception:
            pass


def create_pcg_solver(operator, preconditioner, b, x, multigrid=None):
    return PCGSolver(operator, b, x, preconditioner, multigrid)


class DotReducer:
    """createDotReducer(shape, fieldA, fieldB) (linalg.js:374-512): ``read()``
    returns sum(a * b); ``lanes`` and the per-lane ``partial`` field as in
    the reference."""

    def __init__(self, shape, a, b):
        shape = tuple(int(s) for s in shape)
        self._keep = (a, b)
        h = C.c_void_p()
        sh = (C.c_int * max(1, len(shape)))(*shape)
        _check(_lib.flux_dot_create(len(shape), sh, C.c_void_p(a.h), C.c_void_p(b.h), C.byref(h)), "createDotReducer")
        self.h = h.value
        lanes = C.c_int()
        part = C.c_void_p()
        _check(_lib.flux_dot_info(C.c_void_p(self.h), C.byref(lanes), C.byref(part)))
        self.lanes = lanes.value
        self.partial = Field(F32, (self.lanes,), handle=part.value, owned=False)

    def dispatch(self):
        _check(_lib.flux_dot_dispatch(C.c_void_p(self.h)), "createDotReducer")

    def read(self):
        t = C.c_double()
        _check(_lib.flux_dot_read(C.c_void_p(self.h), C.byref(t)), "createDotReducer")
        return t.value

    def __del__(self):
        try:
            if _initialised and self.h:
                _lib.flux_dot_destroy(C.c_void_p(self.h))
        except Exception:
            pass


def create_dot_reducer(shape, a, b):
    return DotReducer(shape, a, b)


class CGSolver:
    """createConjugateGradientSolver(applyOperator, b, x) (linalg.js:563-728):
    ``operator`` is a LaplacianOperator, or a Field holding the diagonal of a
    diagonal operator."""

    def __init__(self, operator, b, x):
        if b.dtype != x.dtype:
            raise FluxError("conjugateGradient: element type mismatch, b.type(%s) != x.type(%s)." % (b.dtype, x.dtype))
        if b.shape != x.shape:
            raise FluxError("conjugateGradient: shape mismatch, b.shape(%s) != x.shape(%s)." % (b.shape, x.shape))
        self._keep = (operator, b, x)
        diag = isinstance(operator, Field)
        h = C.c_void_p()
        _check(_lib.flux_cg_create(C.c_void_p(None if diag else operator.h), C.c_void_p(operator.h if diag else None),
                                   C.c_void_p(b.h), C.c_void_p(x.h), C.byref(h)), "conjugateGradient")
        self.h = h.value
        self.iterations = 0

    def solve(self, tol, maxiter):
        ok = C.c_int()
        it = C.c_int()
        _check(_lib.flux_cg_solve(C.c_void_p(self.h), C.c_double(tol), int(maxiter), C.byref(ok), C.byref(it)), "solve")
        self.iterations = it.value
        return bool(ok.value)

    def __del__(self):
        try:
            if _initialised and self.h:
                _lib.flux_cg_destroy(C.c_void_p(self.h))
        except Exception:
            pass


def create_cg_solver(operator, b, x):
    return CGSolver(operator, b, x)


class GridPressureSolver3:
    """createGridPressureSolver3({resolution, gridSpacing, dirichlet, faceWeights, colliderWeights, ...}).

    ``dirichlet`` is a pair (active u8 field, target f32 field) evaluated per
    cell; ``face_weights`` / ``collider_weights`` are triples of face fields.
    """

    def __init__(self, resolution, grid_spacing=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0), dirichlet=None,
                 face_weights=None, collider_weights=None, multigrid=None, preconditioner="multigrid",
                 tolerance=1e-6, max_iterations=None, max_iterations_factor=1.5, relative_tolerance=True,
                 max_plausible_pressure=1e6):
        if preconditioner not in PREC:
            raise FluxError("grid_pressure_solver3: unknown preconditioner '%s' (expected multigrid, jacobi, none)."
                            % preconditioner)
        d = Pressure3Desc()
        _lib.flux_pressure3_desc_default(C.byref(d))
        self.resolution = tuple(resolution)
        for a in range(3):
            d.resolution[a] = int(resolution[a])
            d.spacing[a] = float(grid_spacing[a])
            d.origin[a] = float(origin[a])
        if dirichlet is not None:
            d.dirichlet_active = dirichlet[0].h
            d.dirichlet_target = dirichlet[1].h if dirichlet[1] is not None else None
        if face_weights is not None:
            for a in range(3):
                d.face_weights[a] = face_weights[a].h
        if collider_weights is not None:
            for a in range(3):
                d.collider_weights[a] = collider_weights[a].h
        d.multigrid = mg_options(**(multigrid or {}))
        d.preconditioner = PREC[preconditioner]
        d.tolerance = tolerance
        d.max_iterations = int(max_iterations or 0)
        d.max_iterations_factor = max_iterations_factor
        d.relative_tolerance = int(bool(relative_tolerance))
        d.max_plausible_pressure = max_plausible_pressure
        self._keep = (dirichlet, face_weights, collider_weights)
        h = C.c_void_p()
        _check(_lib.flux_pressure3_create(C.byref(d), C.byref(h)), "createGridPressureSolver3")
        self.h = h.value
        self.diagnostics = {}
        n = self.resolution
        self.pressure = self._view(0, F32, n)
        self.b = self._view(1, F32, n)
        self.dirichlet_mask = self._view(2, U8, n)
        self.vent_mask = self._view(3, U8, n)
        self.weights = None
        if face_weights is not None or collider_weights is not None:
            self.weights = [self._view(5 + a, F32, face_shape(n, a)) for a in range(3)]

    def _view(self, which, dtype, shape):
        ptr = _lib.flux_pressure3_field(C.c_void_p(self.h), which)
        return Field(dtype, shape, handle=ptr, owned=False) if ptr else None

    def assemble(self, vin):
        """first half of project(): Dirichlet fields, weights, right-hand side, coarse levels"""
        _check(_lib.flux_pressure3_assemble(C.c_void_p(self.h), *[C.c_void_p(f.h) for f in vin]), "assemble")

    def scaled_guess(self, guess):
        """between assemble() and solve(): pressure += alpha * guess on the free
        cells, alpha minimising the energy-norm error along the guess; returns alpha"""
        alpha = C.c_double()
        _check(_lib.flux_pressure3_scaled_guess(C.c_void_p(self.h), C.c_void_p(guess.h), C.byref(alpha)),
               "scaledGuess")
        return alpha.value

    def set_coarse_space(self, free_indicator=True, fields=()):
        """the additive coarse-space correction of the preconditioner:
        M^-1 r = V(r) + Phi (Phi^T A Phi)^-1 Phi^T r, Phi = [free-cell
        indicator] + fields (each restricted to the free cells); the fields
        stay bound to the solver"""
        fields = list(fields)
        self._cs_keep = fields
        arr = (C.c_void_p * max(1, len(fields)))(*[f.h for f in fields]) if fields else None
        _check(_lib.flux_pressure3_set_coarse_space(C.c_void_p(self.h), int(bool(free_indicator)), len(fields), arr),
               "coarseSpace")

    def set_flexible(self, on=True):
        """the flexible (Polak-Ribiere) beta in the projection's PCG"""
        _check(_lib.flux_pressure3_set_flexible(C.c_void_p(self.h), int(bool(on))), "flexible")

    def solve(self, vin, vout=None):
        """second half of project(): PCG from the current pressure, breaker, correction"""
        vout = vin if vout is None else vout
        dg = PressureDiag()
        _check(_lib.flux_pressure3_solve(C.c_void_p(self.h), *[C.c_void_p(f.h) for f in vin],
                                         *[C.c_void_p(f.h) for f in vout], C.byref(dg)), "solve")
        return self._diag(dg)

    def project(self, vin, vout=None):
        """project(input, output) + dispatch(): vin/vout are (U, V, W) fields."""
        vout = vin if vout is None else vout
        dg = PressureDiag()
        _check(_lib.flux_pressure3_project(C.c_void_p(self.h), *[C.c_void_p(f.h) for f in vin],
                                           *[C.c_void_p(f.h) for f in vout], C.byref(dg)), "project")
        return self._diag(dg)

    def _diag(self, dg):
        self.diagnostics = {"converged": bool(dg.converged), "rejected": bool(dg.rejected), "iterations": dg.iterations,
                            "stoppedBy": STOP_NAMES.get(dg.stopped_by, "unknown"), "residual": dg.residual,
                            "bNorm": dg.b_norm, "badCells": dg.bad_cells}
        return self.diagnostics

    def __del__(self):
        try:
            if _initialised and self.h:
                _lib.flux_pressure3_destroy(C.c_void_p(self.h))
        except Exception:
            pass


def create_grid_pressure_solver3(**kw):
    return GridPressureSolver3(**kw)


def face_shape(shape, axis):
    s = list(shape)
    s[axis] += 1
    return tuple(s)


def timer_start():
    _check(_lib.flux_timer_start())


def timer_stop():
    ms = C.c_double()
    _check(_lib.flux_timer_stop(C.byref(ms)))
    return ms.value


def kprof_enable(name):
    _check(_lib.flux_kprof_enable(name.encode() if name else None))


def kprof_read():
    ms = C.c_double()
    n = C.c_int64()
    _check(_lib.flux_kprof_read(C.byref(ms), C.byref(n)))
    return ms.value, n.value
