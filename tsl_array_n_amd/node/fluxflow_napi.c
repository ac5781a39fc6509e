/*
 * N-API addon: the thin JavaScript binding of include/fluxflow_hip.h.
 *
 * Every export maps one C-ABI entry point; handles are napi externals whose
 * finalizers release the native object (fields the library owns elsewhere,
 * like a pressure solver's own pressure field, are wrapped without one).
 * Errors become JavaScript exceptions carrying flux_last_error(), so the
 * reference's messages ("not divisible", "shape mismatch", ...) surface
 * unchanged.  The ESM module index.mjs builds the reference's API on top.
 */
#define NAPI_VERSION 6
#include <node_api.h>
#include <stdint.h>
#include <stdlib.h>
#include <pthread.h>
#include <string.h>

#include "../../include/fluxflow_hip.h"

#define CHECK_NAPI(call)                                                      \
    do {                                                                      \
        if ((call) != napi_ok) {                                              \
            napi_throw_error(env, NULL, "fluxflow addon: N-API call failed"); \
            return NULL;                                                      \
        }                                                                     \
    } while (0)

#define CHECK_FLUX(call)                                      \
    do {                                                      \
        if ((call) != FLUX_OK) {                              \
            napi_throw_error(env, NULL, flux_last_error());   \
            return NULL;                                      \
        }                                                     \
    } while (0)

enum { K_FIELD = 1, K_FIELD_VIEW, K_OP, K_MG, K_PCG, K_P3, K_EXTRAP, K_BND, K_FLIP, K_SMOKE, K_S2, K_ADV3, K_DOT, K_CG };

typedef struct {
    int kind;
    void* ptr;
} handle_t;

static void fin_handle(napi_env env, void* data, void* hint) {
    (void)env;
    (void)hint;
    handle_t* h = (handle_t*)data;
    if (!h) return;
    switch (h->kind) {
        case K_FIELD: flux_field_destroy((flux_field*)h->ptr); break;
        case K_OP: flux_operator_destroy((flux_operator*)h->ptr); break;
        case K_MG: flux_multigrid_destroy((flux_multigrid*)h->ptr); break;
        case K_PCG: flux_pcg_destroy((flux_pcg*)h->ptr); break;
        case K_P3: flux_pressure3_destroy((flux_pressure3*)h->ptr); break;
        case K_EXTRAP: flux_extrapolator3_destroy((flux_extrapolator*)h->ptr); break;
        case K_BND: flux_boundary3_destroy((flux_boundary3*)h->ptr); break;
        case K_FLIP: flux_flip3_destroy((flux_flip3*)h->ptr); break;
        case K_SMOKE: flux_smoke3_destroy((flux_smoke3*)h->ptr); break;
        case K_S2: flux_solver2_destroy((flux_solver2*)h->ptr); break;
        case K_ADV3: flux_advection3_destroy((flux_advection3*)h->ptr); break;
        case K_DOT: flux_dot_destroy((flux_dot*)h->ptr); break;
        case K_CG: flux_cg_destroy((flux_cg*)h->ptr); break;
        default: break;
    }
    free(h);
}

static napi_value wrap(napi_env env, int kind, void* ptr) {
    handle_t* h = (handle_t*)malloc(sizeof(handle_t));
    h->kind = kind;
    h->ptr = ptr;
    napi_value v;
    if (napi_create_external(env, h, fin_handle, NULL, &v) != napi_ok) return NULL;
    return v;
}

static void* unwrap(napi_env env, napi_value v, int kind_a, int kind_b) {
    napi_valuetype t;
    if (napi_typeof(env, v, &t) != napi_ok || t != napi_external) return NULL;
    handle_t* h = NULL;
    if (napi_get_value_external(env, v, (void**)&h) != napi_ok || !h) return NULL;
    if (h->kind != kind_a && h->kind != kind_b) return NULL;
    return h->ptr;
}

static int is_nullish(napi_env env, napi_value v) {
    napi_valuetype t;
    if (napi_typeof(env, v, &t) != napi_ok) return 1;
    return t == napi_null || t == napi_undefined;
}

static flux_field* get_field(napi_env env, napi_value v) { return (flux_field*)unwrap(env, v, K_FIELD, K_FIELD_VIEW); }

static int get_int(napi_env env, napi_value v, int dflt) {
    int32_t x;
    if (napi_get_value_int32(env, v, &x) != napi_ok) return dflt;
    return x;
}

static double get_double(napi_env env, napi_value v, double dflt) {
    double x;
    if (napi_get_value_double(env, v, &x) != napi_ok) return dflt;
    return x;
}

static napi_value prop(napi_env env, napi_value obj, const char* name) {
    napi_value v = NULL;
    bool has = false;
    if (napi_has_named_property(env, obj, name, &has) != napi_ok || !has) return NULL;
    napi_get_named_property(env, obj, name, &v);
    return v;
}

static int get_int_array(napi_env env, napi_value arr, int* out, int max) {
    uint32_t n = 0;
    if (napi_get_array_length(env, arr, &n) != napi_ok) return -1;
    if ((int)n > max) return -1;
    for (uint32_t i = 0; i < n; ++i) {
        napi_value e;
        napi_get_element(env, arr, i, &e);
        out[i] = get_int(env, e, 0);
    }
    return (int)n;
}

static int get_double_array(napi_env env, napi_value arr, double* out, int max) {
    uint32_t n = 0;
    if (napi_get_array_length(env, arr, &n) != napi_ok) return -1;
    if ((int)n > max) return -1;
    for (uint32_t i = 0; i < n; ++i) {
        napi_value e;
        napi_get_element(env, arr, i, &e);
        out[i] = get_double(env, e, 0.0);
    }
    return (int)n;
}

/* array of up to 3 fields, or null -> returns count, fills out */
static int get_field_array_n(napi_env env, napi_value arr, const flux_field** out, uint32_t max) {
    if (is_nullish(env, arr)) return 0;
    uint32_t n = 0;
    if (napi_get_array_length(env, arr, &n) != napi_ok || n > max) return -1;
    for (uint32_t i = 0; i < n; ++i) {
        napi_value e;
        napi_get_element(env, arr, i, &e);
        out[i] = get_field(env, e);
        if (!out[i]) return -1;
    }
    return (int)n;
}

static int get_field_array(napi_env env, napi_value arr, const flux_field** out) {
    return get_field_array_n(env, arr, out, 3);
}

static napi_value args_n(napi_env env, napi_callback_info info, size_t want, napi_value* argv) {
    size_t argc = want;
    napi_value self;
    if (napi_get_cb_info(env, info, &argc, argv, &self, NULL) != napi_ok) return NULL;
    for (size_t i = argc; i < want; ++i) napi_get_undefined(env, &argv[i]);
    return self;
}

static napi_value undef(napi_env env) {
    napi_value u;
    napi_get_undefined(env, &u);
    return u;
}

static napi_value num(napi_env env, double x) {
    napi_value v;
    napi_create_double(env, x, &v);
    return v;
}

/* ------------------------------------------------------------- context */
static napi_value js_init(napi_env env, napi_callback_info info) {
    napi_value a[1];
    args_n(env, info, 1, a);
    CHECK_FLUX(flux_init(is_nullish(env, a[0]) ? 0 : get_int(env, a[0], 0)));
    return undef(env);
}

static napi_value js_sync(napi_env env, napi_callback_info info) {
    (void)info;
    CHECK_FLUX(flux_sync());
    return undef(env);
}

static napi_value js_device_info(napi_env env, napi_callback_info info) {
    (void)info;
    char name[256];
    int cus = 0;
    CHECK_FLUX(flux_device_info(name, sizeof name, &cus));
    napi_value o, s;
    napi_create_object(env, &o);
    napi_create_string_utf8(env, name, NAPI_AUTO_LENGTH, &s);
    napi_set_named_property(env, o, "name", s);
    napi_set_named_property(env, o, "computeUnits", num(env, cus));
    return o;
}

static napi_value js_last_error(napi_env env, napi_callback_info info) {
    (void)info;
    napi_value s;
    napi_create_string_utf8(env, flux_last_error(), NAPI_AUTO_LENGTH, &s);
    return s;
}

/* -------------------------------------------------------------- fields */
static napi_value js_field_create(napi_env env, napi_callback_info info) {
    napi_value a[2];
    args_n(env, info, 2, a);
    int dtype = get_int(env, a[0], -1);
    int shp[3];
    int nd = get_int_array(env, a[1], shp, 3);
    if (nd <= 0) {
        napi_throw_error(env, NULL, "arrayN: shape must be an array of 1..3 positive integers");
        return NULL;
    }
    int64_t s64[3] = {1, 1, 1};
    for (int i = 0; i < nd; ++i) s64[i] = shp[i];
    flux_field* f = NULL;
    CHECK_FLUX(flux_field_create(dtype, nd, s64, &f));
    return wrap(env, K_FIELD, f);
}

static napi_value js_field_upload(napi_env env, napi_callback_info info) {
    napi_value a[2];
    args_n(env, info, 2, a);
    flux_field* f = get_field(env, a[0]);
    void* data = NULL;
    size_t len = 0;
    napi_typedarray_type tt;
    napi_value ab;
    size_t off = 0;
    if (!f || napi_get_typedarray_info(env, a[1], &tt, &len, &data, &ab, &off) != napi_ok) {
        napi_throw_type_error(env, NULL, "fromArray: expected (field, TypedArray)");
        return NULL;
    }
    size_t esize = (tt == napi_float64_array) ? 8 : (tt == napi_uint8_array || tt == napi_int8_array ||
                                                     tt == napi_uint8_clamped_array) ? 1
                                                   : (tt == napi_int16_array || tt == napi_uint16_array) ? 2 : 4;
    CHECK_FLUX(flux_field_upload(f, data, len * esize));
    return undef(env);
}

static napi_value js_field_download(napi_env env, napi_callback_info info) {
    napi_value a[2];
    args_n(env, info, 2, a);
    flux_field* f = get_field(env, a[0]);
    void* data = NULL;
    size_t len = 0;
    napi_typedarray_type tt;
    napi_value ab;
    size_t off = 0;
    if (!f || napi_get_typedarray_info(env, a[1], &tt, &len, &data, &ab, &off) != napi_ok) {
        napi_throw_type_error(env, NULL, "toArray: expected (field, TypedArray)");
        return NULL;
    }
    size_t esize = (tt == napi_float64_array) ? 8 : (tt == napi_uint8_array || tt == napi_int8_array ||
                                                     tt == napi_uint8_clamped_array) ? 1
                                                   : (tt == napi_int16_array || tt == napi_uint16_array) ? 2 : 4;
    CHECK_FLUX(flux_field_download(f, data, len * esize));
    return a[1];
}

static napi_value js_field_fill(napi_env env, napi_callback_info info) {
    napi_value a[2];
    args_n(env, info, 2, a);
    flux_field* f = get_field(env, a[0]);
    if (!f) {
        napi_throw_type_error(env, NULL, "fill: expected a field");
        return NULL;
    }
    CHECK_FLUX(flux_field_fill(f, get_double(env, a[1], 0.0)));
    return undef(env);
}

static napi_value js_field_copy(napi_env env, napi_callback_info info) {
    napi_value a[2];
    args_n(env, info, 2, a);
    flux_field* d = get_field(env, a[0]);
    flux_field* s = get_field(env, a[1]);
    if (!d || !s) {
        napi_throw_type_error(env, NULL, "copy: expected two fields");
        return NULL;
    }
    CHECK_FLUX(flux_field_copy(d, s));
    CHECK_FLUX(flux_sync());
    return undef(env);
}

/* ------------------------------------------------------------ operator */
static napi_value js_operator_create(napi_env env, napi_callback_info info) {
    napi_value a[4];
    args_n(env, info, 4, a);
    int shp[3];
    double sp[3];
    int nd = get_int_array(env, a[0], shp, 3);
    int ns = get_double_array(env, a[1], sp, 3);
    const flux_field* w[3] = {NULL, NULL, NULL};
    int nw = get_field_array(env, a[3], w);
    if (nd < 0 || ns < 0 || nw < 0) {
        napi_throw_type_error(env, NULL, "createLaplacianOperator: bad arguments");
        return NULL;
    }
    flux_operator* op = NULL;
    CHECK_FLUX(flux_operator_create(nd, shp, sp, is_nullish(env, a[2]) ? NULL : get_field(env, a[2]),
                                    nw ? w : NULL, &op));
    return wrap(env, K_OP, op);
}

static napi_value js_operator_apply(napi_env env, napi_callback_info info) {
    napi_value a[3];
    args_n(env, info, 3, a);
    flux_operator* op = (flux_operator*)unwrap(env, a[0], K_OP, K_OP);
    CHECK_FLUX(flux_operator_apply(op, get_field(env, a[1]), get_field(env, a[2])));
    return undef(env);
}

/* ----------------------------------------------------------- multigrid */
static void mg_opts(napi_env env, napi_value o, flux_mg_options* m) {
    flux_mg_options_default(m);
    if (is_nullish(env, o)) return;
    napi_value v;
    if ((v = prop(env, o, "numberOfLevels"))) m->levels = get_int(env, v, m->levels);
    if ((v = prop(env, o, "numberOfSmoothingIterationsDown"))) m->nu_down = get_int(env, v, m->nu_down);
    if ((v = prop(env, o, "numberOfSmoothingIterationsUp"))) m->nu_up = get_int(env, v, m->nu_up);
    if ((v = prop(env, o, "numberOfCoarsestIterations"))) m->n_coarsest = get_int(env, v, m->n_coarsest);
    if ((v = prop(env, o, "numberOfFinalIterations"))) m->n_final = get_int(env, v, m->n_final);
    if ((v = prop(env, o, "sorFactor"))) m->omega = get_double(env, v, m->omega);
    if ((v = prop(env, o, "storage"))) {
        /* 'f32' | 'f16' | 'bf16' (or the FLUX_STORE_* number) */
        napi_valuetype t;
        char name[16] = {0};
        size_t len = 0;
        if (napi_typeof(env, v, &t) == napi_ok && t == napi_string) {
            napi_get_value_string_utf8(env, v, name, sizeof(name), &len);
            if (strcmp(name, "f32") == 0) m->storage = FLUX_STORE_F32;
            else if (strcmp(name, "f16") == 0) m->storage = FLUX_STORE_F16;
            else if (strcmp(name, "bf16") == 0) m->storage = FLUX_STORE_BF16;
            else m->storage = -1; /* rejected by the library with its own message */
        } else {
            m->storage = get_int(env, v, m->storage);
        }
    }
    if ((v = prop(env, o, "smoother"))) {
        /* 'rbgs' | 'jacobi' | 'brick' (or the FLUX_SMOOTH_* number) */
        napi_valuetype t;
        char name[16] = {0};
        size_t len = 0;
        if (napi_typeof(env, v, &t) == napi_ok && t == napi_string) {
            napi_get_value_string_utf8(env, v, name, sizeof(name), &len);
            if (strcmp(name, "rbgs") == 0) m->smoother = FLUX_SMOOTH_RBGS;
            else if (strcmp(name, "jacobi") == 0) m->smoother = FLUX_SMOOTH_JACOBI;
            else if (strcmp(name, "brick") == 0) m->smoother = FLUX_SMOOTH_BRICK;
            else m->smoother = -1; /* rejected by the library with its own message */
        } else {
            m->smoother = get_int(env, v, m->smoother);
        }
    }
    if ((v = prop(env, o, "omega"))) m->jacobi_omega = get_double(env, v, m->jacobi_omega);
}

static napi_value js_multigrid_create(napi_env env, napi_callback_info info) {
    napi_value a[6];
    args_n(env, info, 6, a);
    int shp[3];
    double sp[3];
    int nd = get_int_array(env, a[0], shp, 3);
    int ns = get_double_array(env, a[1], sp, 3);
    const flux_field* w[3] = {NULL, NULL, NULL};
    int nw = get_field_array(env, a[5], w);
    if (nd < 0 || ns < 0 || nw < 0) {
        napi_throw_type_error(env, NULL, "multigrid: bad arguments");
        return NULL;
    }
    if (nd != ns) {
        char msg[160];
        snprintf(msg, sizeof msg, "multigrid: gridSpacing length (%d) must match shape length (%d).", ns, nd);
        napi_throw_error(env, NULL, msg);
        return NULL;
    }
    flux_mg_options m;
    mg_opts(env, a[2], &m);
    flux_multigrid* mg = NULL;
    CHECK_FLUX(flux_multigrid_create(nd, shp, sp, &m, is_nullish(env, a[3]) ? NULL : get_field(env, a[3]),
                                     is_nullish(env, a[4]) ? NULL : get_field(env, a[4]), nw ? w : NULL, &mg));
    return wrap(env, K_MG, mg);
}

static napi_value js_multigrid_apply(napi_env env, napi_callback_info info) {
    napi_value a[3];
    args_n(env, info, 3, a);
    CHECK_FLUX(flux_multigrid_apply((flux_multigrid*)unwrap(env, a[0], K_MG, K_MG), get_field(env, a[1]),
                                    get_field(env, a[2])));
    return undef(env);
}

static napi_value js_multigrid_refresh(napi_env env, napi_callback_info info) {
    napi_value a[1];
    args_n(env, info, 1, a);
    CHECK_FLUX(flux_multigrid_refresh((flux_multigrid*)unwrap(env, a[0], K_MG, K_MG)));
    return undef(env);
}

/* ----------------------------------------------------------------- PCG */
static const char* stop_name(int s) {
    switch (s) {
        case FLUX_STOP_DEGENERATE_PAP: return "degenerate-pAp";
        case FLUX_STOP_PAP_GROWTH: return "pAp-growth";
        case FLUX_STOP_ALPHA_MAGNITUDE: return "alpha-magnitude";
        case FLUX_STOP_DEGENERATE_OLD_RZ: return "degenerate-oldRZ";
        case FLUX_STOP_CONVERGED: return "converged";
        case FLUX_STOP_RESIDUAL_GROWING: return "residual-growing";
        default: return "none";
    }
}

static napi_value js_pcg_create(napi_env env, napi_callback_info info) {
    napi_value a[5];
    args_n(env, info, 5, a);
    flux_pcg* s = NULL;
    CHECK_FLUX(flux_pcg_create((flux_operator*)unwrap(env, a[0], K_OP, K_OP), get_int(env, a[1], 0),
                               is_nullish(env, a[2]) ? NULL : (flux_multigrid*)unwrap(env, a[2], K_MG, K_MG),
                               get_field(env, a[3]), get_field(env, a[4]), &s));
    return wrap(env, K_PCG, s);
}

static napi_value pcg_state_object(napi_env env, const flux_pcg_state* st);

static napi_value js_pcg_solve(napi_env env, napi_callback_info info) {
    napi_value a[4];
    args_n(env, info, 4, a);
    flux_pcg_state st;
    bool rel = false;
    napi_get_value_bool(env, a[3], &rel);
    CHECK_FLUX(flux_pcg_solve((flux_pcg*)unwrap(env, a[0], K_PCG, K_PCG), get_double(env, a[1], 1e-6),
                              get_int(env, a[2], 100), rel ? 1 : 0, &st));
    return pcg_state_object(env, &st);
}

static napi_value pcg_state_object(napi_env env, const flux_pcg_state* stp) {
    const flux_pcg_state st = *stp;
    napi_value o, s, b;
    napi_create_object(env, &o);
    napi_set_named_property(env, o, "iterations", num(env, st.iterations));
    napi_create_string_utf8(env, stop_name(st.stopped_by), NAPI_AUTO_LENGTH, &s);
    napi_set_named_property(env, o, "stoppedBy", s);
    napi_get_boolean(env, st.converged != 0, &b);
    napi_set_named_property(env, o, "converged", b);
    napi_set_named_property(env, o, "residualSquared", num(env, st.residual * st.residual));
    napi_set_named_property(env, o, "residual", num(env, st.residual));
    napi_set_named_property(env, o, "threshold", num(env, st.threshold));
    napi_set_named_property(env, o, "bNorm", num(env, st.b_norm));
    return o;
}

/* ------------------------------------------------------------ pressure */
static napi_value js_pressure3_create(napi_env env, napi_callback_info info) {
    napi_value a[1];
    args_n(env, info, 1, a);
    napi_value o = a[0];
    flux_pressure3_desc d;
    flux_pressure3_desc_default(&d);
    napi_value v;
    int tmp[3];
    double dt[3];
    if ((v = prop(env, o, "resolution")) && get_int_array(env, v, tmp, 3) == 3)
        for (int i = 0; i < 3; ++i) d.resolution[i] = tmp[i];
    if ((v = prop(env, o, "gridSpacing")) && get_double_array(env, v, dt, 3) == 3)
        for (int i = 0; i < 3; ++i) d.spacing[i] = dt[i];
    if ((v = prop(env, o, "origin")) && get_double_array(env, v, dt, 3) == 3)
        for (int i = 0; i < 3; ++i) d.origin[i] = dt[i];
    if ((v = prop(env, o, "dirichletActive")) && !is_nullish(env, v)) d.dirichlet_active = get_field(env, v);
    if ((v = prop(env, o, "dirichletTarget")) && !is_nullish(env, v)) d.dirichlet_target = get_field(env, v);
    if ((v = prop(env, o, "faceWeights")) && get_field_array(env, v, d.face_weights) < 0) {
        napi_throw_type_error(env, NULL, "createGridPressureSolver3: faceWeights must be three fields");
        return NULL;
    }
    if ((v = prop(env, o, "colliderWeights")) && get_field_array(env, v, d.collider_weights) < 0) {
        napi_throw_type_error(env, NULL, "createGridPressureSolver3: colliderWeights must be three fields");
        return NULL;
    }
    if ((v = prop(env, o, "multigrid"))) mg_opts(env, v, &d.multigrid);
    if ((v = prop(env, o, "preconditionerKind"))) d.preconditioner = get_int(env, v, d.preconditioner);
    if ((v = prop(env, o, "tolerance"))) d.tolerance = get_double(env, v, d.tolerance);
    if ((v = prop(env, o, "maxIterations")) && !is_nullish(env, v)) d.max_iterations = get_int(env, v, 0);
    if ((v = prop(env, o, "maxIterationsFactor"))) d.max_iterations_factor = get_double(env, v, d.max_iterations_factor);
    if ((v = prop(env, o, "relativeTolerance"))) {
        bool b = true;
        napi_get_value_bool(env, v, &b);
        d.relative_tolerance = b ? 1 : 0;
    }
    if ((v = prop(env, o, "maxPlausiblePressure"))) d.max_plausible_pressure = get_double(env, v, d.max_plausible_pressure);
    flux_pressure3* p = NULL;
    CHECK_FLUX(flux_pressure3_create(&d, &p));
    return wrap(env, K_P3, p);
}

static napi_value js_pressure3_field(napi_env env, napi_callback_info info) {
    napi_value a[2];
    args_n(env, info, 2, a);
    flux_field* f = flux_pressure3_field((flux_pressure3*)unwrap(env, a[0], K_P3, K_P3), get_int(env, a[1], 0));
    if (!f) {
        napi_value n;
        napi_get_null(env, &n);
        return n;
    }
    return wrap(env, K_FIELD_VIEW, f);
}

// pressure3SetCoarseSpace(p, freeIndicator, [fields]): the coarse-space
// correction of the projection's preconditioner (no reference counterpart)
static napi_value js_pressure3_set_coarse_space(napi_env env, napi_callback_info info) {
    napi_value a[3];
    args_n(env, info, 3, a);
    flux_pressure3* p = (flux_pressure3*)unwrap(env, a[0], K_P3, K_P3);
    bool is_array = false;
    uint32_t n = 0;
    const flux_field* f[3] = {NULL, NULL, NULL};
    if (napi_is_array(env, a[2], &is_array) == napi_ok && is_array) {
        napi_get_array_length(env, a[2], &n);
        if (n > 3) {
            napi_throw_error(env, NULL, "coarseSpace: at most 3 fields");
            return NULL;
        }
        for (uint32_t q = 0; q < n; ++q) {
            napi_value e;
            napi_get_element(env, a[2], q, &e);
            f[q] = get_field(env, e);
        }
    }
    CHECK_FLUX(flux_pressure3_set_coarse_space(p, get_int(env, a[1], 1), (int)n, f));
    return NULL;
}

// pressure3SetFlexible(p, on): the flexible (Polak-Ribiere) beta
static napi_value js_pressure3_set_flexible(napi_env env, napi_callback_info info) {
    napi_value a[2];
    args_n(env, info, 2, a);
    CHECK_FLUX(flux_pressure3_set_flexible((flux_pressure3*)unwrap(env, a[0], K_P3, K_P3), get_int(env, a[1], 1)));
    return NULL;
}

static napi_value diag_object(napi_env env, const flux_pressure_diag* d);

static napi_value js_pressure3_project(napi_env env, napi_callback_info info) {
    napi_value a[3];
    args_n(env, info, 3, a);
    const flux_field* in[3] = {NULL, NULL, NULL};
    const flux_field* out[3] = {NULL, NULL, NULL};
    if (get_field_array(env, a[1], in) != 3 || get_field_array(env, a[2], out) != 3) {
        napi_throw_type_error(env, NULL, "project: expected (solver, [u,v,w], [u,v,w])");
        return NULL;
    }
    flux_pressure_diag dg;
    CHECK_FLUX(flux_pressure3_project((flux_pressure3*)unwrap(env, a[0], K_P3, K_P3), in[0], in[1], in[2],
                                      (flux_field*)out[0], (flux_field*)out[1], (flux_field*)out[2], &dg));
    return diag_object(env, &dg);
}

static napi_value diag_object(napi_env env, const flux_pressure_diag* d) {
    const flux_pressure_diag dg = *d;
    napi_value o, s, b1, b2;
    napi_create_object(env, &o);
    napi_get_boolean(env, dg.converged != 0, &b1);
    napi_get_boolean(env, dg.rejected != 0, &b2);
    napi_set_named_property(env, o, "converged", b1);
    napi_set_named_property(env, o, "rejected", b2);
    napi_set_named_property(env, o, "iterations", num(env, dg.iterations));
    napi_create_string_utf8(env, stop_name(dg.stopped_by), NAPI_AUTO_LENGTH, &s);
    napi_set_named_property(env, o, "stoppedBy", s);
    napi_set_named_property(env, o, "residual", num(env, dg.residual));
    napi_set_named_property(env, o, "bNorm", num(env, dg.b_norm));
    napi_set_named_property(env, o, "badCells", num(env, (double)dg.bad_cells));
    return o;
}

/* ------------------------------------------------------------ transfers */
static int get_grid(napi_env env, napi_value o, flux_grid3* g) {
    int n[3];
    double h[3] = {1, 1, 1}, org[3] = {0, 0, 0};
    napi_value v;
    if (!(v = prop(env, o, "n")) || get_int_array(env, v, n, 3) != 3) return -1;
    if ((v = prop(env, o, "h"))) get_double_array(env, v, h, 3);
    if ((v = prop(env, o, "origin"))) get_double_array(env, v, org, 3);
    for (int i = 0; i < 3; ++i) {
        g->n[i] = n[i];
        g->h[i] = (float)h[i];
        g->origin[i] = (float)org[i];
    }
    return 0;
}

static napi_value js_bin_particles(napi_env env, napi_callback_info info) {
    napi_value a[6];
    args_n(env, info, 6, a);
    flux_grid3 g;
    if (get_grid(env, a[0], &g)) {
        napi_throw_type_error(env, NULL, "binParticles: bad grid");
        return NULL;
    }
    int64_t outside = 0;
    CHECK_FLUX(flux_bin_particles(&g, get_field(env, a[1]), (int64_t)get_double(env, a[2], 0), get_field(env, a[3]),
                                  get_field(env, a[4]), get_field(env, a[5]), &outside));
    return num(env, (double)outside);
}

static napi_value js_p2g_gather(napi_env env, napi_callback_info info) {
    napi_value a[9];
    args_n(env, info, 9, a);
    flux_grid3 g;
    if (get_grid(env, a[0], &g)) {
        napi_throw_type_error(env, NULL, "p2g: bad grid");
        return NULL;
    }
    CHECK_FLUX(flux_p2g_gather(&g, get_field(env, a[1]), get_field(env, a[2]), get_field(env, a[3]),
                               get_field(env, a[4]), get_int(env, a[5], 0), (float)get_double(env, a[6], 1e-4),
                               get_field(env, a[7]), get_field(env, a[8])));
    return undef(env);
}

static napi_value js_g2p(napi_env env, napi_callback_info info) {
    napi_value a[8];
    args_n(env, info, 8, a);
    flux_grid3 g;
    const flux_field* vn[3];
    const flux_field* vo[3];
    if (get_grid(env, a[0], &g) || get_field_array(env, a[4], vn) != 3 || get_field_array(env, a[5], vo) != 3) {
        napi_throw_type_error(env, NULL, "g2p: bad arguments");
        return NULL;
    }
    CHECK_FLUX(flux_g2p(&g, get_field(env, a[1]), get_field(env, a[2]), (int64_t)get_double(env, a[3], 0), vn, vo,
                        (float)get_double(env, a[6], 0.97), (float)get_double(env, a[7], 0.02)));
    return undef(env);
}

static napi_value js_advect_particles(napi_env env, napi_callback_info info) {
    napi_value a[6];
    args_n(env, info, 6, a);
    flux_grid3 g;
    const flux_field* v[3];
    if (get_grid(env, a[0], &g) || get_field_array(env, a[3], v) != 3) {
        napi_throw_type_error(env, NULL, "advectParticles: bad arguments");
        return NULL;
    }
    CHECK_FLUX(flux_advect_particles(&g, get_field(env, a[1]), (int64_t)get_double(env, a[2], 0), v,
                                     (float)get_double(env, a[4], 0.0), get_int(env, a[5], 32)));
    return undef(env);
}

static napi_value js_advect_faces(napi_env env, napi_callback_info info) {
    napi_value a[5];
    args_n(env, info, 5, a);
    flux_grid3 g;
    const flux_field* vi[3];
    const flux_field* vo[3];
    if (get_grid(env, a[0], &g) || get_field_array(env, a[1], vi) != 3 || get_field_array(env, a[2], vo) != 3) {
        napi_throw_type_error(env, NULL, "advectFaceCentered3: bad arguments");
        return NULL;
    }
    CHECK_FLUX(flux_advect_faces(&g, vi, (flux_field* const*)vo, (float)get_double(env, a[3], 0.0),
                                 get_int(env, a[4], 32)));
    return undef(env);
}

/* createSemiLagrangianAdvectionSolver3: (grid, [u, v, w], dt, maxSubsteps, order) */
static napi_value js_advection3_create(napi_env env, napi_callback_info info) {
    napi_value a[5];
    args_n(env, info, 5, a);
    flux_grid3 g;
    const flux_field* v[3];
    if (get_grid(env, a[0], &g) || get_field_array(env, a[1], v) != 3) {
        napi_throw_type_error(env, NULL, "createSemiLagrangianAdvectionSolver3: bad arguments");
        return NULL;
    }
    flux_advection3* adv = NULL;
    CHECK_FLUX(flux_advection3_create(&g, v, (float)get_double(env, a[2], 0.0), get_int(env, a[3], 32),
                                      get_int(env, a[4], 1), &adv));
    return wrap(env, K_ADV3, adv);
}

static napi_value js_advection3_set_dt(napi_env env, napi_callback_info info) {
    napi_value a[2];
    args_n(env, info, 2, a);
    CHECK_FLUX(flux_advection3_set_dt((flux_advection3*)unwrap(env, a[0], K_ADV3, K_ADV3),
                                      (float)get_double(env, a[1], 0.0)));
    return undef(env);
}

static napi_value js_advection3_set_collider(napi_env env, napi_callback_info info) {
    napi_value a[2];
    args_n(env, info, 2, a);
    CHECK_FLUX(flux_advection3_set_collider((flux_advection3*)unwrap(env, a[0], K_ADV3, K_ADV3),
                                            is_nullish(env, a[1]) ? NULL : get_field(env, a[1])));
    return undef(env);
}

static napi_value js_advection3_faces(napi_env env, napi_callback_info info) {
    napi_value a[3];
    args_n(env, info, 3, a);
    const flux_field* vi[3];
    const flux_field* vo[3];
    if (get_field_array(env, a[1], vi) != 3 || get_field_array(env, a[2], vo) != 3) {
        napi_throw_type_error(env, NULL, "advectFaceCentered3: bad arguments");
        return NULL;
    }
    CHECK_FLUX(flux_advection3_faces((flux_advection3*)unwrap(env, a[0], K_ADV3, K_ADV3), vi,
                                     (flux_field* const*)vo));
    return undef(env);
}

static napi_value js_advection3_scalar(napi_env env, napi_callback_info info) {
    napi_value a[3];
    args_n(env, info, 3, a);
    CHECK_FLUX(flux_advection3_scalar((flux_advection3*)unwrap(env, a[0], K_ADV3, K_ADV3), get_field(env, a[1]),
                                      get_field(env, a[2])));
    return undef(env);
}

static napi_value js_two_phase_beta(napi_env env, napi_callback_info info) {
    napi_value a[6];
    args_n(env, info, 6, a);
    flux_grid3 g;
    const flux_field* beta[3];
    if (get_grid(env, a[0], &g) || get_field_array(env, a[5], beta) != 3) {
        napi_throw_type_error(env, NULL, "twoPhaseBeta: bad arguments");
        return NULL;
    }
    CHECK_FLUX(flux_two_phase_beta(&g, is_nullish(env, a[1]) ? NULL : get_field(env, a[1]),
                                   (float)get_double(env, a[2], 1.0), (float)get_double(env, a[3], 1000.0),
                                   get_field(env, a[4]), (flux_field* const*)beta));
    return undef(env);
}

/* ------------------------------------------------------- FLIP grid stages */
static napi_value js_extrapolator_create(napi_env env, napi_callback_info info) {
    napi_value a[3];
    args_n(env, info, 3, a);
    flux_extrapolator* e = NULL;
    CHECK_FLUX(flux_extrapolator3_create(get_field(env, a[0]), get_field(env, a[1]), get_field(env, a[2]), &e));
    return wrap(env, K_EXTRAP, e);
}

static napi_value js_extrapolator_run(napi_env env, napi_callback_info info) {
    napi_value a[2];
    args_n(env, info, 2, a);
    CHECK_FLUX(flux_extrapolator3_run((flux_extrapolator*)unwrap(env, a[0], K_EXTRAP, K_EXTRAP), get_int(env, a[1], 5)));
    return undef(env);
}

static napi_value js_mark_fluid_cells(napi_env env, napi_callback_info info) {
    napi_value a[4];
    args_n(env, info, 4, a);
    flux_grid3 g;
    if (get_grid(env, a[0], &g)) {
        napi_throw_type_error(env, NULL, "markFluidCells: bad grid");
        return NULL;
    }
    CHECK_FLUX(flux_mark_fluid_cells(&g, get_field(env, a[1]), (int64_t)get_double(env, a[2], 0), get_field(env, a[3])));
    return undef(env);
}

static napi_value js_fluid_adjacent(napi_env env, napi_callback_info info) {
    napi_value a[4];
    args_n(env, info, 4, a);
    flux_grid3 g;
    if (get_grid(env, a[0], &g)) {
        napi_throw_type_error(env, NULL, "fluidAdjacent: bad grid");
        return NULL;
    }
    CHECK_FLUX(flux_fluid_adjacent(&g, get_field(env, a[1]), get_int(env, a[2], 0), get_field(env, a[3])));
    return undef(env);
}

static int get_grid3d(napi_env env, napi_value o, flux_grid3d* g) {
    napi_value v;
    if (!(v = prop(env, o, "n")) || get_int_array(env, v, g->n, 3) != 3) return -1;
    for (int i = 0; i < 3; ++i) {
        g->h[i] = 1.0;
        g->origin[i] = 0.0;
    }
    if ((v = prop(env, o, "h"))) get_double_array(env, v, g->h, 3);
    if ((v = prop(env, o, "origin"))) get_double_array(env, v, g->origin, 3);
    return 0;
}

/* shapes: [{type: 'sphere'|'box', center, radius, halfExtents, rotation: {axis, angle}}] */
static int get_shapes(napi_env env, napi_value arr, flux_primitive3** out, int* count) {
    uint32_t n = 0;
    if (napi_get_array_length(env, arr, &n) != napi_ok) return -1;
    flux_primitive3* p = (flux_primitive3*)calloc(n > 0 ? n : 1, sizeof(flux_primitive3));
    for (uint32_t i = 0; i < n; ++i) {
        napi_value e, v;
        napi_get_element(env, arr, i, &e);
        char type[16] = {0};
        size_t len = 0;
        if ((v = prop(env, e, "type"))) napi_get_value_string_utf8(env, v, type, sizeof(type), &len);
        p[i].type = strcmp(type, "sphere") == 0 ? 0 : (strcmp(type, "box") == 0 ? 1 : -1);
        if ((v = prop(env, e, "center"))) get_double_array(env, v, p[i].center, 3);
        if ((v = prop(env, e, "radius"))) p[i].radius = get_double(env, v, 0.0);
        if ((v = prop(env, e, "halfExtents"))) get_double_array(env, v, p[i].half, 3);
        if ((v = prop(env, e, "rotation")) && !is_nullish(env, v)) {
            napi_value r;
            p[i].rotated = 1;
            if ((r = prop(env, v, "axis"))) get_double_array(env, r, p[i].axis, 3);
            if ((r = prop(env, v, "angle"))) p[i].angle = get_double(env, r, 0.0);
        }
    }
    *out = p;
    *count = (int)n;
    return 0;
}

static napi_value js_collider_face_weights(napi_env env, napi_callback_info info) {
    napi_value a[4];
    args_n(env, info, 4, a);
    flux_grid3d g;
    const flux_field* w[3];
    flux_primitive3* sh = NULL;
    int count = 0;
    bool inv = false;
    napi_get_value_bool(env, a[2], &inv);
    if (get_grid3d(env, a[0], &g) || get_field_array(env, a[3], w) != 3 || get_shapes(env, a[1], &sh, &count)) {
        free(sh);
        napi_throw_type_error(env, NULL, "computeFaceWeights: bad arguments");
        return NULL;
    }
    int rc = flux_collider_face_weights(&g, sh, count, inv ? 1 : 0, (flux_field* const*)w);
    free(sh);
    CHECK_FLUX(rc);
    return undef(env);
}

static napi_value js_collider_sdf(napi_env env, napi_callback_info info) {
    napi_value a[4];
    args_n(env, info, 4, a);
    flux_grid3d g;
    flux_primitive3* sh = NULL;
    int count = 0;
    bool inv = false;
    napi_get_value_bool(env, a[2], &inv);
    if (get_grid3d(env, a[0], &g) || get_shapes(env, a[1], &sh, &count)) {
        free(sh);
        napi_throw_type_error(env, NULL, "addShapes: bad arguments");
        return NULL;
    }
    int rc = flux_collider_sdf(&g, sh, count, inv ? 1 : 0, get_field(env, a[3]));
    free(sh);
    CHECK_FLUX(rc);
    return undef(env);
}

static napi_value js_boundary3_create(napi_env env, napi_callback_info info) {
    napi_value a[5];
    args_n(env, info, 5, a);
    flux_grid3 g;
    const flux_field* vel[3];
    const flux_field* w[3] = {NULL, NULL, NULL};
    int nw = get_field_array(env, a[3], w);
    if (get_grid(env, a[0], &g) || get_field_array(env, a[1], vel) != 3 || (nw != 0 && nw != 3)) {
        napi_throw_type_error(env, NULL, "createGridBlockedBoundaryConditionSolver3: bad arguments");
        return NULL;
    }
    flux_boundary3* b = NULL;
    CHECK_FLUX(flux_boundary3_create(&g, (flux_field* const*)vel, is_nullish(env, a[2]) ? NULL : get_field(env, a[2]),
                                     nw == 3 ? w : NULL, (float)get_double(env, a[4], 0.5), &b));
    return wrap(env, K_BND, b);
}

// boundary3AddInflow(b, sdfField, [vx, vy, vz], mode): a createSDFInflow3 object
static napi_value js_boundary3_add_inflow(napi_env env, napi_callback_info info) {
    napi_value a[4];
    args_n(env, info, 4, a);
    double v[3] = {0.0, 0.0, 0.0};
    if (get_double_array(env, a[2], v, 3) != 3) {
        napi_throw_type_error(env, NULL, "setInflows: velocity must be [vx, vy, vz]");
        return NULL;
    }
    CHECK_FLUX(flux_boundary3_add_inflow((flux_boundary3*)unwrap(env, a[0], K_BND, K_BND), get_field(env, a[1]),
                                         (float)v[0], (float)v[1], (float)v[2], get_int(env, a[3], 0)));
    return undef(env);
}

// outflow3Dirichlet(grid, [sdf fields], userActive | null, userTarget | null, active, target)
static napi_value js_outflow3_dirichlet(napi_env env, napi_callback_info info) {
    napi_value a[6];
    args_n(env, info, 6, a);
    flux_grid3 g;
    const flux_field* o[4] = {NULL, NULL, NULL, NULL};
    const int no = get_field_array_n(env, a[1], o, 4);
    if (get_grid(env, a[0], &g) || no < 0) {
        napi_throw_type_error(env, NULL, "createOutflowPressureDirichlet3: bad arguments");
        return NULL;
    }
    CHECK_FLUX(flux_outflow3_dirichlet(&g, o, no, is_nullish(env, a[2]) ? NULL : get_field(env, a[2]),
                                       is_nullish(env, a[3]) ? NULL : get_field(env, a[3]), get_field(env, a[4]),
                                       get_field(env, a[5])));
    return undef(env);
}

static napi_value js_boundary3_set_closed(napi_env env, napi_callback_info info) {
    napi_value a[2];
    args_n(env, info, 2, a);
    CHECK_FLUX(flux_boundary3_set_closed((flux_boundary3*)unwrap(env, a[0], K_BND, K_BND), get_int(env, a[1], 63)));
    return undef(env);
}

static napi_value js_boundary3_constrain(napi_env env, napi_callback_info info) {
    napi_value a[2];
    args_n(env, info, 2, a);
    CHECK_FLUX(flux_boundary3_constrain((flux_boundary3*)unwrap(env, a[0], K_BND, K_BND), get_int(env, a[1], 5)));
    return undef(env);
}

/* createGridFlipSolver3 options object (grid_flip_solver3.js:135-163) */
static napi_value js_flip3_create(napi_env env, napi_callback_info info) {
    napi_value a[1];
    args_n(env, info, 1, a);
    napi_value o = a[0], v;
    flux_flip3_desc d;
    flux_flip3_desc_default(&d);
    int tmp[3];
    double dv[3];
    if ((v = prop(env, o, "resolution")) && get_int_array(env, v, tmp, 3) == 3)
        for (int i = 0; i < 3; ++i) d.resolution[i] = tmp[i];
    if ((v = prop(env, o, "gridSpacing")) && get_double_array(env, v, dv, 3) == 3)
        for (int i = 0; i < 3; ++i) d.spacing[i] = dv[i];
    if ((v = prop(env, o, "origin")) && get_double_array(env, v, dv, 3) == 3)
        for (int i = 0; i < 3; ++i) d.origin[i] = dv[i];
    if ((v = prop(env, o, "gravity")) && get_double_array(env, v, dv, 3) == 3)
        for (int i = 0; i < 3; ++i) d.gravity[i] = (float)dv[i];
    if ((v = prop(env, o, "maxParticles"))) d.max_particles = (int64_t)get_double(env, v, 0);
    if ((v = prop(env, o, "dt"))) d.dt = (float)get_double(env, v, 0);
    if ((v = prop(env, o, "reducedPressure"))) {
        bool b = true;
        napi_get_value_bool(env, v, &b);
        d.reduced_pressure = b ? 1 : 0;
    }
    if ((v = prop(env, o, "flipRatio"))) d.flip_ratio = (float)get_double(env, v, d.flip_ratio);
    if ((v = prop(env, o, "velocityDamping"))) d.velocity_damping = (float)get_double(env, v, d.velocity_damping);
    if ((v = prop(env, o, "p2gAtomicScale"))) d.p2g_atomic_scale = (float)get_double(env, v, d.p2g_atomic_scale);
    if ((v = prop(env, o, "weightEpsilon"))) d.weight_epsilon = (float)get_double(env, v, d.weight_epsilon);
    if ((v = prop(env, o, "closedDomainBoundaryFlag")) && !is_nullish(env, v))
        d.closed_domain_flags = get_int(env, v, d.closed_domain_flags);
    if ((v = prop(env, o, "resampleEnabled"))) {
        bool b = true;
        napi_get_value_bool(env, v, &b);
        d.resample = b ? 1 : 0;
    }
    if ((v = prop(env, o, "minParticlesPerCell"))) d.min_particles_per_cell = get_int(env, v, d.min_particles_per_cell);
    if ((v = prop(env, o, "maxParticlesPerCell"))) d.max_particles_per_cell = get_int(env, v, d.max_particles_per_cell);
    if ((v = prop(env, o, "colliderSdf")) && !is_nullish(env, v)) d.collider_sdf = get_field(env, v);
    if ((v = prop(env, o, "colliderWeights")) && get_field_array(env, v, d.collider_weights) < 0) {
        napi_throw_type_error(env, NULL, "createGridFlipSolver3: colliderWeights must be three fields");
        return NULL;
    }
    if ((v = prop(env, o, "friction"))) d.friction = (float)get_double(env, v, d.friction);
    if ((v = prop(env, o, "colliderPushThresh"))) d.collider_push_thresh = (float)get_double(env, v, 0.0);
    if ((v = prop(env, o, "colliderPushShift"))) d.collider_push_shift = (float)get_double(env, v, 0.0);
    if ((v = prop(env, o, "multigrid"))) mg_opts(env, v, &d.pressure.multigrid);
    if ((v = prop(env, o, "preconditionerKind"))) d.pressure.preconditioner = get_int(env, v, d.pressure.preconditioner);
    if ((v = prop(env, o, "tolerance"))) d.pressure.tolerance = get_double(env, v, d.pressure.tolerance);
    if ((v = prop(env, o, "maxIterations")) && !is_nullish(env, v)) d.pressure.max_iterations = get_int(env, v, 0);
    if ((v = prop(env, o, "maxPlausiblePressure")) && !is_nullish(env, v))
        d.pressure.max_plausible_pressure = get_double(env, v, 0.0);
    if ((v = prop(env, o, "carryConcentration"))) {
        bool b = false;
        napi_get_value_bool(env, v, &b);
        d.carry_concentration = b ? 1 : 0;
    }
    if ((v = prop(env, o, "mixing"))) d.mixing = (float)get_double(env, v, 0.0);
    if ((v = prop(env, o, "fade"))) d.fade = (float)get_double(env, v, 0.0);
    if ((v = prop(env, o, "concentrationAtomicScale")))
        d.concentration_atomic_scale = (float)get_double(env, v, d.concentration_atomic_scale);
    if ((v = prop(env, o, "ambientDensity")) && !is_nullish(env, v)) d.ambient_density = (float)get_double(env, v, 0.0);
    if ((v = prop(env, o, "componentDensity")) && !is_nullish(env, v))
        d.component_density = (float)get_double(env, v, 0.0);
    if ((v = prop(env, o, "massWeightedTransfer"))) {
        bool b = false;
        napi_get_value_bool(env, v, &b);
        d.mass_weighted_transfer = b ? 1 : 0;
    }
    if ((v = prop(env, o, "pressurePinMode"))) d.pressure_pin = get_int(env, v, 0);
    if ((v = prop(env, o, "pinCell")) && !is_nullish(env, v) && get_int_array(env, v, tmp, 3) == 3)
        for (int i = 0; i < 3; ++i) d.pin_cell[i] = tmp[i];
    if ((v = prop(env, o, "phaseAware"))) {
        bool b = false;
        napi_get_value_bool(env, v, &b);
        d.phase_resample = b ? 1 : 0;
    }
    flux_flip3* s = NULL;
    CHECK_FLUX(flux_flip3_create(&d, &s));
    return wrap(env, K_FLIP, s);
}

static napi_value js_flip3_step(napi_env env, napi_callback_info info) {
    napi_value a[1];
    args_n(env, info, 1, a);
    flux_pressure_diag dg;
    CHECK_FLUX(flux_flip3_step((flux_flip3*)unwrap(env, a[0], K_FLIP, K_FLIP), &dg));
    return diag_object(env, &dg);
}

// flip3SetDt(s, dt): the step's dt from the next step on (the reference's
// live dt node, written by createGridAdaptiveTimeStep3)
static napi_value js_flip3_set_dt(napi_env env, napi_callback_info info) {
    napi_value a[2];
    args_n(env, info, 2, a);
    CHECK_FLUX(flux_flip3_set_dt((flux_flip3*)unwrap(env, a[0], K_FLIP, K_FLIP), (float)get_double(env, a[1], 0.0)));
    return NULL;
}

// flip3SetCoarseSpace(s, freeIndicator): the coarse-space correction of the
// step's projection (flux_flip3_pressure; no reference counterpart)
static napi_value js_flip3_set_coarse_space(napi_env env, napi_callback_info info) {
    napi_value a[2];
    args_n(env, info, 2, a);
    flux_pressure3* p = flux_flip3_pressure((flux_flip3*)unwrap(env, a[0], K_FLIP, K_FLIP));
    CHECK_FLUX(flux_pressure3_set_coarse_space(p, get_int(env, a[1], 1), 0, NULL));
    return NULL;
}

static napi_value js_flip3_field(napi_env env, napi_callback_info info) {
    napi_value a[2];
    args_n(env, info, 2, a);
    flux_field* f = flux_flip3_field((flux_flip3*)unwrap(env, a[0], K_FLIP, K_FLIP), get_int(env, a[1], 0));
    if (!f) {
        napi_value n;
        napi_get_null(env, &n);
        return n;
    }
    return wrap(env, K_FIELD_VIEW, f);
}

/* createGridSmokeSolver3 options object (grid_smoke_solver3.js:25-41) */
static napi_value js_smoke3_create(napi_env env, napi_callback_info info) {
    napi_value a[1];
    args_n(env, info, 1, a);
    napi_value o = a[0], v;
    flux_smoke3_desc d;
    flux_smoke3_desc_default(&d);
    int tmp[3];
    double dv[3];
    if ((v = prop(env, o, "resolution")) && get_int_array(env, v, tmp, 3) == 3)
        for (int i = 0; i < 3; ++i) d.resolution[i] = tmp[i];
    if ((v = prop(env, o, "gridSpacing")) && get_double_array(env, v, dv, 3) == 3)
        for (int i = 0; i < 3; ++i) d.spacing[i] = dv[i];
    if ((v = prop(env, o, "origin")) && get_double_array(env, v, dv, 3) == 3)
        for (int i = 0; i < 3; ++i) d.origin[i] = dv[i];
    if ((v = prop(env, o, "up")) && get_double_array(env, v, dv, 3) == 3)
        for (int i = 0; i < 3; ++i) d.up[i] = (float)dv[i];
    if ((v = prop(env, o, "dt"))) d.dt = (float)get_double(env, v, 0);
    if ((v = prop(env, o, "buoyancySmokeDensityFactor")))
        d.buoyancy_smoke_density_factor = (float)get_double(env, v, d.buoyancy_smoke_density_factor);
    if ((v = prop(env, o, "buoyancyTemperatureFactor")))
        d.buoyancy_temperature_factor = (float)get_double(env, v, d.buoyancy_temperature_factor);
    if ((v = prop(env, o, "ambientTemperature"))) d.ambient_temperature = (float)get_double(env, v, 0);
    if ((v = prop(env, o, "smokeDecay"))) d.smoke_decay = (float)get_double(env, v, d.smoke_decay);
    if ((v = prop(env, o, "temperatureDecay"))) d.temperature_decay = (float)get_double(env, v, d.temperature_decay);
    if ((v = prop(env, o, "closedDomainBoundaryFlag")) && !is_nullish(env, v))
        d.closed_domain_flags = get_int(env, v, d.closed_domain_flags);
    if ((v = prop(env, o, "multigrid"))) mg_opts(env, v, &d.pressure.multigrid);
    if ((v = prop(env, o, "preconditionerKind"))) d.pressure.preconditioner = get_int(env, v, d.pressure.preconditioner);
    if ((v = prop(env, o, "tolerance"))) d.pressure.tolerance = get_double(env, v, d.pressure.tolerance);
    if ((v = prop(env, o, "maxIterations")) && !is_nullish(env, v)) d.pressure.max_iterations = get_int(env, v, 0);
    if ((v = prop(env, o, "advectionOrder"))) d.advection_order = get_int(env, v, 1);
    if ((v = prop(env, o, "maxSubsteps"))) d.max_substeps = get_int(env, v, d.max_substeps);
    flux_smoke3* s = NULL;
    CHECK_FLUX(flux_smoke3_create(&d, &s));
    return wrap(env, K_SMOKE, s);
}

static napi_value js_smoke3_step(napi_env env, napi_callback_info info) {
    napi_value a[1];
    args_n(env, info, 1, a);
    flux_pressure_diag dg;
    CHECK_FLUX(flux_smoke3_step((flux_smoke3*)unwrap(env, a[0], K_SMOKE, K_SMOKE), &dg));
    return diag_object(env, &dg);
}

static napi_value js_smoke3_field(napi_env env, napi_callback_info info) {
    napi_value a[2];
    args_n(env, info, 2, a);
    flux_field* f = flux_smoke3_field((flux_smoke3*)unwrap(env, a[0], K_SMOKE, K_SMOKE), get_int(env, a[1], 0));
    if (!f) {
        napi_value n;
        napi_get_null(env, &n);
        return n;
    }
    return wrap(env, K_FIELD_VIEW, f);
}

static napi_value js_smoke3_source(napi_env env, napi_callback_info info) {
    napi_value a[5];
    args_n(env, info, 5, a);
    double c[3] = {0, 0, 0};
    get_double_array(env, a[1], c, 3);
    float cf[3] = {(float)c[0], (float)c[1], (float)c[2]};
    CHECK_FLUX(flux_smoke3_source_sphere((flux_smoke3*)unwrap(env, a[0], K_SMOKE, K_SMOKE), cf,
                                         (float)get_double(env, a[2], 1.0), (float)get_double(env, a[3], 1.0),
                                         (float)get_double(env, a[4], 1.0)));
    return undef(env);
}

/* ---------------------------------------------------------------- 2-D */
static int get_grid2(napi_env env, napi_value o, flux_grid2* g) {
    napi_value v;
    int n[2];
    double d[2];
    if (!(v = prop(env, o, "resolution")) || get_int_array(env, v, n, 2) != 2) return -1;
    g->n[0] = n[0];
    g->n[1] = n[1];
    g->h[0] = g->h[1] = 1.0f;
    g->origin[0] = g->origin[1] = 0.0f;
    if ((v = prop(env, o, "gridSpacing")) && get_double_array(env, v, d, 2) == 2) {
        g->h[0] = (float)d[0];
        g->h[1] = (float)d[1];
    }
    if ((v = prop(env, o, "origin")) && get_double_array(env, v, d, 2) == 2) {
        g->origin[0] = (float)d[0];
        g->origin[1] = (float)d[1];
    }
    return 0;
}

/* addPolygons(field, grid, polygons = [[[x, y], ...], ...], invert) */
static napi_value js_polygon_sdf2(napi_env env, napi_callback_info info) {
    napi_value a[4];
    args_n(env, info, 4, a);
    flux_grid2 g;
    uint32_t rings = 0;
    if (get_grid2(env, a[1], &g) || napi_get_array_length(env, a[2], &rings) != napi_ok || rings == 0) {
        napi_throw_type_error(env, NULL, "addPolygons: bad arguments");
        return NULL;
    }
    int* starts = (int*)calloc(rings + 1, sizeof(int));
    uint32_t total = 0;
    for (uint32_t r = 0; r < rings; ++r) {
        napi_value ring;
        uint32_t len = 0;
        napi_get_element(env, a[2], r, &ring);
        napi_get_array_length(env, ring, &len);
        total += len;
        starts[r + 1] = (int)total;
    }
    double* xy = (double*)calloc(2 * (size_t)(total > 0 ? total : 1), sizeof(double));
    for (uint32_t r = 0; r < rings; ++r) {
        napi_value ring;
        napi_get_element(env, a[2], r, &ring);
        for (int q = starts[r]; q < starts[r + 1]; ++q) {
            napi_value pt;
            napi_get_element(env, ring, (uint32_t)(q - starts[r]), &pt);
            get_double_array(env, pt, xy + 2 * q, 2);
        }
    }
    bool inv = false;
    napi_get_value_bool(env, a[3], &inv);
    int rc = flux_polygon_sdf2(get_field(env, a[0]), &g, xy, starts, (int)rings, inv ? 1 : 0);
    free(xy);
    free(starts);
    CHECK_FLUX(rc);
    return undef(env);
}

/* createMaxAbsReducer(fields).read() */
static napi_value js_max_abs(napi_env env, napi_callback_info info) {
    napi_value a[1];
    args_n(env, info, 1, a);
    const flux_field* f[8];
    int n = get_field_array_n(env, a[0], f, 8);
    if (n <= 0 || n > 8) {
        napi_throw_type_error(env, NULL, "createMaxAbsReducer: expected 1 to 8 fields");
        return NULL;
    }
    float m = 0.0f;
    CHECK_FLUX(flux_max_abs(f, n, &m));
    napi_value out;
    napi_create_double(env, (double)m, &out);
    return out;
}

/* computeAdaptiveSubSteps(maxVelocity, minGridSpacing, frameDt, courantNumber, maxSubSteps) */
static napi_value js_adaptive_substeps(napi_env env, napi_callback_info info) {
    napi_value a[5];
    args_n(env, info, 5, a);
    int steps = 1;
    double sub = 0.0;
    CHECK_FLUX(flux_adaptive_substeps(get_double(env, a[0], 0.0), get_double(env, a[1], 1.0), get_double(env, a[2], 0.0),
                                      get_double(env, a[3], 5.0), get_int(env, a[4], 32), &steps, &sub));
    napi_value out, v;
    napi_create_object(env, &out);
    napi_create_int32(env, steps, &v);
    napi_set_named_property(env, out, "numSubSteps", v);
    napi_create_double(env, sub, &v);
    napi_set_named_property(env, out, "subDt", v);
    return out;
}

/* createGridSolver2 options (grid_solver2.js:131-159) */
static napi_value js_solver2_create(napi_env env, napi_callback_info info) {
    napi_value a[1];
    args_n(env, info, 1, a);
    napi_value o = a[0], v;
    flux_solver2_desc d;
    flux_solver2_desc_default(&d);
    if (get_grid2(env, o, &d.grid)) {
        napi_throw_type_error(env, NULL, "createGridSolver2: resolution must be [nx, ny]");
        return NULL;
    }
    if ((v = prop(env, o, "colliderSdf")) && !is_nullish(env, v)) d.collider_sdf = get_field(env, v);
    if ((v = prop(env, o, "friction"))) d.friction = (float)get_double(env, v, d.friction);
    if ((v = prop(env, o, "advectCollider"))) {
        bool b = false;
        napi_get_value_bool(env, v, &b);
        d.advect_collider = b ? 1 : 0;
    }
    if ((v = prop(env, o, "maxSubsteps"))) d.max_substeps = get_int(env, v, d.max_substeps);
    if ((v = prop(env, o, "advectionOrder"))) d.advection_order = get_int(env, v, 1);
    if ((v = prop(env, o, "inflows")) && !is_nullish(env, v)) {
        uint32_t n = 0;
        napi_get_array_length(env, v, &n);
        if (n > 4) {
            napi_throw_error(env, NULL, "createGridSolver2: at most 4 inflows");
            return NULL;
        }
        d.n_inflows = (int)n;
        for (uint32_t q = 0; q < n; ++q) {
            napi_value e, w;
            double vel[2] = {0.0, 0.0};
            napi_get_element(env, v, q, &e);
            if ((w = prop(env, e, "sdf"))) d.inflow_sdf[q] = get_field(env, w);
            if ((w = prop(env, e, "velocity"))) get_double_array(env, w, vel, 2);
            d.inflow_velocity[q][0] = (float)vel[0];
            d.inflow_velocity[q][1] = (float)vel[1];
            if ((w = prop(env, e, "mode"))) d.inflow_mode[q] = get_int(env, w, 0);
        }
    }
    if ((v = prop(env, o, "outflows")) && !is_nullish(env, v)) {
        int n = get_field_array_n(env, v, d.outflow_sdf, 4);
        if (n < 0) {
            napi_throw_error(env, NULL, "createGridSolver2: at most 4 outflows");
            return NULL;
        }
        d.n_outflows = n;
    }
    if ((v = prop(env, o, "outflowVelocityBC"))) {
        bool b = true;
        napi_get_value_bool(env, v, &b);
        d.outflow_velocity_bc = b ? 1 : 0;
    }
    if ((v = prop(env, o, "dirichletActive")) && !is_nullish(env, v)) d.dirichlet_active = get_field(env, v);
    if ((v = prop(env, o, "dirichletTarget")) && !is_nullish(env, v)) d.dirichlet_target = get_field(env, v);
    if ((v = prop(env, o, "closedDomainBoundaryFlag")) && !is_nullish(env, v))
        d.closed_flags = get_int(env, v, d.closed_flags);
    if ((v = prop(env, o, "multigrid"))) mg_opts(env, v, &d.multigrid);
    if ((v = prop(env, o, "preconditionerKind"))) d.preconditioner = get_int(env, v, d.preconditioner);
    if ((v = prop(env, o, "tolerance"))) d.tolerance = get_double(env, v, d.tolerance);
    if ((v = prop(env, o, "maxIterations")) && !is_nullish(env, v)) d.max_iterations = get_int(env, v, 0);
    flux_solver2* s = NULL;
    CHECK_FLUX(flux_solver2_create(&d, &s));
    return wrap(env, K_S2, s);
}

static napi_value js_solver2_step(napi_env env, napi_callback_info info) {
    napi_value a[2];
    args_n(env, info, 2, a);
    flux_pressure_diag dg;
    CHECK_FLUX(flux_solver2_step((flux_solver2*)unwrap(env, a[0], K_S2, K_S2), (float)get_double(env, a[1], 0.0), &dg));
    return diag_object(env, &dg);
}

static napi_value js_solver2_field(napi_env env, napi_callback_info info) {
    napi_value a[2];
    args_n(env, info, 2, a);
    flux_field* f = flux_solver2_field((flux_solver2*)unwrap(env, a[0], K_S2, K_S2), get_int(env, a[1], 0));
    if (!f) {
        napi_value n;
        napi_get_null(env, &n);
        return n;
    }
    return wrap(env, K_FIELD_VIEW, f);
}


/* ------------------------------------------------------------ async work */
/* The reference's solver calls are async (project() returns an async
 * dispatch, solve() and onAdvanceTimeStep() are awaited,
 * grid_pressure_solver3.js:502-556): here the blocking library call runs on
 * a libuv worker thread (napi_async_work) and the returned Promise settles
 * on the JS thread, so the event loop keeps running meanwhile.  Library
 * calls must not interleave (one stream, linalg.js:1793-1797): one mutex
 * serialises every worker job. */
enum { JOB_PCG, JOB_P3, JOB_FLIP, JOB_SMOKE, JOB_S2, JOB_DOT, JOB_CG };
typedef struct {
    napi_async_work work;
    napi_deferred deferred;
    napi_ref keep; /* the JS values the job reads, alive until it settles */
    int kind;
    void* obj;
    double tol;
    int maxiter, rel, converged, iterations;
    float dt;
    const flux_field* in[3];
    flux_field* out[3];
    int rc;
    char err[512];
    flux_pcg_state st;
    flux_pressure_diag dg;
    double value;
} job_t;

static pthread_mutex_t g_job_mutex = PTHREAD_MUTEX_INITIALIZER;

static void job_execute(napi_env env, void* data) {
    (void)env;
    job_t* j = (job_t*)data;
    pthread_mutex_lock(&g_job_mutex);
    j->rc = flux_thread_bind();
    if (j->rc == FLUX_OK) {
        switch (j->kind) {
            case JOB_PCG: j->rc = flux_pcg_solve((flux_pcg*)j->obj, j->tol, j->maxiter, j->rel, &j->st); break;
            case JOB_P3:
                j->rc = flux_pressure3_project((flux_pressure3*)j->obj, j->in[0], j->in[1], j->in[2], j->out[0],
                                               j->out[1], j->out[2], &j->dg);
                break;
            case JOB_FLIP: j->rc = flux_flip3_step((flux_flip3*)j->obj, &j->dg); break;
            case JOB_SMOKE: j->rc = flux_smoke3_step((flux_smoke3*)j->obj, &j->dg); break;
            case JOB_S2: j->rc = flux_solver2_step((flux_solver2*)j->obj, j->dt, &j->dg); break;
            case JOB_DOT: j->rc = flux_dot_read((flux_dot*)j->obj, &j->value); break;
            case JOB_CG: j->rc = flux_cg_solve((flux_cg*)j->obj, j->tol, j->maxiter, &j->converged, &j->iterations); break;
            default: j->rc = FLUX_ERR_ARG; break;
        }
    }
    if (j->rc != FLUX_OK) snprintf(j->err, sizeof(j->err), "%s", flux_last_error());
    pthread_mutex_unlock(&g_job_mutex);
}

static napi_value pcg_state_object(napi_env env, const flux_pcg_state* st);

static void job_complete(napi_env env, napi_status status, void* data) {
    job_t* j = (job_t*)data;
    if (status != napi_ok || j->rc != FLUX_OK) {
        napi_value msg, err;
        napi_create_string_utf8(env, status != napi_ok ? "fluxflow addon: async work failed" : j->err,
                                NAPI_AUTO_LENGTH, &msg);
        napi_create_error(env, NULL, msg, &err);
        napi_reject_deferred(env, j->deferred, err);
    } else {
        napi_value v = NULL;
        switch (j->kind) {
            case JOB_PCG: v = pcg_state_object(env, &j->st); break;
            case JOB_DOT: napi_create_double(env, j->value, &v); break;
            case JOB_CG: {
                napi_value b;
                napi_create_object(env, &v);
                napi_get_boolean(env, j->converged != 0, &b);
                napi_set_named_property(env, v, "converged", b);
                napi_set_named_property(env, v, "iterations", num(env, j->iterations));
                break;
            }
            default: v = diag_object(env, &j->dg); break;
        }
        napi_resolve_deferred(env, j->deferred, v);
    }
    if (j->keep) napi_delete_reference(env, j->keep);
    napi_delete_async_work(env, j->work);
    free(j);
}

/* queues `j` and returns its Promise; `keep` (may be NULL) stays referenced
 * until the job settles */
static napi_value job_start(napi_env env, job_t* j, napi_value keep) {
    napi_value promise, name;
    if (napi_create_promise(env, &j->deferred, &promise) != napi_ok) {
        free(j);
        napi_throw_error(env, NULL, "fluxflow addon: napi_create_promise failed");
        return NULL;
    }
    if (keep) napi_create_reference(env, keep, 1, &j->keep);
    napi_create_string_utf8(env, "fluxflow", NAPI_AUTO_LENGTH, &name);
    napi_create_async_work(env, NULL, name, job_execute, job_complete, j, &j->work);
    napi_queue_async_work(env, j->work);
    return promise;
}

static job_t* job_new(int kind, void* obj) {
    job_t* j = (job_t*)calloc(1, sizeof(job_t));
    j->kind = kind;
    j->obj = obj;
    return j;
}

static napi_value js_pcg_solve_async(napi_env env, napi_callback_info info) {
    napi_value a[4];
    args_n(env, info, 4, a);
    flux_pcg* s = (flux_pcg*)unwrap(env, a[0], K_PCG, K_PCG);
    if (!s) {
        napi_throw_type_error(env, NULL, "solve: not a solver");
        return NULL;
    }
    job_t* j = job_new(JOB_PCG, s);
    j->tol = get_double(env, a[1], 1e-6);
    j->maxiter = get_int(env, a[2], 100);
    bool rel = false;
    napi_get_value_bool(env, a[3], &rel);
    j->rel = rel ? 1 : 0;
    return job_start(env, j, a[0]);
}

static napi_value js_pressure3_project_async(napi_env env, napi_callback_info info) {
    napi_value a[3];
    args_n(env, info, 3, a);
    flux_pressure3* p = (flux_pressure3*)unwrap(env, a[0], K_P3, K_P3);
    job_t* j = job_new(JOB_P3, p);
    const flux_field* out[3] = {NULL, NULL, NULL};
    if (!p || get_field_array(env, a[1], j->in) != 3 || get_field_array(env, a[2], out) != 3) {
        free(j);
        napi_throw_type_error(env, NULL, "project: expected (solver, [u,v,w], [u,v,w])");
        return NULL;
    }
    for (int i = 0; i < 3; ++i) j->out[i] = (flux_field*)out[i];
    napi_value keep;
    napi_create_array_with_length(env, 3, &keep);
    napi_set_element(env, keep, 0, a[0]);
    napi_set_element(env, keep, 1, a[1]);
    napi_set_element(env, keep, 2, a[2]);
    return job_start(env, j, keep);
}

static napi_value step_async(napi_env env, napi_callback_info info, int kind, int hkind) {
    napi_value a[2];
    args_n(env, info, 2, a);
    void* h = unwrap(env, a[0], hkind, hkind);
    if (!h) {
        napi_throw_type_error(env, NULL, "onAdvanceTimeStep: not a solver");
        return NULL;
    }
    job_t* j = job_new(kind, h);
    if (kind == JOB_S2) j->dt = (float)get_double(env, a[1], 0.0);
    return job_start(env, j, a[0]);
}

static napi_value js_flip3_step_async(napi_env env, napi_callback_info info) {
    return step_async(env, info, JOB_FLIP, K_FLIP);
}
static napi_value js_smoke3_step_async(napi_env env, napi_callback_info info) {
    return step_async(env, info, JOB_SMOKE, K_SMOKE);
}
static napi_value js_solver2_step_async(napi_env env, napi_callback_info info) {
    return step_async(env, info, JOB_S2, K_S2);
}

/* ------------------------------------------------- dot reducer and CG */
static napi_value js_dot_create(napi_env env, napi_callback_info info) {
    napi_value a[3];
    args_n(env, info, 3, a);
    int shp[3];
    int nd = get_int_array(env, a[0], shp, 3);
    flux_dot* d = NULL;
    CHECK_FLUX(flux_dot_create(nd, shp, get_field(env, a[1]), get_field(env, a[2]), &d));
    return wrap(env, K_DOT, d);
}

static napi_value js_dot_read_async(napi_env env, napi_callback_info info) {
    napi_value a[1];
    args_n(env, info, 1, a);
    flux_dot* d = (flux_dot*)unwrap(env, a[0], K_DOT, K_DOT);
    if (!d) {
        napi_throw_type_error(env, NULL, "read: not a dot reducer");
        return NULL;
    }
    return job_start(env, job_new(JOB_DOT, d), a[0]);
}

static napi_value js_dot_dispatch(napi_env env, napi_callback_info info) {
    napi_value a[1];
    args_n(env, info, 1, a);
    CHECK_FLUX(flux_dot_dispatch((flux_dot*)unwrap(env, a[0], K_DOT, K_DOT)));
    return undef(env);
}

static napi_value js_dot_info(napi_env env, napi_callback_info info) {
    napi_value a[1];
    args_n(env, info, 1, a);
    int lanes = 0;
    flux_field* part = NULL;
    CHECK_FLUX(flux_dot_info((flux_dot*)unwrap(env, a[0], K_DOT, K_DOT), &lanes, &part));
    napi_value o;
    napi_create_object(env, &o);
    napi_set_named_property(env, o, "lanes", num(env, lanes));
    napi_set_named_property(env, o, "partial", wrap(env, K_FIELD_VIEW, part));
    return o;
}

static napi_value js_cg_create(napi_env env, napi_callback_info info) {
    napi_value a[4];
    args_n(env, info, 4, a);
    flux_cg* s = NULL;
    flux_operator* op = is_nullish(env, a[0]) ? NULL : (flux_operator*)unwrap(env, a[0], K_OP, K_OP);
    const flux_field* diag = is_nullish(env, a[1]) ? NULL : get_field(env, a[1]);
    CHECK_FLUX(flux_cg_create(op, diag, get_field(env, a[2]), get_field(env, a[3]), &s));
    return wrap(env, K_CG, s);
}

static napi_value js_cg_solve_async(napi_env env, napi_callback_info info) {
    napi_value a[3];
    args_n(env, info, 3, a);
    flux_cg* s = (flux_cg*)unwrap(env, a[0], K_CG, K_CG);
    if (!s) {
        napi_throw_type_error(env, NULL, "solve: not a CG solver");
        return NULL;
    }
    job_t* j = job_new(JOB_CG, s);
    j->tol = get_double(env, a[1], 1e-6);
    j->maxiter = get_int(env, a[2], 100);
    return job_start(env, j, a[0]);
}

static napi_value js_cg_fields(napi_env env, napi_callback_info info) {
    napi_value a[1];
    args_n(env, info, 1, a);
    flux_field *p = NULL, *r = NULL, *ap = NULL, *ax = NULL;
    CHECK_FLUX(flux_cg_fields((flux_cg*)unwrap(env, a[0], K_CG, K_CG), &p, &r, &ap, &ax));
    napi_value o;
    napi_create_object(env, &o);
    napi_set_named_property(env, o, "p", wrap(env, K_FIELD_VIEW, p));
    napi_set_named_property(env, o, "r", wrap(env, K_FIELD_VIEW, r));
    napi_set_named_property(env, o, "Ap", wrap(env, K_FIELD_VIEW, ap));
    napi_set_named_property(env, o, "Ax", wrap(env, K_FIELD_VIEW, ax));
    return o;
}

#define EXPORT(name, fn)                                                   \
    do {                                                                   \
        napi_value f_;                                                     \
        napi_create_function(env, name, NAPI_AUTO_LENGTH, fn, NULL, &f_);  \
        napi_set_named_property(env, exports, name, f_);                   \
    } while (0)

static napi_value module_init(napi_env env, napi_value exports) {
    EXPORT("init", js_init);
    EXPORT("sync", js_sync);
    EXPORT("deviceInfo", js_device_info);
    EXPORT("lastError", js_last_error);
    EXPORT("fieldCreate", js_field_create);
    EXPORT("fieldUpload", js_field_upload);
    EXPORT("fieldDownload", js_field_download);
    EXPORT("fieldFill", js_field_fill);
    EXPORT("fieldCopy", js_field_copy);
    EXPORT("operatorCreate", js_operator_create);
    EXPORT("operatorApply", js_operator_apply);
    EXPORT("multigridCreate", js_multigrid_create);
    EXPORT("multigridApply", js_multigrid_apply);
    EXPORT("multigridRefresh", js_multigrid_refresh);
    EXPORT("pcgCreate", js_pcg_create);
    EXPORT("pcgSolve", js_pcg_solve);
    EXPORT("pcgSolveAsync", js_pcg_solve_async);
    EXPORT("pressure3ProjectAsync", js_pressure3_project_async);
    EXPORT("flip3StepAsync", js_flip3_step_async);
    EXPORT("smoke3StepAsync", js_smoke3_step_async);
    EXPORT("solver2StepAsync", js_solver2_step_async);
    EXPORT("dotCreate", js_dot_create);
    EXPORT("dotReadAsync", js_dot_read_async);
    EXPORT("dotDispatch", js_dot_dispatch);
    EXPORT("dotInfo", js_dot_info);
    EXPORT("cgCreate", js_cg_create);
    EXPORT("cgSolveAsync", js_cg_solve_async);
    EXPORT("cgFields", js_cg_fields);
    EXPORT("pressure3Create", js_pressure3_create);
    EXPORT("pressure3Field", js_pressure3_field);
    EXPORT("pressure3SetCoarseSpace", js_pressure3_set_coarse_space);
    EXPORT("pressure3SetFlexible", js_pressure3_set_flexible);
    EXPORT("pressure3Project", js_pressure3_project);
    EXPORT("binParticles", js_bin_particles);
    EXPORT("p2gGather", js_p2g_gather);
    EXPORT("g2p", js_g2p);
    EXPORT("advectParticles", js_advect_particles);
    EXPORT("advectFaces", js_advect_faces);
    EXPORT("advection3Create", js_advection3_create);
    EXPORT("advection3SetDt", js_advection3_set_dt);
    EXPORT("advection3Faces", js_advection3_faces);
    EXPORT("advection3Scalar", js_advection3_scalar);
    EXPORT("twoPhaseBeta", js_two_phase_beta);
    EXPORT("extrapolatorCreate", js_extrapolator_create);
    EXPORT("extrapolatorRun", js_extrapolator_run);
    EXPORT("markFluidCells", js_mark_fluid_cells);
    EXPORT("fluidAdjacent", js_fluid_adjacent);
    EXPORT("colliderFaceWeights", js_collider_face_weights);
    EXPORT("colliderSdf", js_collider_sdf);
    EXPORT("boundary3Create", js_boundary3_create);
    EXPORT("boundary3SetClosed", js_boundary3_set_closed);
    EXPORT("boundary3AddInflow", js_boundary3_add_inflow);
    EXPORT("outflow3Dirichlet", js_outflow3_dirichlet);
    EXPORT("advection3SetCollider", js_advection3_set_collider);
    EXPORT("boundary3Constrain", js_boundary3_constrain);
    EXPORT("flip3Create", js_flip3_create);
    EXPORT("flip3Step", js_flip3_step);
    EXPORT("flip3Field", js_flip3_field);
    EXPORT("flip3SetDt", js_flip3_set_dt);
    EXPORT("flip3SetCoarseSpace", js_flip3_set_coarse_space);
    EXPORT("smoke3Create", js_smoke3_create);
    EXPORT("smoke3Step", js_smoke3_step);
    EXPORT("smoke3Field", js_smoke3_field);
    EXPORT("smoke3Source", js_smoke3_source);
    EXPORT("polygonSdf2", js_polygon_sdf2);
    EXPORT("maxAbs", js_max_abs);
    EXPORT("adaptiveSubSteps", js_adaptive_substeps);
    EXPORT("solver2Create", js_solver2_create);
    EXPORT("solver2Step", js_solver2_step);
    EXPORT("solver2Field", js_solver2_field);
    return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, module_init)
