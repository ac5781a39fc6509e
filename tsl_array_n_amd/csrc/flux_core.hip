// Context, fields, timers: the tsl_array_n side of the boundary
// (packages/tsl_array_n/src/context.js, array.js).
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "flux_internal.h"

namespace flux {

hipStream_t g_stream = nullptr;
bool g_init = false;
int g_device = 0;
int g_cus = 0;
unsigned g_march_blocks = 4096;
bool g_march_flip = true;
unsigned g_march_toggle = 0;

static thread_local std::string t_err;

int set_error(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    t_err = buf;
    return code;
}

void clear_error() { t_err.clear(); }

size_t dtype_size(int dtype) {
    switch (dtype) {
        case FLUX_F32: return 4;
        case FLUX_I32: return 4;
        case FLUX_U8: return 1;
        case FLUX_F64: return 8;
        case FLUX_F16: return 2;
        case FLUX_BF16: return 2;
        case FLUX_VEC3: return 12;
        default: return 0;
    }
}

void* dmalloc(size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 16) != hipSuccess) return nullptr;
    return p;
}

void dfree(void* p) {
    if (p) (void)hipFree(p);
}

int check_field(const flux_field* f, int dtype, int64_t count, const char* what) {
    if (!f) return set_error(FLUX_ERR_ARG, "%s: field is NULL", what);
    if (f->dtype != dtype) return set_error(FLUX_ERR_ARG, "%s: element type mismatch (%d != %d)", what, f->dtype, dtype);
    if (count >= 0 && f->count != count)
        return set_error(FLUX_ERR_ARG, "%s: shape mismatch (%lld elements, expected %lld)", what, (long long)f->count,
                         (long long)count);
    return FLUX_OK;
}

// ---------------------------------------------------------------- kprof
struct KProf {
    std::string name;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pool;
    size_t used = 0;
    bool open = false;
};
static KProf g_kprof;

bool kprof_on(const char* name) { return !g_kprof.name.empty() && g_kprof.name == name; }

void kprof_begin(const char* name) {
    if (!kprof_on(name)) return;
    if (g_kprof.used == g_kprof.pool.size()) {
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess) return;
        if (hipEventCreate(&b) != hipSuccess) return;
        g_kprof.pool.push_back({a, b});
    }
    auto& pr = g_kprof.pool[g_kprof.used];
    (void)hipEventRecord(pr.first, g_stream);
    g_kprof.open = true;
}

void kprof_end(const char* name) {
    if (!kprof_on(name) || !g_kprof.open) return;
    auto& pr = g_kprof.pool[g_kprof.used++];
    (void)hipEventRecord(pr.second, g_stream);
    g_kprof.open = false;
}

}  // namespace flux

using namespace flux;

extern "C" {

const char* flux_last_error(void) { return t_err.c_str(); }

int flux_init(int device) {
    clear_error();
    if (g_init) {
        if (device == g_device) return FLUX_OK;
        return set_error(FLUX_ERR_STATE, "flux_init: already initialised on device %d", g_device);
    }
    int n = 0;
    FLUX_HIP(hipGetDeviceCount(&n));
    FLUX_REQUIRE(device >= 0 && device < n, "flux_init: device %d out of range (%d devices)", device, n);
    FLUX_HIP(hipSetDevice(device));
    // default priority stream (never a prioritised one)
    FLUX_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    hipDeviceProp_t prop;
    FLUX_HIP(hipGetDeviceProperties(&prop, device));
    g_cus = prop.multiProcessorCount;
    g_march_blocks = 4096;
    if (const char* e = getenv("FLUX_MARCH_BLOCKS")) {
        long v = strtol(e, nullptr, 10);
        if (v >= 64 && v <= (1 << 20)) g_march_blocks = (unsigned)v;
    }
    g_march_flip = true;
    if (const char* e = getenv("FLUX_MARCH_FLIP")) g_march_flip = strtol(e, nullptr, 10) != 0;
    g_device = device;
    g_init = true;
    return FLUX_OK;
}

int flux_shutdown(void) {
    if (!g_init) return FLUX_OK;
    (void)hipStreamSynchronize(g_stream);
    for (auto& pr : g_kprof.pool) {
        (void)hipEventDestroy(pr.first);
        (void)hipEventDestroy(pr.second);
    }
    g_kprof = KProf();
    (void)hipStreamDestroy(g_stream);
    g_stream = nullptr;
    g_init = false;
    return FLUX_OK;
}

int flux_thread_bind(void) {
    FLUX_REQUIRE_INIT();
    FLUX_HIP(hipSetDevice(g_device));
    return FLUX_OK;
}

int flux_sync(void) {
    FLUX_REQUIRE_INIT();
    FLUX_HIP(hipStreamSynchronize(g_stream));
    return FLUX_OK;
}

int flux_device_info(char* name, size_t len, int* cus) {
    FLUX_REQUIRE_INIT();
    hipDeviceProp_t prop;
    FLUX_HIP(hipGetDeviceProperties(&prop, g_device));
    if (name && len) {
        snprintf(name, len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (cus) *cus = prop.multiProcessorCount;
    return FLUX_OK;
}

int flux_field_create(int dtype, int ndim, const int64_t* shape, flux_field** out) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(out, "arrayN: out is NULL");
    FLUX_REQUIRE(dtype_size(dtype) > 0, "arrayN: unknown element type %d", dtype);
    FLUX_REQUIRE(ndim >= 1 && ndim <= 3, "arrayN: only 1D/2D/3D shapes are supported, got %dD", ndim);
    flux_field* f = new flux_field();
    f->dtype = dtype;
    f->ndim = ndim;
    int64_t count = 1;
    for (int a = 0; a < ndim; ++a) {
        if (shape[a] <= 0) {
            delete f;
            return set_error(FLUX_ERR_ARG, "arrayN: shape[%d] = %lld must be positive", a, (long long)shape[a]);
        }
        f->shape[a] = shape[a];
        count *= shape[a];
    }
    f->count = count;
    f->bytes = (size_t)count * dtype_size(dtype);
    f->d = dmalloc(f->bytes);
    if (!f->d) {
        delete f;
        return set_error(FLUX_ERR_NOMEM, "arrayN: hipMalloc of %zu bytes failed", (size_t)count * dtype_size(dtype));
    }
    // fields start zeroed, like instancedArray
    if (hipMemsetAsync(f->d, 0, f->bytes, g_stream) != hipSuccess || hipStreamSynchronize(g_stream) != hipSuccess) {
        dfree(f->d);
        delete f;
        return set_error(FLUX_ERR_HIP, "arrayN: clearing the new field failed");
    }
    *out = f;
    return FLUX_OK;
}

int flux_field_destroy(flux_field* f) {
    if (!f) return FLUX_OK;
    if (g_init) (void)hipStreamSynchronize(g_stream);
    dfree(f->d);
    delete f;
    return FLUX_OK;
}

int flux_field_upload(flux_field* f, const void* src, size_t bytes) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(f && src, "fromArray: NULL argument");
    FLUX_REQUIRE(bytes == f->bytes, "fromArray: %zu bytes given, field holds %zu", bytes, f->bytes);
    FLUX_HIP(hipMemcpyAsync(f->d, src, bytes, hipMemcpyHostToDevice, g_stream));
    FLUX_HIP(hipStreamSynchronize(g_stream));
    return FLUX_OK;
}

int flux_field_download(const flux_field* f, void* dst, size_t bytes) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(f && dst, "toArray: NULL argument");
    FLUX_REQUIRE(bytes == f->bytes, "toArray: %zu bytes requested, field holds %zu", bytes, f->bytes);
    FLUX_HIP(hipMemcpyAsync(dst, f->d, bytes, hipMemcpyDeviceToHost, g_stream));
    FLUX_HIP(hipStreamSynchronize(g_stream));
    return FLUX_OK;
}

int flux_field_copy(flux_field* dst, const flux_field* src) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(dst && src, "copy: NULL argument");
    FLUX_REQUIRE(dst->bytes == src->bytes, "copy: size mismatch (%zu vs %zu bytes)", dst->bytes, src->bytes);
    FLUX_HIP(hipMemcpyAsync(dst->d, src->d, src->bytes, hipMemcpyDeviceToDevice, g_stream));
    return FLUX_OK;
}

int flux_field_info(const flux_field* f, int* dtype, int* ndim, int64_t* shape, int64_t* count) {
    FLUX_REQUIRE(f, "info: field is NULL");
    if (dtype) *dtype = f->dtype;
    if (ndim) *ndim = f->ndim;
    if (shape)
        for (int a = 0; a < 3; ++a) shape[a] = f->shape[a];
    if (count) *count = f->count;
    return FLUX_OK;
}

void* flux_field_data(const flux_field* f) { return f ? f->d : nullptr; }

static hipEvent_t g_t0 = nullptr, g_t1 = nullptr;

int flux_timer_start(void) {
    FLUX_REQUIRE_INIT();
    if (!g_t0) {
        FLUX_HIP(hipEventCreate(&g_t0));
        FLUX_HIP(hipEventCreate(&g_t1));
    }
    FLUX_HIP(hipEventRecord(g_t0, g_stream));
    return FLUX_OK;
}

int flux_timer_stop(double* ms) {
    FLUX_REQUIRE_INIT();
    FLUX_REQUIRE(g_t0, "timer: flux_timer_start not called");
    FLUX_HIP(hipEventRecord(g_t1, g_stream));
    FLUX_HIP(hipEventSynchronize(g_t1));
    float t = 0;
    FLUX_HIP(hipEventElapsedTime(&t, g_t0, g_t1));
    if (ms) *ms = t;
    return FLUX_OK;
}

int flux_kprof_enable(const char* name) {
    FLUX_REQUIRE_INIT();
    g_kprof.name = name ? name : "";
    g_kprof.used = 0;
    g_kprof.open = false;
    return FLUX_OK;
}

int flux_kprof_read(double* total_ms, int64_t* launches) {
    FLUX_REQUIRE_INIT();
    FLUX_HIP(hipStreamSynchronize(g_stream));
    double sum = 0;
    for (size_t q = 0; q < g_kprof.used; ++q) {
        float t = 0;
        FLUX_HIP(hipEventElapsedTime(&t, g_kprof.pool[q].first, g_kprof.pool[q].second));
        sum += t;
    }
    if (total_ms) *total_ms = sum;
    if (launches) *launches = (int64_t)g_kprof.used;
    g_kprof.used = 0;
    return FLUX_OK;
}

}  // extern "C"
