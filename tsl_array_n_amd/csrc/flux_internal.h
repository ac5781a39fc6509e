// Internal declarations shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/fluxflow_hip.h"

struct flux_field {
    void* d = nullptr;
    int dtype = FLUX_F32;
    int ndim = 1;
    int64_t shape[3] = {1, 1, 1};
    int64_t count = 0;
    size_t bytes = 0;
};

namespace flux {

extern hipStream_t g_stream;
extern bool g_init;
extern int g_device;
extern int g_cus;
// Blocks a column-march grid aims at (FLUX_MARCH_BLOCKS, default 4096:
// 16 blocks of 256 threads per CU on 256 CUs).
extern unsigned g_march_blocks;
// Consecutive finest-level column-march launches alternate the direction in
// which their blocks walk the z-chunks (FLUX_MARCH_FLIP=0 turns it off): a
// launch starts where the previous one ended, whose lines are still in the
// memory-side cache.  The toggle is reset at the start of every solve and
// V-cycle apply so the block order, and with it every partial sum, is the
// same from run to run.
extern bool g_march_flip;
extern unsigned g_march_toggle;

int set_error(int code, const char* fmt, ...);
void clear_error();
size_t dtype_size(int dtype);

// Device allocation owned by library objects (never exposed as a field).
void* dmalloc(size_t bytes);
void dfree(void* p);

// Per-kernel event timing (flux_kprof_*).
bool kprof_on(const char* name);
void kprof_begin(const char* name);
void kprof_end(const char* name);

inline int64_t cells_of(const int* n) { return (int64_t)n[0] * n[1] * n[2]; }
inline int64_t faces_of(const int* n, int axis) {
    return (int64_t)(n[0] + (axis == 0)) * (n[1] + (axis == 1)) * (n[2] + (axis == 2));
}
inline unsigned blocks_for(int64_t work, int threads) { return (unsigned)((work + threads - 1) / threads); }

// Checks a field against an expected element type and element count.
int check_field(const flux_field* f, int dtype, int64_t count, const char* what);

}  // namespace flux

#define FLUX_HIP(call)                                                                                   \
    do {                                                                                                 \
        hipError_t e_ = (call);                                                                          \
        if (e_ != hipSuccess)                                                                            \
            return ::flux::set_error(FLUX_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                                     __FILE__, __LINE__);                                                \
    } while (0)

#define FLUX_LAUNCH_CHECK()                                                                                  \
    do {                                                                                                     \
        hipError_t e_ = hipGetLastError();                                                                   \
        if (e_ != hipSuccess)                                                                                \
            return ::flux::set_error(FLUX_ERR_HIP, "kernel launch failed: %s (%s:%d)", hipGetErrorString(e_), \
                                     __FILE__, __LINE__);                                                    \
    } while (0)

#define FLUX_REQUIRE(cond, ...)                                        \
    do {                                                               \
        if (!(cond)) return ::flux::set_error(FLUX_ERR_ARG, __VA_ARGS__); \
    } while (0)

#define FLUX_REQUIRE_INIT()                                                                    \
    do {                                                                                       \
        if (!::flux::g_init) return ::flux::set_error(FLUX_ERR_STATE, "flux_init() not called"); \
    } while (0)

#define FLUX_TRY(expr)          \
    do {                        \
        int rc_ = (expr);       \
        if (rc_) return rc_;    \
    } while (0)
