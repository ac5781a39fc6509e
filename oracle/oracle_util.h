/* Shared index helpers for the oracle (test infrastructure only). */
#ifndef FLUX_ORACLE_UTIL_H
#define FLUX_ORACLE_UTIL_H
#include <stdint.h>
#include "oracle.h"

/* cell (i,j,k) -> flat index, first index fastest (TA/src/array.js:24-38) */
static inline int64_t orc_cidx(const int* n, int i, int j, int k) {
    return (int64_t)i + (int64_t)n[0] * ((int64_t)j + (int64_t)n[1] * k);
}

/* face array of `axis` (shape n with +1 on that axis), index (i,j,k) */
static inline int64_t orc_fidx(const int* n, int axis, int i, int j, int k) {
    int64_t nx = n[0] + (axis == 0), ny = n[1] + (axis == 1);
    return (int64_t)i + nx * ((int64_t)j + ny * k);
}

static inline int64_t orc_cells(const int* n) { return (int64_t)n[0] * n[1] * n[2]; }

static inline int64_t orc_faces(const int* n, int axis) {
    return (int64_t)(n[0] + (axis == 0)) * (n[1] + (axis == 1)) * (n[2] + (axis == 2));
}

#endif
