/* FP32 instantiation of the oracle solver (test infrastructure only). */
#define REAL float
#define SUF f32
#include "solver_body.h"
