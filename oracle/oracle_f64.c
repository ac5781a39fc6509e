/* FP64 instantiation of the oracle solver (test infrastructure only). */
#define REAL double
#define SUF f64
#include "solver_body.h"
