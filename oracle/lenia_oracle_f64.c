/* Double-precision instantiation of the CPU oracle: the FP64 reference for error
 * statements (test infrastructure only). */
#define REAL double
#define PFX(n) lo_f64_##n
#include "lenia_oracle_impl.h"

/* The analytic-growth exp, exported so its accuracy can be pinned. */
double lo_canon_exp(double x) { return canon_exp(x); }
