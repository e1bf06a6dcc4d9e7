// Mixed-precision solve (FP64 iterative refinement around an FP32 / BF16
// inner PCG) and the learned coarse correction.
#pragma once

#include "gputo_internal.h"

namespace gputo {

void mixed_create(gputo_handle *h);
void mixed_destroy(gputo_handle *h);
int mixed_bench_operator(gputo_handle *h, int precision, int reps, double *avg_ms);
void mixed_stats(gputo_handle *h, gputo_to_stats *st);
int mixed_solve(gputo_handle *h, const double *F, double *U, double tol, int maxit, double *relres, int *nit);

}  // namespace gputo
