// Mixed-precision solve: FP64 iterative refinement around an inner PCG in
// reduced precision.
#include "mixed.h"

namespace gputo {

void mixed_create(gputo_handle *) {}
void mixed_destroy(gputo_handle *) {}
void mixed_stats(gputo_handle *, gputo_to_stats *) {}
int mixed_bench_operator(gputo_handle *, int, int, double *) { return GPUTO_ERR_ARGUMENT; }

int mixed_solve(gputo_handle *, const double *, double *, double, int, double *, int *nit)
{
    *nit = GPUTO_ERR_ARGUMENT;
    return GPUTO_ERR_ARGUMENT;
}

}  // namespace gputo
