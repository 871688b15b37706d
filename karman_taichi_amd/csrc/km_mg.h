// Multigrid V-cycle preconditioner of the p' system (km_mg.hip): a
// comparator for the two-level and learned preconditioners in the spirit of
// benchmark/mgpcg.py:91-135 (V-cycle, red-black Gauss-Seidel smoothing,
// summed restriction, injection prolongation), on an aggregation hierarchy
// of the variable-coefficient 5-point operator.
#pragma once

#include "km_common.h"

namespace kmmg {

// Build (or rebuild) the hierarchy of the NX x NY system coef on stream s:
// each level aggregates 2 x 2 interior cells of the level above, with the
// Galerkin operator P^T A P (P: the free cells' indicator), until the
// interior has at most 400 cells or a side has fewer than 4 cells.
int setup(km_ctx* ctx, int NX, int NY, const double* coef, hipStream_t s);

// z = M(r): one V-cycle from z = 0 (r, z: the finest level's FP64 vectors),
// kernel launches only (graph capture may record it).  skip (may be null): a
// device word; the launches do nothing while it is non-zero.
int vcycle(km_ctx* ctx, const double* r, double* z, hipStream_t s, const int* skip);

// Levels of the current hierarchy (0: none).
int levels(const km_ctx* ctx);

// The device arena of the levels below the finest (changes when the
// hierarchy is rebuilt for another system shape; captured graphs key on it).
const void* arena(const km_ctx* ctx);

}  // namespace kmmg

void km_mg_free(km_ctx* ctx);
