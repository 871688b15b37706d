// Internal interface of the aggregation coarse level (km_coarse.hip) used by
// the flexible BiCGSTAB (km_fbicg.hip).  Not part of the C-ABI.
#pragma once

#include "km_common.h"

namespace kmcoarse {

// Aggregate edge length for an nx x ny interior (req > 0: that size; else the
// smallest of 8, 16, ..., 256 that keeps the coarse level at <= 1024 unknowns).
int choose_block(int nx, int ny, int req);

// Galerkin coarse operator P^T A P of the FP64 system, its band LU and its
// explicit inverse (enqueued on s).
int setup(km_ctx* ctx, int NX, int NY, const double* coef, hipStream_t s);

// z1 = P A_c^-1 P^T r and r1 = r - A z1 (A in T for the residual, the FP64
// coefficients for the free-cell mask); Prr (may be null) receives the
// per-workgroup partials of r1.r1 of a make_flat(NX * NY) launch.  omega > 0
// fuses the first damped Jacobi sweep: z1 <- z1 + omega D^-1 r1, r1 unused.
template <typename T>
int apply(km_ctx* ctx, const double* coef, const T* coefT, const T* r, T* z1, T* r1, double* Prr, const int* skip,
          hipStream_t s, double omega = 0.0);

// z_out = z_in + omega D^-1 res, with res = r - A z_in computed here when null.
template <typename T>
int jacobi(km_ctx* ctx, int NX, int NY, const T* coefT, const T* r, const T* res, const T* zin, T* zout, double omega,
           const int* skip, hipStream_t s);

}  // namespace kmcoarse

void km_coarse_free(km_ctx* ctx);
