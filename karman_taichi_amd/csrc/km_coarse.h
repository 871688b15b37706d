// Internal interface of the aggregation coarse level (km_coarse.hip) and the
// fused two-level flexible BiCGSTAB (km_twolevel.hip), used by the flexible
// solver (km_fbicg.hip).  Not part of the C-ABI.
#pragma once

#include "km_common.h"

#include <map>
#include <tuple>

// Coarse level of one p' system: B x B aggregates of the interior, i-major.
struct km_coarse {
    int NX = 0, NY = 0, B = 0, mI = 0, mJ = 0, m = 0, bw = 0;
    int block_req = 0;        // km_set_coarse (0: automatic)
    char* arena = nullptr;
    size_t bytes = 0;
    double* cc = nullptr;     // 5 planes of m coarse coefficients
    double* band = nullptr;   // m x (2 bw + 1) band of A_c, then its LU factors
    double* inv = nullptr;    // m x m A_c^-1, row-major
    double* rc = nullptr;     // kMaxSlices x m restricted residual (per-slice partial sums)
    double* yc = nullptr;     // m coarse solution
    double* fy = nullptr;     // fused iteration: Y(r), Y(p), Y(v), Y(s) (4 m)
    double* frc = nullptr;    // fused iteration: slot partials of P^T v, then of P^T t (2 kMaxVS m)
    uint8_t* mask = nullptr;  // NX x NY: 1 on the free interior cells (the support of P)
    size_t lds_max = 0;       // LDS per workgroup on this device (the whole-band factorisation)
    bool held = false;        // inv holds A_c^-1 of a coarse level of the current shape
    int update = 32;          // km_set_coarse_update: 0 factor, 1 Newton-Schulz update each solve, k every k-th
    int since = 0;            // solves that reused the held inverse since it was last updated
    const double* coef = nullptr;   // the coefficient array of the last set-up (reuse needs the same one)
    bool refresh = false;     // a solve ran long: factor on the next set-up
    int totals = -1;          // km_set_coarse_totals: -1 automatic, 0 never, 1 always
    int recent[8] = {};       // iteration counts of the last fused solves of this shape (0: none)
    int chunk_big = 8, chunk_small = 2, chunk_lead = 2;   // km_set_coarse_chunks
    uint64_t used = 0;        // coarse_tick of the last solve that used this level
    // tile shape of the fused smoothing launches per (NX, NY, sizeof T)
    std::map<std::tuple<int, int, int>, int> tiles;
};

namespace kmcoarse {

// f on every coarse level of the context (a setting applies to all)
void for_each_level(km_ctx* ctx, const std::function<void(km_coarse*)>& f);

// Red-black SOR sweeps of the two-level smoother (each a red then a black
// half-sweep) and their relaxation factor: on the 1280 x 256 channel's p'
// systems two sweeps at 1.3 take the flexible solve from 45 iterations with
// four damped Jacobi sweeps (omega 0.8, the same halo) to 27..34 in the
// oracle; 1.15 .. 1.35 do about as well, 1.5 and above worse.
constexpr int kSweeps = 2;
constexpr double kOmega = 1.3;
// The held inverse X of the last coarse operator is updated to
// X (2 I - A_c X) for the next one when ||I - A_c X||_inf < kNsTol (then
// I - A_c X' = (I - A_c X)^2 and the error is below kNsTol^2); otherwise A_c
// is factored anew.
constexpr double kNsTol = 0.5;
// Row slices per aggregate of a restriction: the partial restrictions
// rc[s m + k] are summed by the coarse solve.
constexpr int kMaxSlices = 8;
// Restriction slots per aggregate in the fused iteration: the tiles that
// overlap one aggregate (32 x 32 tiles on 256 x 256 aggregates: 64).
constexpr int kMaxVS = 64;

// Aggregate edge length for an nx x ny interior (req > 0: that size; else the
// smallest of 8, 12, 16, 24, ..., 256 that keeps the coarse level at <= 2048
// unknowns, or on larger grids at as many as keep the two reads of the dense
// coarse inverse per iteration (16 m^2 bytes) within ~40% of the fine grid's
// ~210 bytes per cell, 4 m^2 <= 21 nx ny, at most 8192), skipping sizes with
// more than kMaxBandwidth aggregate columns.  On the 1280 x 256
// channel 16 x 16 aggregates (1280 unknowns) beat 12, 14 and 20 (436 against
// 329-372 SIMPLE steps/s); on the 4096 x 820 channel 32 x 32 (3328) beat the
// 48 x 48 of the 2048 bound, 24 x 24 and 40 x 40 (45.4 against 38.0-39.0).
int choose_block(int nx, int ny, int req);
// The widest coarse band (aggregate columns) the factorisation's LDS ring of
// (bw + 1) x (2 bw + 1) doubles holds in 64 KB; choose_block skips sizes
// with more.
constexpr int kMaxBandwidth = 63;

// Galerkin coarse operator P^T A P of the FP64 system, its band LU, its
// explicit inverse and the free-cell mask (enqueued on s).  solver: set up for
// a flexible solve, which may run with a lagged level (km_set_coarse_update).
int setup(km_ctx* ctx, int NX, int NY, const double* coef, hipStream_t s, bool solver = false);

// rc = P^T r (one workgroup per aggregate) and yc = A_c^-1 rc.
template <typename T>
int restrict_solve(km_ctx* ctx, const T* r, const int* skip, hipStream_t s);
// yc = A_c^-1 rc with rc given as S slice partials.
int coarse_solve(km_ctx* ctx, int S, const int* skip, hipStream_t s);

// z1 = P A_c^-1 P^T r and r1 = r - A z1 (A in T for the residual, the FP64
// coefficients for the free-cell mask); Prr (may be null) receives the
// per-workgroup partials of r1.r1 of a make_flat(NX * NY) launch.
template <typename T>
int apply(km_ctx* ctx, const double* coef, const T* coefT, const T* r, T* z1, T* r1, double* Prr, const int* skip,
          hipStream_t s);

// z = M(r): z1 = P yc from the current coarse solution yc, then kSweeps
// red-black SOR sweeps on A z = r (one tiled launch).
template <typename T>
int smooth(km_ctx* ctx, const T* coefT, const T* r, T* z, hipStream_t s);

// Inner flexible BiCGSTAB for A d = rhs with the two-level preconditioner
// (red-black SOR smoothing), four launches per iteration; ws holds 8n values of T.
// sym: the caller asserts aW(i, j) = aE(i - 1, j) and aS(i, j) = aN(i, j - 1)
// bit for bit (the smoothing tiles then read three coefficient planes).
// coef64: the FP64 coefficients (required with T = float: the FP32-vector
// smoothing tiles compute with them).
template <typename T>
int fused_inner(km_ctx* ctx, int NX, int NY, const T* coefT, const T* rhs, T* xd, T* ws, double eps, int max_iter,
                bool use_graph, KState* out, const std::function<int()>* tail = nullptr, bool sym = false,
                const double* coef64 = nullptr);

// Group solves: up to kGroupSystems independent symmetric FP64 systems of
// fused_inner (from x = 0), each on its own coarse level C[k] (set up by the
// caller; ctx->coarse is left on the last), all with smoothing tile `tile`,
// iterated in the same launches until every one stops; system k keeps its
// state in ctx->st[4 + k] and out[k] receives it.  Each system's arithmetic
// is that of its own fused_inner with the same tile.
constexpr int kGroupSystems = 4;
int fused_inner_n(km_ctx* ctx, int n, km_coarse* const* C, const int* NX, const int* NY, const double* const* coef,
                  const double* const* rhs, double* const* xd, double* const* ws, double eps, const int* max_iter,
                  bool use_graph, KState* out, const std::function<int()>* tail, int tile);
// The smoothing tile of level C's symmetric FP64 system (timed on its first
// use; ws: 5n scratch values); leaves ctx->coarse = C.
int group_tile(km_ctx* ctx, km_coarse* C, const double* coef, const double* rhs, double* ws, int* tile);

// Inner flexible BiCGSTAB for A d = rhs with one or two red-black Gauss-Seidel sweeps
// as the preconditioner and no coarse level (the momentum systems), three
// launches per iteration; ws holds 8n values of T.
template <typename T>
int smooth_inner(km_ctx* ctx, int NX, int NY, const T* coefT, const T* rhs, T* xd, T* ws, double eps, int max_iter,
                 bool use_graph, KState* out, const std::function<int()>* tail = nullptr);
// Two independent systems of smooth_inner (arrays of two: shapes, operands,
// 8n-value workspaces) iterated in the same launches until both stop; out
// receives both states, tail runs after the last chunk.
template <typename T>
int smooth_inner2(km_ctx* ctx, const int* NX, const int* NY, const T* const* coefT, const T* const* rhs, T* const* xd,
                  T* const* ws, double eps, const int* max_iter, bool use_graph, KState* out,
                  const std::function<int()>* tail = nullptr);

// Tile shape (rows, columns) of the fused smoothing kernels chosen for the
// coarse level's system in FP64 (32 x 32 before the first fused solve).
int tile_info(km_ctx* ctx, int* ti, int* tj, int* threads);
int set_tile(km_ctx* ctx, int which);

// km_time_kernel for the fused kernels (which 4..7: K_PV, K_CV, K_ST, K_CT;
// sym: the three-plane smoothing kernels of a symmetric system).
int time_kernel(km_ctx* ctx, int which, int NX, int NY, const double* coef, const double* b, int reps, bool b2b,
                double* avg_us, bool sym = false);

}  // namespace kmcoarse

void km_coarse_free(km_ctx* ctx);
