// Multigrid V-cycle preconditioner of the p' system.
//
// The reference's multigrid (benchmark/mgpcg.py:91-135) is a V-cycle for the
// constant-coefficient Poisson operator: red-black Gauss-Seidel smoothing
// (:118-123), summed restriction of the residual (:103-110), injection
// prolongation (:112-115), 50 sweeps on the coarsest level (:127-128), the
// post-smoothing with the colours reversed (:131-135).  Here the operator is
// the variable-coefficient 5-point p' system (links may vanish at walls and
// obstacles, identity rows inside them), so each level is built from the one
// above: 2 x 2 aggregates of interior cells, the Galerkin operator P^T A P
// with P the indicator of the free cells (a cell with a non-zero link), which
// is again a 5-point operator.  The correction is prolonged with weight 1.5:
// piecewise-constant prolongation leaves the Galerkin correction about half
// too small on smooth errors (per-cycle contraction on a Poisson problem
// 0.67 / 0.71 at 64^2 / 128^2 cells with weight 1, 0.45 / 0.52 with 1.5).
//
// One V-cycle from z = 0: per level one red-black sweep (red first), the
// residual summed into the aggregates, the next level, the correction added,
// one sweep black first; the coarsest level (at most kMinCells interior
// cells) takes kCoarseSweeps sweeps in one workgroup with z and r in LDS.
//
// Every level is laid out as the finest: padded (X, Y) arrays, index
// i * Y + j, coefficient plane q at q * X * Y.  The levels below the finest
// live in one arena per context (coefficients, r, z: 7 X Y doubles per
// level), allocated when the system shape changes; the coefficients are
// rebuilt by the caller's set-up for every new system.
//
// All arithmetic is FP64 with contraction off, each expression in the order
// the oracle's restatement (oracle/krylov.py mg_*) writes it, so the GPU's
// M(r) equals the oracle's to rounding of the division alone.
#include "km_mg.h"

#include <algorithm>

struct km_mg {
    int NX = 0, NY = 0;
    int L = 0;                         // levels, finest first
    int X[12] = {}, Y[12] = {};        // padded dimensions per level
    const double* coef0 = nullptr;     // the finest level's coefficients (the caller's, per set-up)
    double* coef[12] = {};             // levels >= 1 (arena)
    double* r[12] = {};
    double* z[12] = {};
    void* arena = nullptr;
};

namespace {

constexpr int kMaxLevels = 12;
constexpr int kMinCells = 400;        // interior cells at which coarsening stops
constexpr int kCoarsestMax = 3072;    // padded cells of the coarsest level (z and r in 48 KB of LDS)
constexpr int kCoarseSweeps = 50;
constexpr int kCoarseThreads = 1024;
constexpr double kOmega = 1.0;
constexpr double kKappa = 1.5;

inline unsigned grid_of(int64_t n) { return (unsigned)((n + KM_BLOCK - 1) / KM_BLOCK); }

// (A z)_k summed as c0 z + c1 z_W + c2 z_E + c3 z_N + c4 z_S, zero outside the array
__device__ __forceinline__ double stencil_at(const double* __restrict__ c, const double* __restrict__ z, int64_t n,
                                             int64_t k, int i, int j, int X, int Y) {
#pragma clang fp contract(off)
    const double zw = i > 0 ? z[k - Y] : 0.0;
    const double ze = i < X - 1 ? z[k + Y] : 0.0;
    const double zn = j < Y - 1 ? z[k + 1] : 0.0;
    const double zs = j > 0 ? z[k - 1] : 0.0;
    return c[k] * z[k] + c[n + k] * zw + c[2 * n + k] * ze + c[3 * n + k] * zn + c[4 * n + k] * zs;
}

__device__ __forceinline__ bool free_at(const double* __restrict__ c, int64_t n, int64_t k) {
    return c[n + k] != 0.0 || c[2 * n + k] != 0.0 || c[3 * n + k] != 0.0 || c[4 * n + k] != 0.0;
}

__device__ __forceinline__ double inv_diag(double d) { return d != 0.0 ? 1.0 / d : 0.0; }

// The Galerkin operator of the 2 x 2 aggregates of level (cf, Xf, Yf): one
// thread per padded coarse cell.  A free fine cell contributes its aP, its
// links inside the aggregate to aP, its links across an aggregate edge (or to
// the halo) to that side's coarse link; an aggregate with no free cell is an
// identity row, the halo zero.
__global__ void __launch_bounds__(KM_BLOCK) k_mg_coarsen(const double* __restrict__ cf, int Xf, int Yf,
                                                       double* __restrict__ cc, int Xc, int Yc) {
#pragma clang fp contract(off)
    const int64_t nc = (int64_t)Xc * Yc;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nc) return;
    const int I = (int)(k / Yc), J = (int)(k % Yc);
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, a4 = 0.0;
    if (I > 0 && I < Xc - 1 && J > 0 && J < Yc - 1) {
        const int nx = Xf - 2, ny = Yf - 2;
        const int64_t nf = (int64_t)Xf * Yf;
        bool hit = false;
        for (int di = 0; di < 2; ++di)
            for (int dj = 0; dj < 2; ++dj) {
                const int i = 2 * I - 1 + di, j = 2 * J - 1 + dj;
                if (i > nx || j > ny) continue;
                const int64_t f = (int64_t)i * Yf + j;
                const double c1 = cf[nf + f], c2 = cf[2 * nf + f], c3 = cf[3 * nf + f], c4 = cf[4 * nf + f];
                if (!(c1 != 0.0 || c2 != 0.0 || c3 != 0.0 || c4 != 0.0)) continue;
                hit = true;
                a0 += cf[f];
                if (di == 1) a0 += c1; else a1 += c1;
                if (di == 0 && i + 1 <= nx) a0 += c2; else a2 += c2;
                if (dj == 0 && j + 1 <= ny) a0 += c3; else a3 += c3;
                if (dj == 1) a0 += c4; else a4 += c4;
            }
        if (!hit) a0 = 1.0;
    }
    cc[k] = a0;
    cc[nc + k] = a1;
    cc[2 * nc + k] = a2;
    cc[3 * nc + k] = a3;
    cc[4 * nc + k] = a4;
}

// z = 0, then the red half of the first sweep (z_red = w D^-1 r: A z = 0)
__global__ void __launch_bounds__(KM_BLOCK) k_mg_first(const double* __restrict__ c, const double* __restrict__ r,
                                                     double* __restrict__ z, int X, int Y, const int* skip) {
#pragma clang fp contract(off)
    if (skip && *skip) return;
    const int64_t n = (int64_t)X * Y;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int i = (int)(k / Y), j = (int)(k % Y);
    double v = 0.0;
    if (((i + j) & 1) == 0) v = 0.0 + (kOmega * inv_diag(c[k])) * (r[k] - 0.0);
    z[k] = v;
}

// One colour's half of a red-black SOR sweep: z += w D^-1 (r - A z) on the
// cells with (i + j) % 2 == colour (they read only the other colour's values)
__global__ void __launch_bounds__(KM_BLOCK) k_mg_half(const double* __restrict__ c, const double* __restrict__ r,
                                                    double* __restrict__ z, int X, int Y, int colour,
                                                    const int* skip) {
#pragma clang fp contract(off)
    if (skip && *skip) return;
    const int64_t n = (int64_t)X * Y;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int i = (int)(k / Y), j = (int)(k % Y);
    if (((i + j) & 1) != colour) return;
    const double az = stencil_at(c, z, n, k, i, j, X, Y);
    z[k] = z[k] + (kOmega * inv_diag(c[k])) * (r[k] - az);
}

// rc = the residual r - A z of the free interior fine cells summed into their
// aggregates (children in the order (0,0), (0,1), (1,0), (1,1)); halo zero
__global__ void __launch_bounds__(KM_BLOCK) k_mg_restrict(const double* __restrict__ c, const double* __restrict__ r,
                                                        const double* __restrict__ z, int X, int Y,
                                                        double* __restrict__ rc, int Xc, int Yc, const int* skip) {
#pragma clang fp contract(off)
    if (skip && *skip) return;
    const int64_t nc = (int64_t)Xc * Yc;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nc) return;
    const int I = (int)(k / Yc), J = (int)(k % Yc);
    double acc = 0.0;
    if (I > 0 && I < Xc - 1 && J > 0 && J < Yc - 1) {
        const int64_t n = (int64_t)X * Y;
        for (int di = 0; di < 2; ++di)
            for (int dj = 0; dj < 2; ++dj) {
                const int i = 2 * I - 1 + di, j = 2 * J - 1 + dj;
                if (i > X - 2 || j > Y - 2) continue;
                const int64_t f = (int64_t)i * Y + j;
                const double res = r[f] - stencil_at(c, z, n, f, i, j, X, Y);
                acc = acc + (free_at(c, n, f) ? res : 0.0);
            }
    }
    rc[k] = acc;
}

// z += kappa zc of the aggregate on the free interior fine cells
__global__ void __launch_bounds__(KM_BLOCK) k_mg_prolong(const double* __restrict__ c, double* __restrict__ z,
                                                       const double* __restrict__ zc, int X, int Y, int Yc,
                                                       const int* skip) {
#pragma clang fp contract(off)
    if (skip && *skip) return;
    const int64_t n = (int64_t)X * Y;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int i = (int)(k / Y), j = (int)(k % Y);
    if (i < 1 || i > X - 2 || j < 1 || j > Y - 2 || !free_at(c, n, k)) return;
    z[k] = z[k] + kKappa * zc[(int64_t)((i - 1) / 2 + 1) * Yc + (j - 1) / 2 + 1];
}

// The coarsest level: z = 0, then `sweeps` red-black sweeps (red first), in
// one workgroup with z and r in LDS (X * Y <= kCoarsestMax)
__global__ void __launch_bounds__(kCoarseThreads) k_mg_coarsest(const double* __restrict__ c,
                                                              const double* __restrict__ r, double* __restrict__ z,
                                                              int X, int Y, int sweeps, const int* skip) {
#pragma clang fp contract(off)
    if (skip && *skip) return;
    extern __shared__ double lds[];
    const int n = X * Y;
    double* zs = lds;
    double* rs = lds + n;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        zs[k] = 0.0;
        rs[k] = r[k];
    }
    __syncthreads();
    for (int s = 0; s < sweeps; ++s)
        for (int colour = 0; colour < 2; ++colour) {
            for (int k = threadIdx.x; k < n; k += blockDim.x) {
                const int i = k / Y, j = k % Y;
                if (((i + j) & 1) != colour) continue;
                const double az = stencil_at(c, zs, n, k, i, j, X, Y);
                zs[k] = zs[k] + (kOmega * inv_diag(c[k])) * (rs[k] - az);
            }
            __syncthreads();
        }
    for (int k = threadIdx.x; k < n; k += blockDim.x) z[k] = zs[k];
}

}  // namespace

namespace kmmg {

int setup(km_ctx* ctx, int NX, int NY, const double* coef, hipStream_t s) {
    KM_REQUIRE(ctx, coef && NX >= 3 && NY >= 3, "multigrid: system %d x %d", NX, NY);
    km_mg* M = ctx->mg;
    if (!M || M->NX != NX || M->NY != NY) {
        int X[kMaxLevels], Y[kMaxLevels];
        int L = 1;
        X[0] = NX;
        Y[0] = NY;
        while (L < kMaxLevels) {
            const int nx = X[L - 1] - 2, ny = Y[L - 1] - 2;
            if ((int64_t)nx * ny <= kMinCells || nx < 4 || ny < 4) break;
            X[L] = (nx + 1) / 2 + 2;
            Y[L] = (ny + 1) / 2 + 2;
            ++L;
        }
        KM_REQUIRE(ctx, (int64_t)X[L - 1] * Y[L - 1] <= kCoarsestMax,
                   "multigrid: coarsest level %d x %d holds more than %d cells", X[L - 1], Y[L - 1], kCoarsestMax);
        size_t total = 0;
        for (int l = 1; l < L; ++l) total += (size_t)7 * X[l] * Y[l];
        km_mg_free(ctx);
        M = new km_mg;
        M->NX = NX;
        M->NY = NY;
        M->L = L;
        for (int l = 0; l < L; ++l) {
            M->X[l] = X[l];
            M->Y[l] = Y[l];
        }
        if (total > 0) {
            hipError_t e;
            {
                KmHostLock lk(km_host_mutex());
                e = hipMalloc(&M->arena, total * sizeof(double));
            }
            if (e != hipSuccess) {
                delete M;
                return km_fail(ctx, KM_EHIP, "multigrid arena of %zu bytes: %s", total * sizeof(double),
                               hipGetErrorString(e));
            }
        }
        double* p = static_cast<double*>(M->arena);
        for (int l = 1; l < L; ++l) {
            const size_t n = (size_t)X[l] * Y[l];
            M->coef[l] = p;
            M->r[l] = p + 5 * n;
            M->z[l] = p + 6 * n;
            p += 7 * n;
        }
        ctx->mg = M;
    }
    M->coef0 = coef;
    for (int l = 0; l + 1 < M->L; ++l) {
        const double* cf = l == 0 ? coef : M->coef[l];
        hipLaunchKernelGGL(k_mg_coarsen, dim3(grid_of((int64_t)M->X[l + 1] * M->Y[l + 1])), dim3(KM_BLOCK), 0, s, cf,
                           M->X[l], M->Y[l], M->coef[l + 1], M->X[l + 1], M->Y[l + 1]);
    }
    KM_CHECK_LAUNCH(ctx);
    return KM_OK;
}

int vcycle(km_ctx* ctx, const double* r, double* z, hipStream_t s, const int* skip) {
    const km_mg* M = ctx->mg;
    KM_REQUIRE(ctx, M && M->L > 0 && M->coef0, "multigrid hierarchy not set up");
    const int L = M->L;
    auto C = [&](int l) { return l == 0 ? M->coef0 : (const double*)M->coef[l]; };
    auto R = [&](int l) { return l == 0 ? r : (const double*)M->r[l]; };
    auto Z = [&](int l) { return l == 0 ? z : M->z[l]; };
    for (int l = 0; l + 1 < L; ++l) {
        const int X = M->X[l], Y = M->Y[l];
        const unsigned g = grid_of((int64_t)X * Y);
        hipLaunchKernelGGL(k_mg_first, dim3(g), dim3(KM_BLOCK), 0, s, C(l), R(l), Z(l), X, Y, skip);
        hipLaunchKernelGGL(k_mg_half, dim3(g), dim3(KM_BLOCK), 0, s, C(l), R(l), Z(l), X, Y, 1, skip);
        hipLaunchKernelGGL(k_mg_restrict, dim3(grid_of((int64_t)M->X[l + 1] * M->Y[l + 1])), dim3(KM_BLOCK), 0, s,
                           C(l), R(l), (const double*)Z(l), X, Y, M->r[l + 1], M->X[l + 1], M->Y[l + 1], skip);
    }
    {
        const int X = M->X[L - 1], Y = M->Y[L - 1];
        hipLaunchKernelGGL(k_mg_coarsest, dim3(1), dim3(kCoarseThreads), (size_t)2 * X * Y * sizeof(double), s,
                           C(L - 1), R(L - 1), Z(L - 1), X, Y, kCoarseSweeps, skip);
    }
    for (int l = L - 2; l >= 0; --l) {
        const int X = M->X[l], Y = M->Y[l];
        const unsigned g = grid_of((int64_t)X * Y);
        hipLaunchKernelGGL(k_mg_prolong, dim3(g), dim3(KM_BLOCK), 0, s, C(l), Z(l), (const double*)M->z[l + 1], X, Y,
                           M->Y[l + 1], skip);
        hipLaunchKernelGGL(k_mg_half, dim3(g), dim3(KM_BLOCK), 0, s, C(l), R(l), Z(l), X, Y, 1, skip);
        hipLaunchKernelGGL(k_mg_half, dim3(g), dim3(KM_BLOCK), 0, s, C(l), R(l), Z(l), X, Y, 0, skip);
    }
    KM_CHECK_LAUNCH(ctx);
    return KM_OK;
}

int levels(const km_ctx* ctx) { return ctx->mg ? ctx->mg->L : 0; }

const void* arena(const km_ctx* ctx) { return ctx->mg ? ctx->mg->arena : nullptr; }

}  // namespace kmmg

void km_mg_free(km_ctx* ctx) {
    if (!ctx || !ctx->mg) return;
    if (ctx->mg->arena) {
        KmHostLock lk(km_host_mutex());
        (void)hipFree(ctx->mg->arena);
    }
    delete ctx->mg;
    ctx->mg = nullptr;
}

extern "C" {

int km_mg_vcycle(km_ctx* ctx, int NX, int NY, const double* coef, const double* r, double* z, int* levels) {
    if (!ctx) return KM_EINVAL;
    KM_REQUIRE(ctx, coef && r && z, "null pointer");
    KM_REQUIRE(ctx, NX >= 3 && NY >= 3 && (int64_t)NX * NY < ((int64_t)1 << 31), "system shape %d x %d", NX, NY);
    if (int rc = kmmg::setup(ctx, NX, NY, coef, ctx->stream)) return rc;
    if (int rc = kmmg::vcycle(ctx, r, z, ctx->stream, nullptr)) return rc;
    if (levels) *levels = ctx->mg->L;
    KM_HIP_SYNC(ctx, hipStreamSynchronize(ctx->stream));
    return KM_OK;
}

}  // extern "C"
