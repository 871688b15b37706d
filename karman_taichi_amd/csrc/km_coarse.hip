// Aggregation coarse level of the two-level p' preconditioner.
//
// The fine interior is cut into B x B aggregates; P is the indicator of the
// free cells (a link is non-zero) of each aggregate and the coarse operator is
// the Galerkin product A_c = P^T A P, again a 5-point operator on the
// aggregate grid.  A_c is factored (banded LU, bandwidth = aggregate columns)
// and inverted explicitly once per solve, so one application
//   z1 = P A_c^-1 P^T r,   r1 = r - A z1
// is a restriction, a small dense matvec and a prolongation fused with the
// residual.  The fine-level smoother (the network or damped Jacobi) then acts
// on r1 and its output is added to z1.
#include "km_coarse.h"

#include <cmath>

struct km_coarse {
    int NX = 0, NY = 0, B = 0, mI = 0, mJ = 0, m = 0, bw = 0;
    int block_req = 0;       // km_set_coarse (0: automatic)
    char* arena = nullptr;
    size_t bytes = 0;
    double* cc = nullptr;     // 5 planes of m coarse coefficients
    double* band = nullptr;   // m x (2 bw + 1) band of A_c, then its LU factors
    double* inv = nullptr;    // m x m A_c^-1 (symmetric)
    double* rc = nullptr;     // m restricted residual
    double* yc = nullptr;     // m coarse solution
};

namespace {

__device__ __forceinline__ bool free_link(const double* __restrict__ c, int64_t n, int64_t k) {
    return c[n + k] != 0.0 || c[2 * n + k] != 0.0 || c[3 * n + k] != 0.0 || c[4 * n + k] != 0.0;
}

// One workgroup per aggregate (I, J): sums of aP and of the links of its free
// cells, split into the diagonal (links inside the aggregate) and the four
// coarse links (links that cross its edges).  Fixed order: per-thread sums
// over a strided cell range, then the workgroup sum.
__global__ void __launch_bounds__(KM_BLOCK)
k_coarse_coef(int NX, int NY, int B, int mI, int mJ, const double* __restrict__ coef, double* __restrict__ cc) {
    __shared__ double sh[4];
    const int agg = blockIdx.x;
    const int I = agg / mJ, J = agg - I * mJ;
    const int nx = NX - 2, ny = NY - 2;
    const int i0 = 1 + I * B, j0 = 1 + J * B;
    const int i1 = i0 + B - 1 < nx ? i0 + B - 1 : nx, j1 = j0 + B - 1 < ny ? j0 + B - 1 : ny;
    const int bi = i1 - i0 + 1, bj = j1 - j0 + 1;
    const int64_t n = (int64_t)NX * NY;
    double d = 0.0, w = 0.0, e = 0.0, nn = 0.0, s = 0.0, cnt = 0.0;
    for (int t = threadIdx.x; t < bi * bj; t += KM_BLOCK) {
        const int i = i0 + t / bj, j = j0 + t % bj;
        const int64_t k = (int64_t)i * NY + j;
        if (!free_link(coef, n, k)) continue;
        cnt += 1.0;
        d += coef[k];
        const double aw = coef[n + k], ae = coef[2 * n + k], an = coef[3 * n + k], as = coef[4 * n + k];
        if (i == i0) w += aw; else d += aw;
        if (i == i1) e += ae; else d += ae;
        if (j == j1) nn += an; else d += an;
        if (j == j0) s += as; else d += as;
    }
    d = block_sum(d, sh);
    w = block_sum(w, sh);
    e = block_sum(e, sh);
    nn = block_sum(nn, sh);
    s = block_sum(s, sh);
    cnt = block_sum(cnt, sh);
    if (threadIdx.x == 0) {
        const int64_t m = (int64_t)mI * mJ;
        const bool empty = cnt == 0.0;
        cc[agg] = empty ? 1.0 : d;
        cc[m + agg] = empty || I == 0 ? 0.0 : w;
        cc[2 * m + agg] = empty || I == mI - 1 ? 0.0 : e;
        cc[3 * m + agg] = empty || J == mJ - 1 ? 0.0 : nn;
        cc[4 * m + agg] = empty || J == 0 ? 0.0 : s;
    }
}

// Band storage: band[r * (2 bw + 1) + bw + (c - r)] = A_c[r][c], |c - r| <= bw.
// Neighbours of aggregate r in i-major order: r -+ 1 (S / N) and r -+ mJ (W / E).
__global__ void __launch_bounds__(1024)
k_coarse_band(int m, int mJ, int bw, const double* __restrict__ cc, double* __restrict__ band) {
    const int W = 2 * bw + 1;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < (int64_t)m * W; t += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(t / W), o = (int)(t % W) - bw;
        const int c = r + o;
        double v = 0.0;
        if (c >= 0 && c < m) {
            if (o == 0) v = cc[r];
            else if (o == -mJ) v = cc[m + r];
            else if (o == mJ) v = cc[2 * m + r];
            else if (o == 1 && (r % mJ) != mJ - 1) v = cc[3 * m + r];
            else if (o == -1 && (r % mJ) != 0) v = cc[4 * m + r];
        }
        band[t] = v;
    }
}

// Banded LU without pivoting (A_c is SPD) in one workgroup: step k updates the
// (bw x bw) window below and right of the pivot; L (unit) overwrites the
// strict lower band, U the rest.
__global__ void __launch_bounds__(1024)
k_coarse_factor(int m, int bw, double* __restrict__ band) {
    const int W = 2 * bw + 1;
    for (int k = 0; k < m; ++k) {
        const int rows = (k + bw < m - 1 ? k + bw : m - 1) - k;   // rows k+1 .. k+rows
        const double piv = band[(int64_t)k * W + bw];
        for (int t = threadIdx.x; t < rows * rows; t += blockDim.x) {
            const int i = k + 1 + t / rows, j = k + 1 + t % rows;
            const double l = band[(int64_t)i * W + bw + (k - i)] / piv;
            const int o = j - i;
            if (o >= -bw && o <= bw) band[(int64_t)i * W + bw + o] -= l * band[(int64_t)k * W + bw + (j - k)];
        }
        __syncthreads();
        for (int t = threadIdx.x; t < rows; t += blockDim.x) {
            const int i = k + 1 + t;
            band[(int64_t)i * W + bw + (k - i)] /= piv;
        }
        __syncthreads();
    }
}

// Column c of A_c^-1 (= row c, A_c symmetric) by forward and back substitution
// with the band factors; one thread per column.
__global__ void __launch_bounds__(KM_BLOCK)
k_coarse_inverse(int m, int bw, const double* __restrict__ band, double* __restrict__ inv) {
    const int c = blockIdx.x * KM_BLOCK + threadIdx.x;
    if (c >= m) return;
    const int W = 2 * bw + 1;
    double* x = inv + (int64_t)c * m;
    for (int i = 0; i < c; ++i) x[i] = 0.0;
    for (int i = c; i < m; ++i) {   // L y = e_c (unit lower)
        double v = i == c ? 1.0 : 0.0;
        const int k0 = i - bw > c ? i - bw : c;
        for (int k = k0; k < i; ++k) v -= band[(int64_t)i * W + bw + (k - i)] * x[k];
        x[i] = v;
    }
    for (int i = m - 1; i >= 0; --i) {   // U x = y
        double v = x[i];
        const int k1 = i + bw < m - 1 ? i + bw : m - 1;
        for (int k = i + 1; k <= k1; ++k) v -= band[(int64_t)i * W + bw + (k - i)] * x[k];
        x[i] = v / band[(int64_t)i * W + bw];
    }
}

// rc[agg] = sum of r over the aggregate's free cells (one workgroup per aggregate)
template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_coarse_restrict(int NX, int NY, int B, int mJ, const double* __restrict__ coef, const T* __restrict__ r,
                  double* __restrict__ rc, const int* __restrict__ skip) {
    if (skip && *skip) return;
    __shared__ double sh[4];
    const int agg = blockIdx.x;
    const int I = agg / mJ, J = agg - I * mJ;
    const int nx = NX - 2, ny = NY - 2;
    const int i0 = 1 + I * B, j0 = 1 + J * B;
    const int bi = (i0 + B - 1 < nx ? i0 + B - 1 : nx) - i0 + 1, bj = (j0 + B - 1 < ny ? j0 + B - 1 : ny) - j0 + 1;
    const int64_t n = (int64_t)NX * NY;
    double a = 0.0;
    for (int t = threadIdx.x; t < bi * bj; t += KM_BLOCK) {
        const int64_t k = (int64_t)(i0 + t / bj) * NY + (j0 + t % bj);
        if (free_link(coef, n, k)) a += (double)r[k];
    }
    a = block_sum(a, sh);
    if (threadIdx.x == 0) rc[agg] = a;
}

// yc = inv rc: one wave per row
__global__ void __launch_bounds__(KM_BLOCK)
k_coarse_matvec(int m, const double* __restrict__ inv, const double* __restrict__ rc, double* __restrict__ yc,
                const int* __restrict__ skip) {
    if (skip && *skip) return;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (row >= m) return;
    const double* a = inv + (int64_t)row * m;
    double s = 0.0;
    for (int k = lane; k < m; k += 64) s += a[k] * rc[k];
    s = wave_sum(s);
    if (lane == 0) yc[row] = s;
}

__device__ __forceinline__ double coarse_value(int i, int j, int B, int mJ, int nx, int ny, const double* yc) {
    if (i < 1 || i > nx || j < 1 || j > ny) return 0.0;
    return yc[((i - 1) / B) * mJ + (j - 1) / B];
}

// z1 = P yc on free cells; r1 = r - A z1; partials of r1.r1 (P_rr, may be null).
// omega > 0: the first damped Jacobi sweep fused in -- z1 receives
// z1 + omega D^-1 r1 and r1 is not written.
template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_coarse_prolong(Flat fl, int NX, int NY, int B, int mJ, const double* __restrict__ coef,
                 const T* __restrict__ coefT, const T* __restrict__ r, const double* __restrict__ yc,
                 T* __restrict__ z1, T* __restrict__ r1, double* __restrict__ Prr, double omega,
                 const int* __restrict__ skip) {
    if (skip && *skip) return;
    __shared__ double sh[4];
    const int nx = NX - 2, ny = NY - 2;
    const int64_t n = fl.n;
    const int64_t lo = (int64_t)blockIdx.x * fl.chunk;
    const int64_t hi = lo + fl.chunk < n ? lo + fl.chunk : n;
    double a = 0.0;
    for (int64_t k = lo + threadIdx.x; k < hi; k += KM_BLOCK) {
        const int i = (int)(k / NY), j = (int)(k % NY);
        // z1 at the cell and its neighbours (zero on cells that are not free)
        auto zc = [&](int ii, int jj) -> T {
            if (ii < 0 || ii >= NX || jj < 0 || jj >= NY) return T(0);
            const int64_t kk = (int64_t)ii * NY + jj;
            return free_link(coef, n, kk) ? (T)coarse_value(ii, jj, B, mJ, nx, ny, yc) : T(0);
        };
        const T zP = zc(i, j);
        T az = coefT[k] * zP;
        az = az + coefT[n + k] * zc(i - 1, j);
        az = az + coefT[2 * n + k] * zc(i + 1, j);
        az = az + coefT[3 * n + k] * zc(i, j + 1);
        az = az + coefT[4 * n + k] * zc(i, j - 1);
        const T rv = r[k] - az;
        if (omega > 0.0) {
            const T d = coefT[k];
            z1[k] = zP + (d != T(0) ? (T)omega * rv / d : T(0));
        } else {
            z1[k] = zP;
            r1[k] = rv;
        }
        a += (double)rv * (double)rv;
    }
    if (Prr) {
        a = block_sum(a, sh);
        if (threadIdx.x == 0) Prr[blockIdx.x] = a;
    }
}

// z_out = z_in + omega D^-1 (r - A z_in); with res (r - A z_in already known) the stencil is skipped
template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_jacobi_sweep(int64_t n, int NX, int NY, const T* __restrict__ c, const T* __restrict__ r, const T* __restrict__ res,
               const T* __restrict__ zin, T* __restrict__ zout, double omega, const int* __restrict__ skip) {
    if (skip && *skip) return;
    for (int64_t k = (int64_t)blockIdx.x * KM_BLOCK + threadIdx.x; k < n; k += (int64_t)gridDim.x * KM_BLOCK) {
        T rk;
        if (res) {
            rk = res[k];
        } else {
            const int i = (int)(k / NY), j = (int)(k % NY);
            T az = c[k] * zin[k];
            az = az + c[n + k] * (i > 0 ? zin[k - NY] : T(0));
            az = az + c[2 * n + k] * (i < NX - 1 ? zin[k + NY] : T(0));
            az = az + c[3 * n + k] * (j < NY - 1 ? zin[k + 1] : T(0));
            az = az + c[4 * n + k] * (j > 0 ? zin[k - 1] : T(0));
            rk = r[k] - az;
        }
        const T d = c[k];
        zout[k] = zin[k] + (d != T(0) ? (T)omega * rk / d : T(0));
    }
}

int flat_blocks2(int64_t n) {
    int64_t g = (n + KM_BLOCK - 1) / KM_BLOCK;
    if (g > 2048) g = 2048;
    return g < 1 ? 1 : (int)g;
}

size_t a256(size_t v) { return (v + 255) & ~size_t(255); }

km_coarse* get_coarse(km_ctx* ctx) {
    if (!ctx->coarse) ctx->coarse = new km_coarse();
    return ctx->coarse;
}

}  // namespace

namespace kmcoarse {

int choose_block(int nx, int ny, int req) {
    if (req > 0) return req;
    for (int B : {8, 16, 32, 64, 128, 256}) {
        const int64_t m = (int64_t)((nx + B - 1) / B) * ((ny + B - 1) / B);
        if (m <= 1024) return B;
    }
    return 512;
}

int setup(km_ctx* ctx, int NX, int NY, const double* coef, hipStream_t s) {
    km_coarse* C = get_coarse(ctx);
    const int nx = NX - 2, ny = NY - 2;
    KM_REQUIRE(ctx, nx >= 1 && ny >= 1, "system too small for a coarse level");
    const int B = choose_block(nx, ny, C->block_req);
    const int mI = (nx + B - 1) / B, mJ = (ny + B - 1) / B, m = mI * mJ;
    KM_REQUIRE(ctx, m <= 8192, "coarse level of %d unknowns is too large (aggregate %d)", m, B);
    const int bw = mJ;
    const size_t need = a256(5 * (size_t)m * 8) + a256((size_t)m * (2 * bw + 1) * 8) + a256((size_t)m * m * 8) +
                        2 * a256((size_t)m * 8);
    if (!C->arena || C->bytes < need) {
        KM_HIP(ctx, hipStreamSynchronize(s));
        km_drop_graphs(ctx);
        if (C->arena) KM_HIP(ctx, hipFree(C->arena));
        C->arena = nullptr;
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, need);
        if (e != hipSuccess) return km_fail(ctx, KM_ENOMEM, "coarse hipMalloc(%zu): %s", need, hipGetErrorString(e));
        C->arena = static_cast<char*>(p);
        C->bytes = need;
    }
    char* q = C->arena;
    C->cc = reinterpret_cast<double*>(q); q += a256(5 * (size_t)m * 8);
    C->band = reinterpret_cast<double*>(q); q += a256((size_t)m * (2 * bw + 1) * 8);
    C->inv = reinterpret_cast<double*>(q); q += a256((size_t)m * m * 8);
    C->rc = reinterpret_cast<double*>(q); q += a256((size_t)m * 8);
    C->yc = reinterpret_cast<double*>(q);
    C->NX = NX; C->NY = NY; C->B = B; C->mI = mI; C->mJ = mJ; C->m = m; C->bw = bw;
    hipLaunchKernelGGL(k_coarse_coef, dim3(m), dim3(KM_BLOCK), 0, s, NX, NY, B, mI, mJ, coef, C->cc);
    hipLaunchKernelGGL(k_coarse_band, dim3(flat_blocks2((int64_t)m * (2 * bw + 1))), dim3(1024), 0, s, m, mJ, bw,
                       (const double*)C->cc, C->band);
    hipLaunchKernelGGL(k_coarse_factor, dim3(1), dim3(1024), 0, s, m, bw, C->band);
    hipLaunchKernelGGL(k_coarse_inverse, dim3((m + KM_BLOCK - 1) / KM_BLOCK), dim3(KM_BLOCK), 0, s, m, bw,
                       (const double*)C->band, C->inv);
    KM_CHECK_LAUNCH(ctx);
    return KM_OK;
}

template <typename T>
int apply(km_ctx* ctx, const double* coef, const T* coefT, const T* r, T* z1, T* r1, double* Prr, const int* skip,
          hipStream_t s, double omega) {
    km_coarse* C = ctx->coarse;
    KM_REQUIRE(ctx, C && C->arena && C->m > 0, "coarse level not set up");
    const Flat fl = make_flat((int64_t)C->NX * C->NY);
    hipLaunchKernelGGL(k_coarse_restrict<T>, dim3(C->m), dim3(KM_BLOCK), 0, s, C->NX, C->NY, C->B, C->mJ, coef, r,
                       C->rc, skip);
    hipLaunchKernelGGL(k_coarse_matvec, dim3((C->m + 3) / 4), dim3(KM_BLOCK), 0, s, C->m, (const double*)C->inv,
                       (const double*)C->rc, C->yc, skip);
    hipLaunchKernelGGL(k_coarse_prolong<T>, dim3(fl.G), dim3(KM_BLOCK), 0, s, fl, C->NX, C->NY, C->B, C->mJ, coef,
                       coefT, r, (const double*)C->yc, z1, r1, Prr, omega, skip);
    KM_CHECK_LAUNCH(ctx);
    return KM_OK;
}

template <typename T>
int jacobi(km_ctx* ctx, int NX, int NY, const T* coefT, const T* r, const T* res, const T* zin, T* zout, double omega,
           const int* skip, hipStream_t s) {
    const int64_t n = (int64_t)NX * NY;
    hipLaunchKernelGGL(k_jacobi_sweep<T>, dim3(flat_blocks2(n)), dim3(KM_BLOCK), 0, s, n, NX, NY, coefT, r, res, zin,
                       zout, omega, skip);
    KM_CHECK_LAUNCH(ctx);
    return KM_OK;
}

template int apply<float>(km_ctx*, const double*, const float*, const float*, float*, float*, double*, const int*,
                          hipStream_t, double);
template int apply<double>(km_ctx*, const double*, const double*, const double*, double*, double*, double*,
                           const int*, hipStream_t, double);
template int jacobi<float>(km_ctx*, int, int, const float*, const float*, const float*, const float*, float*, double,
                           const int*, hipStream_t);
template int jacobi<double>(km_ctx*, int, int, const double*, const double*, const double*, const double*, double*,
                            double, const int*, hipStream_t);

}  // namespace kmcoarse

void km_coarse_free(km_ctx* ctx) {
    if (!ctx || !ctx->coarse) return;
    if (ctx->coarse->arena) (void)hipFree(ctx->coarse->arena);
    delete ctx->coarse;
    ctx->coarse = nullptr;
}

extern "C" {

int km_set_coarse(km_ctx* ctx, int block) {
    if (!ctx) return KM_EINVAL;
    KM_REQUIRE(ctx, block >= 0 && block <= 4096, "aggregate size %d out of range", block);
    get_coarse(ctx)->block_req = block;
    return KM_OK;
}

int km_coarse_correct(km_ctx* ctx, int NX, int NY, const double* coef, const double* r, double* z1, double* r1) {
    if (!ctx) return KM_EINVAL;
    KM_REQUIRE(ctx, coef && r && z1 && r1, "null pointer");
    KM_REQUIRE(ctx, NX >= 3 && NY >= 3 && (int64_t)NX * NY < ((int64_t)1 << 31), "system shape %d x %d", NX, NY);
    if (int rc = kmcoarse::setup(ctx, NX, NY, coef, ctx->stream)) return rc;
    return kmcoarse::apply<double>(ctx, coef, coef, r, z1, r1, nullptr, nullptr, ctx->stream, 0.0);
}

int km_coarse_info(km_ctx* ctx, int* block, int* unknowns) {
    if (!ctx || !block || !unknowns) return KM_EINVAL;
    KM_REQUIRE(ctx, ctx->coarse != nullptr, "no coarse level");
    *block = ctx->coarse->B;
    *unknowns = ctx->coarse->m;
    return KM_OK;
}

}  // extern "C"
