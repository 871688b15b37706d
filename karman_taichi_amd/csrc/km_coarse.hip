// Aggregation coarse level of the two-level p' preconditioner.
//
// The fine interior is cut into B x B aggregates; P is the indicator of the
// free cells (a link is non-zero) of each aggregate and the coarse operator is
// the Galerkin product A_c = P^T A P, again a 5-point operator on the
// aggregate grid.  A_c is factored (banded LU, bandwidth = aggregate columns)
// and inverted explicitly once per solve, so one application
//   z1 = P A_c^-1 P^T r,   r1 = r - A z1
// is a restriction, a small dense matvec and a prolongation fused with the
// residual.  The fine-level smoother (the network, or red-black SOR sweeps
// in the fused iteration of km_twolevel.hip) then acts
// on r1 and its output is added to z1.
#include "km_coarse.h"

#include <rocblas/rocblas.h>

#include <cmath>

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

// mask[k] = 1 on interior cells with a non-zero link
__global__ void __launch_bounds__(KM_BLOCK)
k_coarse_mask(int NX, int NY, const double* __restrict__ coef, uint8_t* __restrict__ mask) {
    const int64_t n = (int64_t)NX * NY;
    for (int64_t k = (int64_t)blockIdx.x * KM_BLOCK + threadIdx.x; k < n; k += (int64_t)gridDim.x * KM_BLOCK) {
        const int i = (int)(k / NY), j = (int)(k % NY);
        const bool interior = i > 0 && i < NX - 1 && j > 0 && j < NY - 1;
        mask[k] = interior && free_link(coef, n, k) ? 1 : 0;
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

// The whole band in LDS, one barrier per elimination step: step k reads
// column k (the multipliers l = A[i][k] / A[k][k]) and updates columns > k;
// the multipliers wait in a side buffer and are written over column k during
// step k + 1, which reads no column below k + 1.
__global__ void __launch_bounds__(64)
k_coarse_factor_lds(int m, int bw, double* __restrict__ band, const int* __restrict__ skip) {
    extern __shared__ double a[];
    if (*skip) return;
    const int W = 2 * bw + 1;
    double* pend = a + (size_t)m * W;   // [2][bw]: multipliers of the last two steps
    for (int e = threadIdx.x; e < m * W; e += blockDim.x) a[e] = band[e];
    __syncthreads();
    int prev_rows = 0;
    for (int k = 0; k < m; ++k) {
        const double* pl = pend + ((k + 1) & 1) * bw;   // step k - 1's multipliers, rows k .. k - 1 + prev_rows
        for (int t = threadIdx.x; t < prev_rows; t += blockDim.x) {
            const int i = k + t;
            a[(size_t)i * W + bw + (k - 1 - i)] = pl[t];
        }
        const int rows = (k + bw < m - 1 ? k + bw : m - 1) - k;   // rows k+1 .. k+rows
        const double* pk = a + (size_t)k * W + bw;                  // pk[o] = A[k][k + o]
        const double rp = 1.0 / pk[0];
        double* pn = pend + (k & 1) * bw;
        // thread t updates column j = k + 1 + t of rows k + 1 .. k + rows (bw <= 64)
        const int t = threadIdx.x;
        if (t < rows) {
            const int j = k + 1 + t;
            const double pkj = pk[j - k];
            for (int i = k + 1; i <= k + rows; ++i) {
                double* ri = a + (size_t)i * W + bw;                // ri[o] = A[i][i + o]
                const double l = ri[k - i] * rp;
                ri[j - i] -= l * pkj;
                if (t == 0) pn[i - k - 1] = l;
            }
        }
        __syncthreads();
        prev_rows = rows;
    }
    for (int e = threadIdx.x; e < m * W; e += blockDim.x) band[e] = a[e];
}

// Banded LU without pivoting (A_c is SPD) in one wave: step k updates the
// (bw x bw) window below and right of the pivot; L (unit) overwrites the
// strict lower band, U the rest.  The bw + 1 rows that steps k .. k + bw touch
// live in an LDS ring (row r in slot r mod (bw + 1)); row k goes back to the
// band once step k is done and row k + bw + 1, untouched so far, takes its slot.
__global__ void __launch_bounds__(64)
k_coarse_factor(int m, int bw, double* __restrict__ band, const int* __restrict__ skip) {
    extern __shared__ double ring[];
    if (*skip) return;
    const int W = 2 * bw + 1, R = bw + 1;
    for (int e = threadIdx.x; e < R * W; e += blockDim.x) {
        const int row = e / W;
        ring[e] = row < m ? band[(int64_t)row * W + (e - row * W)] : 0.0;
    }
    // the row entering the ring after step k, loaded during step k (thread e
    // holds entries e, e + 64, ...; W <= 4 * 64 when the ring fits 64 KiB)
    double inc[4];
    auto fetch = [&](int row) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = threadIdx.x + u * 64;
            inc[u] = row < m && e < W ? band[(int64_t)row * W + e] : 0.0;
        }
    };
    fetch(R);
    __syncthreads();
    for (int k = 0; k < m; ++k) {
        const int rows = (k + bw < m - 1 ? k + bw : m - 1) - k;   // rows k+1 .. k+rows
        const double* pk = ring + (k % R) * W + bw;                 // pk[o] = A[k][k + o]
        const double piv = pk[0];
        // thread t updates column j = k + 1 + t of rows k + 1 .. k + rows (bw <= 64)
        if ((int)threadIdx.x < rows) {
            const int j = k + 1 + threadIdx.x;
            const double pkj = pk[j - k];
            for (int i = k + 1; i <= k + rows; ++i) {
                double* ri = ring + (i % R) * W + bw;               // ri[o] = A[i][i + o]
                ri[j - i] -= ri[k - i] / piv * pkj;
            }
        }
        __syncthreads();
        for (int t = threadIdx.x; t < rows; t += blockDim.x) {
            const int i = k + 1 + t;
            ring[(i % R) * W + bw + (k - i)] /= piv;
        }
        __syncthreads();
        double* slot = ring + (k % R) * W;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = threadIdx.x + u * 64;
            if (e < W) {
                band[(int64_t)k * W + e] = slot[e];
                slot[e] = inc[u];
            }
        }
        fetch(k + R + 1);
        __syncthreads();
    }
}

// A_c^-1, row-major, from the band factors: thread c owns column c and all
// threads of a workgroup walk the rows together, so the factor entries of a
// row are the same for every thread (read from an LDS copy of CH rows at a
// time; the next CH rows are loaded into registers meanwhile) and the last
// BWT values of the column's substitution stay in registers (BWT >= bw; the
// LDS rows are padded to 2 BWT + 1 entries with zeros, so every row is read
// without a branch, and hold the reciprocal of U's diagonal).
// Forward: L y = e_c for rows from the workgroup's first column on (y = 0
// above it); back: U x = y, the rows in reverse, y read back from inv CH rows
// ahead.  The bw products of a row are summed as a tree.
template <int BWT>
__device__ __forceinline__ double tree_dot(const double* __restrict__ f, int stride, const double (&w)[BWT]) {
    double t[BWT];
#pragma unroll
    for (int d = 0; d < BWT; ++d) t[d] = f[d * stride] * w[d];
#pragma unroll
    for (int h = BWT / 2; h >= 1; h /= 2)
#pragma unroll
        for (int d = 0; d < h; ++d) t[d] += t[d + h];
    return t[0];
}

template <int BWT>
__global__ void __launch_bounds__(KM_BLOCK)
k_coarse_inverse(int m, int bw, const double* __restrict__ band, double* __restrict__ inv,
                 const int* __restrict__ skip) {
    constexpr int CH = 32;                        // rows per LDS chunk
    constexpr int WMAX = 2 * BWT + 1;
    constexpr int PF = (CH * WMAX + KM_BLOCK - 1) / KM_BLOCK;   // chunk values per thread
    __shared__ double fb[CH * WMAX];
    if (*skip) return;
    const int c0 = blockIdx.x * KM_BLOCK, c = c0 + threadIdx.x;
    const bool own = c < m;
    const int W = 2 * bw + 1;   // the band's row length
    double pf[PF];
    // rows [r0, r0 + CH) of the band into pf, padded to WMAX entries per row
    // (zero past the bandwidth and past m) / from pf into fb
    auto fetch = [&](int r0) {
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int e = threadIdx.x + u * KM_BLOCK;
            const int row = r0 + e / WMAX, o = e % WMAX - BWT;
            const bool in = e < CH * WMAX && row >= 0 && row < m && o >= -bw && o <= bw;
            const double val = in ? band[(int64_t)row * W + bw + o] : 0.0;
            pf[u] = o == 0 ? (val != 0.0 ? 1.0 / val : 0.0) : val;   // the diagonal as its reciprocal
        }
    };
    auto stage = [&]() {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int e = threadIdx.x + u * KM_BLOCK;
            if (e < CH * WMAX) fb[e] = pf[u];
        }
        __syncthreads();
    };
    double w[BWT];
#pragma unroll
    for (int d = 0; d < BWT; ++d) w[d] = 0.0;
    fetch(c0);
    for (int i0 = c0; i0 < m; i0 += CH) {   // w[d] = y[i - 1 - d]
        stage();
        fetch(i0 + CH);
        const int n = m - i0 < CH ? m - i0 : CH;
        for (int u = 0; u < n; ++u) {
            const int i = i0 + u;
            const double* row = fb + u * WMAX + BWT;   // row[o] = factor entry (i, i + o), |o| <= BWT
            const double v = (i == c ? 1.0 : 0.0) - tree_dot<BWT>(row - 1, -1, w);
#pragma unroll
            for (int d = BWT - 1; d > 0; --d) w[d] = w[d - 1];
            w[0] = v;
            if (own) inv[(int64_t)i * m + c] = v;
        }
    }
#pragma unroll
    for (int d = 0; d < BWT; ++d) w[d] = 0.0;
    const int top = ((m - 1) / CH) * CH;     // chunks [i0, i0 + CH) from the last one down
    double yv[CH];
    auto fetch_y = [&](int r0) {
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            const int i = r0 + u;
            yv[u] = own && i >= c0 && i < m ? inv[(int64_t)i * m + c] : 0.0;
        }
    };
    fetch(top);
    __syncthreads();   // the forward stores of this thread's column precede the reads below
    fetch_y(top);
    for (int i0 = top; i0 >= 0; i0 -= CH) {   // w[d] = x[i + 1 + d]
        stage();
        double yc_[CH];
#pragma unroll
        for (int u = 0; u < CH; ++u) yc_[u] = yv[u];
        if (i0 >= CH) {
            fetch(i0 - CH);
            fetch_y(i0 - CH);
        }
        const int n = m - i0 < CH ? m - i0 : CH;
#pragma unroll
        for (int u = CH - 1; u >= 0; --u) {
            if (u >= n) continue;
            const int i = i0 + u;
            const double* row = fb + u * WMAX + BWT;
            const double v = (yc_[u] - tree_dot<BWT>(row + 1, 1, w)) * row[0];
#pragma unroll
            for (int d = BWT - 1; d > 0; --d) w[d] = w[d - 1];
            w[0] = v;
            if (own) inv[(int64_t)i * m + c] = v;
        }
    }
}

// The same for any bandwidth: one thread per column with the substitution in
// the output row, each step reading its own factor entries.
__global__ void __launch_bounds__(KM_BLOCK)
k_coarse_inverse_wide(int m, int bw, const double* __restrict__ band, double* __restrict__ x,
                      const int* __restrict__ skip) {
    if (*skip) return;
    const int c = blockIdx.x * KM_BLOCK + threadIdx.x;
    if (c >= m) return;
    const int W = 2 * bw + 1;
    auto X = [&](int i) -> double& { return x[(int64_t)i * m + c]; };
    for (int i = 0; i < c; ++i) X(i) = 0.0;
    for (int i = c; i < m; ++i) {   // L y = e_c (unit lower)
        double v = i == c ? 1.0 : 0.0;
        const int k0 = i - bw > c ? i - bw : c;
        for (int k = k0; k < i; ++k) v -= band[(int64_t)i * W + bw + (k - i)] * X(k);
        X(i) = v;
    }
    for (int i = m - 1; i >= 0; --i) {   // U x = y
        double v = X(i);
        const int k1 = i + bw < m - 1 ? i + bw : m - 1;
        for (int k = i + 1; k <= k1; ++k) v -= band[(int64_t)i * W + bw + (k - i)] * X(k);
        X(i) = v / band[(int64_t)i * W + bw];
    }
}

// Newton-Schulz update of the held inverse X for the new A_c: one wave per
// row i forms (A_c X)[i][:] from the 5-point coarse coefficients, writes
// M = 2 I - A_c X and the row sum of |I - A_c X| (its infinity norm, row by
// row); X (2 I - A_c X) then has error at most the square of that norm.
__global__ void __launch_bounds__(KM_BLOCK)
k_coarse_ns_residual(int m, int mJ, const double* __restrict__ cc, const double* __restrict__ X,
                     double* __restrict__ M, double* __restrict__ rowerr) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= m) return;
    const double cp = cc[i], cw = cc[m + i], ce = cc[2 * m + i], cn = cc[3 * m + i], cs = cc[4 * m + i];
    const double* x0 = X + (int64_t)i * m;
    double e = 0.0;
    for (int j = lane; j < m; j += 64) {
        double v = cp * x0[j];
        if (i >= mJ) v += cw * x0[j - (int64_t)mJ * m];
        if (i + mJ < m) v += ce * x0[j + (int64_t)mJ * m];
        if (i + 1 < m) v += cn * x0[j + m];
        if (i > 0) v += cs * x0[j - m];
        const double d = (i == j ? 1.0 : 0.0) - v;
        M[(int64_t)i * m + j] = d + (i == j ? 1.0 : 0.0);
        e += fabs(d);
    }
    e = wave_sum(e);
    if (lane == 0) rowerr[i] = e;
}

// ok = 1 when max_i rowerr[i] < tol (the update is kept and the
// factorisation skipped), else 0; NaN rows count as too large
__global__ void __launch_bounds__(KM_BLOCK)
k_coarse_ns_check(int m, const double* __restrict__ rowerr, double tol, int* __restrict__ ok) {
    __shared__ double sh[4];
    double v = 0.0;
    for (int i = threadIdx.x; i < m; i += KM_BLOCK) {
        const double e = rowerr[i];
        v = fmax(v, e == e ? e : 1e300);
    }
    v = block_max(v, sh);
    if (threadIdx.x == 0) *ok = v < tol ? 1 : 0;
}

__global__ void __launch_bounds__(KM_BLOCK)
k_coarse_ns_accept(int64_t n, const double* __restrict__ Xn, double* __restrict__ X, const int* __restrict__ ok) {
    if (!*ok) return;
    for (int64_t k = (int64_t)blockIdx.x * KM_BLOCK + threadIdx.x; k < n; k += (int64_t)gridDim.x * KM_BLOCK)
        X[k] = Xn[k];
}

__global__ void k_coarse_flag(int* __restrict__ f, int v) {
    if (threadIdx.x == 0) *f = v;
}

// rc[agg] = sum of r over the aggregate's free cells (one workgroup per aggregate)
template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_coarse_restrict(int NX, int NY, int B, int mJ, const uint8_t* __restrict__ mask, const T* __restrict__ r,
                  double* __restrict__ rc, const int* __restrict__ skip) {
    if (skip && *skip) return;
    __shared__ double sh[4];
    const int agg = blockIdx.x;
    const int I = agg / mJ, J = agg - I * mJ;
    const int nx = NX - 2, ny = NY - 2;
    const int i0 = 1 + I * B, j0 = 1 + J * B;
    const int bi = (i0 + B - 1 < nx ? i0 + B - 1 : nx) - i0 + 1, bj = (j0 + B - 1 < ny ? j0 + B - 1 : ny) - j0 + 1;
    double a = 0.0;
    for (int t = threadIdx.x; t < bi * bj; t += KM_BLOCK) {
        const int64_t k = (int64_t)(i0 + t / bj) * NY + (j0 + t % bj);
        if (mask[k]) a += (double)r[k];
    }
    a = block_sum(a, sh);
    if (threadIdx.x == 0) rc[agg] = a;
}

// yc = inv rc, rc[k] = sum over the S slices of rc[s m + k]: the workgroup
// first sums the slices into LDS (all loads in flight: slices past S re-read
// the last one and are masked), then one wave per row.
__global__ void __launch_bounds__(KM_BLOCK)
k_coarse_matvec(int m, int S, const double* __restrict__ inv, const double* __restrict__ rc, double* __restrict__ yc,
                const int* __restrict__ skip) {
    extern __shared__ double rs_[];
    if (skip && *skip) return;
    // one slice: read rc directly (L2), no staging; else sum the slices in LDS
    const double* rs = S == 1 ? rc : rs_;
    for (int k = threadIdx.x; S > 1 && k < m; k += KM_BLOCK) {
        double v[kmcoarse::kMaxSlices];
#pragma unroll
        for (int q = 0; q < kmcoarse::kMaxSlices; ++q) v[q] = rc[(int64_t)(q < S ? q : S - 1) * m + k];
        double a = v[0];
#pragma unroll
        for (int q = 1; q < kmcoarse::kMaxSlices; ++q) a += q < S ? v[q] : 0.0;
        rs_[k] = a;
    }
    if (S > 1) __syncthreads();
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (row >= m) return;
    const double* a = inv + (int64_t)row * m;
    double s = 0.0;
    // eight row loads in flight per lane, summed in k order
    for (int k0 = lane; k0 < m; k0 += 8 * 64) {
        double av[8], rv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const bool in = k0 + 64 * u < m;
            av[u] = in ? a[k0 + 64 * u] : 0.0;
            rv[u] = in ? rs[k0 + 64 * u] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (k0 + 64 * u < m) s += av[u] * rv[u];
    }
    s = wave_sum(s);
    if (lane == 0) yc[row] = s;
}

__device__ __forceinline__ double coarse_value(int i, int j, int B, int mJ, int nx, int ny, const double* yc) {
    if (i < 1 || i > nx || j < 1 || j > ny) return 0.0;
    return yc[((i - 1) / B) * mJ + (j - 1) / B];
}

// z1 = P yc on free cells; r1 = r - A z1; partials of r1.r1 (P_rr, may be null)
template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_coarse_prolong(Flat fl, int NX, int NY, int B, int mJ, const double* __restrict__ coef,
                 const T* __restrict__ coefT, const T* __restrict__ r, const double* __restrict__ yc,
                 T* __restrict__ z1, T* __restrict__ r1, double* __restrict__ Prr, const int* __restrict__ skip) {
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
        z1[k] = zP;
        r1[k] = rv;
        a += (double)rv * (double)rv;
    }
    if (Prr) {
        a = block_sum(a, sh);
        if (threadIdx.x == 0) Prr[blockIdx.x] = a;
    }
}

int flat_blocks2(int64_t n) {
    int64_t g = (n + KM_BLOCK - 1) / KM_BLOCK;
    if (g > 2048) g = 2048;
    return g < 1 ? 1 : (int)g;
}

size_t a256(size_t v) { return (v + 255) & ~size_t(255); }

// Xn = X M for row-major m x m X, M: the column-major product M^T X^T
// (a plain library GEMM)
int ns_gemm(km_ctx* ctx, int m, const double* M, const double* X, double* Xn, hipStream_t s) {
    KmHostLock lk(km_host_mutex());   // handle creation and the first GEMM allocate and load code
    if (!ctx->blas) {
        rocblas_handle h = nullptr;
        if (rocblas_create_handle(&h) != rocblas_status_success)
            return km_fail(ctx, KM_EHIP, "rocblas_create_handle failed");
        ctx->blas = h;
    }
    rocblas_handle h = static_cast<rocblas_handle>(ctx->blas);
    if (rocblas_set_stream(h, s) != rocblas_status_success) return km_fail(ctx, KM_EHIP, "rocblas_set_stream failed");
    const double one = 1.0, zero = 0.0;
    const rocblas_status st = rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, m, m, m, &one, M, m, X,
                                            m, &zero, Xn, m);
    if (st != rocblas_status_success) return km_fail(ctx, KM_EHIP, "rocblas_dgemm: status %d", (int)st);
    return KM_OK;
}

km_coarse* get_coarse(km_ctx* ctx) {
    if (!ctx->coarse) ctx->coarse = new km_coarse();
    return ctx->coarse;
}

// The settings a new level takes over from the current one.
void copy_settings(const km_coarse* from, km_coarse* to) {
    to->block_req = from->block_req;
    to->update = from->update;
    to->totals = from->totals;
    to->chunk_big = from->chunk_big;
    to->chunk_small = from->chunk_small;
    to->chunk_lead = from->chunk_lead;
    to->tiles = from->tiles;
    to->lds_max = from->lds_max;
}

constexpr size_t kMaxLevels = 128;

// The level of a fused solve of `coef` (NX x NY): each coefficient array
// keeps its own, so the cases of a batch each keep their held inverse across
// outer steps; past kMaxLevels the least recently used level is recycled.
km_coarse* level_for(km_ctx* ctx, int NX, int NY, const double* coef) {
    km_coarse* cur = get_coarse(ctx);
    auto& pool = ctx->coarse_pool;
    if (pool.empty()) pool.push_back(cur);
    const uint64_t tick = ++ctx->coarse_tick;
    km_coarse* pick = nullptr;
    for (km_coarse* C : pool)
        if (C->arena && C->coef == coef && C->NX == NX && C->NY == NY) pick = C;
    if (!pick)
        for (km_coarse* C : pool)
            if (!C->arena) pick = C;   // a level that never ran
    if (!pick && pool.size() < kMaxLevels) {
        pick = new km_coarse();
        copy_settings(cur, pick);
        pool.push_back(pick);
    }
    if (!pick) {
        pick = pool[0];
        for (km_coarse* C : pool)
            if (C->used < pick->used) pick = C;
    }
    pick->used = tick;
    ctx->coarse = pick;
    return pick;
}


}  // namespace

namespace kmcoarse {

int choose_block(int nx, int ny, int req) {
    if (req > 0) return req;
    for (int B : {8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256}) {
        const int64_t mJ = (ny + B - 1) / B, m = (int64_t)((nx + B - 1) / B) * mJ;
        if (mJ > kMaxBandwidth) continue;   // the factorisation's LDS ring
        if (m <= 2048 || (m <= 8192 && 4 * m * m <= 21 * (int64_t)nx * ny)) return B;
    }
    return 512;
}

void for_each_level(km_ctx* ctx, const std::function<void(km_coarse*)>& f) {
    km_coarse* cur = get_coarse(ctx);
    bool seen = false;
    for (km_coarse* C : ctx->coarse_pool) {
        f(C);
        seen = seen || C == cur;
    }
    if (!seen) f(cur);
}

int setup(km_ctx* ctx, int NX, int NY, const double* coef, hipStream_t s, bool solver) {
    km_coarse* C = solver ? level_for(ctx, NX, NY, coef) : get_coarse(ctx);
    const int nx = NX - 2, ny = NY - 2;
    KM_REQUIRE(ctx, nx >= 1 && ny >= 1, "system too small for a coarse level");
    // the smoothing tiles address vectors with 32-bit byte offsets
    KM_REQUIRE(ctx, (int64_t)NX * NY < ((int64_t)1 << 29), "system of %lld cells is too large for the tile kernels",
               (long long)NX * NY);
    const int B = choose_block(nx, ny, C->block_req);
    const int mI = (nx + B - 1) / B, mJ = (ny + B - 1) / B, m = mI * mJ;
    KM_REQUIRE(ctx, m <= 8192, "coarse level of %d unknowns is too large (aggregate %d)", m, B);
    const int bw = mJ;
    // the factorisation's LDS ring: (bw + 1) rows of 2 bw + 1
    KM_REQUIRE(ctx, (size_t)(bw + 1) * (2 * bw + 1) * sizeof(double) <= 64 * 1024,
               "coarse band of width %d is too wide (aggregate %d)", bw, B);
    const int64_t n = (int64_t)NX * NY;
    const size_t mm = (size_t)m * m * 8;
    const size_t need = a256(5 * (size_t)m * 8) + a256((size_t)m * (2 * bw + 1) * 8) + 3 * a256(mm) +
                        a256((size_t)kMaxSlices * m * 8) + 2 * a256((size_t)m * 8) + a256((size_t)4 * m * 8) +
                        a256((size_t)2 * kMaxVS * m * 8) + a256((size_t)n) + 256;
    if (!C->arena || C->bytes < need) {
        KM_HIP_SYNC(ctx, hipStreamSynchronize(s));
        km_drop_graphs(ctx);
        if (C->arena) KM_HIP_SYNC(ctx, hipFree(C->arena));
        C->arena = nullptr;
        void* p = nullptr;
        hipError_t e;
        {
            KmHostLock lk(km_host_mutex());
            e = hipMalloc(&p, need);
        }
        if (e != hipSuccess) return km_fail(ctx, KM_ENOMEM, "coarse hipMalloc(%zu): %s", need, hipGetErrorString(e));
        C->arena = static_cast<char*>(p);
        C->bytes = need;
        C->held = false;
    }
    char* q = C->arena;
    C->cc = reinterpret_cast<double*>(q); q += a256(5 * (size_t)m * 8);
    C->band = reinterpret_cast<double*>(q); q += a256((size_t)m * (2 * bw + 1) * 8);
    C->inv = reinterpret_cast<double*>(q); q += a256(mm);
    double* nsM = reinterpret_cast<double*>(q); q += a256(mm);
    double* nsX = reinterpret_cast<double*>(q); q += a256(mm);
    C->rc = reinterpret_cast<double*>(q); q += a256((size_t)kMaxSlices * m * 8);
    C->yc = reinterpret_cast<double*>(q); q += a256((size_t)m * 8);
    double* rowerr = reinterpret_cast<double*>(q); q += a256((size_t)m * 8);
    C->fy = reinterpret_cast<double*>(q); q += a256((size_t)4 * m * 8);
    C->frc = reinterpret_cast<double*>(q); q += a256((size_t)2 * kMaxVS * m * 8);
    int* ok = reinterpret_cast<int*>(q); q += 256;
    C->mask = reinterpret_cast<uint8_t*>(q);
    // the held inverse belongs to a coarse level of the same shape: try the update
    const bool same = C->held && C->NX == NX && C->NY == NY && C->B == B;
    if (!(C->NX == NX && C->NY == NY && C->B == B))   // another system shape
        for (int& c : C->recent) c = 0;
    // a solve of the system set up last time (same coefficient array) may run
    // with a lagged coarse inverse: the flexible solve only needs an
    // approximate one
    const bool lagged = solver && same && coef == C->coef && C->update > 1 && !C->refresh;
    if (lagged && C->since + 1 < C->update) {
        C->since++;   // reuse the held level as it is
        return KM_OK;
    }
    C->since = 0;
    C->coef = coef;
    C->refresh = false;
    const bool update = same && C->update > 0;
    C->NX = NX; C->NY = NY; C->B = B; C->mI = mI; C->mJ = mJ; C->m = m; C->bw = bw;
    hipLaunchKernelGGL(k_coarse_mask, dim3(flat_blocks2(n)), dim3(KM_BLOCK), 0, s, NX, NY, coef, C->mask);
    hipLaunchKernelGGL(k_coarse_coef, dim3(m), dim3(KM_BLOCK), 0, s, NX, NY, B, mI, mJ, coef, C->cc);
    hipLaunchKernelGGL(k_coarse_band, dim3(flat_blocks2((int64_t)m * (2 * bw + 1))), dim3(1024), 0, s, m, mJ, bw,
                       (const double*)C->cc, C->band);
    if (update) {
        // X <- X (2 I - A_c X) with the previous solve's X, kept when
        // ||I - A_c X||_inf < kNsTol (then the error is below kNsTol^2)
        hipLaunchKernelGGL(k_coarse_ns_residual, dim3((m + 3) / 4), dim3(KM_BLOCK), 0, s, m, mJ, (const double*)C->cc,
                           (const double*)C->inv, nsM, rowerr);
        hipLaunchKernelGGL(k_coarse_ns_check, dim3(1), dim3(KM_BLOCK), 0, s, m, (const double*)rowerr, kNsTol, ok);
        if (int rc = ns_gemm(ctx, m, nsM, C->inv, nsX, s)) return rc;
        hipLaunchKernelGGL(k_coarse_ns_accept, dim3(flat_blocks2((int64_t)m * m)), dim3(KM_BLOCK), 0, s,
                           (int64_t)m * m, (const double*)nsX, C->inv, (const int*)ok);
    } else {
        hipLaunchKernelGGL(k_coarse_flag, dim3(1), dim3(64), 0, s, ok, 0);
    }
    // one wave (the ring's register prefetch assumes 64 threads): a step
    // updates at most bw x bw entries
    const size_t lds_bytes = ((size_t)m * (2 * bw + 1) + 2 * (size_t)bw) * sizeof(double);
    if (C->lds_max == 0) {
        int dev = 0, v = 0;
        KM_HIP(ctx, hipGetDevice(&dev));
        KM_HIP(ctx, hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        C->lds_max = v > 0 ? (size_t)v : 64 * 1024;
    }
    // the factorisation and explicit inverse run unless the update was kept;
    // a lagged level keeps its inverse when the update is rejected
    const int* skip = ok;
    if (lagged) {
        C->held = true;
        return KM_OK;
    }
    if (lds_bytes <= C->lds_max)
        hipLaunchKernelGGL(k_coarse_factor_lds, dim3(1), dim3(64), lds_bytes, s, m, bw, C->band, skip);
    else
        hipLaunchKernelGGL(k_coarse_factor, dim3(1), dim3(64), (size_t)(bw + 1) * (2 * bw + 1) * sizeof(double), s,
                           m, bw, C->band, skip);
    const dim3 gi((m + KM_BLOCK - 1) / KM_BLOCK);
    const double* band = C->band;
    if (bw <= 8)
        hipLaunchKernelGGL(k_coarse_inverse<8>, gi, dim3(KM_BLOCK), 0, s, m, bw, band, C->inv, skip);
    else if (bw <= 16)
        hipLaunchKernelGGL(k_coarse_inverse<16>, gi, dim3(KM_BLOCK), 0, s, m, bw, band, C->inv, skip);
    else if (bw <= 32)
        hipLaunchKernelGGL(k_coarse_inverse<32>, gi, dim3(KM_BLOCK), 0, s, m, bw, band, C->inv, skip);
    else
        hipLaunchKernelGGL(k_coarse_inverse_wide, gi, dim3(KM_BLOCK), 0, s, m, bw, band, C->inv, skip);
    KM_CHECK_LAUNCH(ctx);
    // the first factorisation also runs one update product into scratch, so
    // that rocBLAS loads its kernels here rather than in a later solve
    if (!update && !ctx->blas && C->update > 0)
        if (int rc = ns_gemm(ctx, m, C->inv, C->inv, nsX, s)) return rc;
    C->held = true;
    return KM_OK;
}

int coarse_solve(km_ctx* ctx, int S, const int* skip, hipStream_t s) {
    km_coarse* C = ctx->coarse;
    KM_REQUIRE(ctx, C && C->arena && C->m > 0, "coarse level not set up");
    KM_REQUIRE(ctx, S >= 1 && S <= kMaxSlices, "restriction slices %d", S);
    hipLaunchKernelGGL(k_coarse_matvec, dim3((C->m + 3) / 4), dim3(KM_BLOCK), S > 1 ? (size_t)C->m * sizeof(double) : 0,
                       s, C->m, S, (const double*)C->inv, (const double*)C->rc, C->yc, skip);
    KM_CHECK_LAUNCH(ctx);
    return KM_OK;
}

template <typename T>
int restrict_solve(km_ctx* ctx, const T* r, const int* skip, hipStream_t s) {
    km_coarse* C = ctx->coarse;
    KM_REQUIRE(ctx, C && C->arena && C->m > 0, "coarse level not set up");
    hipLaunchKernelGGL(k_coarse_restrict<T>, dim3(C->m), dim3(KM_BLOCK), 0, s, C->NX, C->NY, C->B, C->mJ,
                       (const uint8_t*)C->mask, r, C->rc, skip);
    KM_CHECK_LAUNCH(ctx);
    return coarse_solve(ctx, 1, skip, s);
}

template <typename T>
int apply(km_ctx* ctx, const double* coef, const T* coefT, const T* r, T* z1, T* r1, double* Prr, const int* skip,
          hipStream_t s) {
    if (int rc = restrict_solve<T>(ctx, r, skip, s)) return rc;
    km_coarse* C = ctx->coarse;
    const Flat fl = make_flat((int64_t)C->NX * C->NY);
    hipLaunchKernelGGL(k_coarse_prolong<T>, dim3(fl.G), dim3(KM_BLOCK), 0, s, fl, C->NX, C->NY, C->B, C->mJ, coef,
                       coefT, r, (const double*)C->yc, z1, r1, Prr, skip);
    KM_CHECK_LAUNCH(ctx);
    return KM_OK;
}

template int restrict_solve<float>(km_ctx*, const float*, const int*, hipStream_t);
template int restrict_solve<double>(km_ctx*, const double*, const int*, hipStream_t);
template int apply<float>(km_ctx*, const double*, const float*, const float*, float*, float*, double*, const int*,
                          hipStream_t);
template int apply<double>(km_ctx*, const double*, const double*, const double*, double*, double*, double*,
                           const int*, hipStream_t);

}  // namespace kmcoarse

void km_coarse_free(km_ctx* ctx) {
    if (!ctx) return;
    KmHostLock lk(km_host_mutex());
    if (ctx->blas) (void)rocblas_destroy_handle(static_cast<rocblas_handle>(ctx->blas));
    ctx->blas = nullptr;
    bool cur_in_pool = false;
    for (km_coarse* C : ctx->coarse_pool) {
        cur_in_pool = cur_in_pool || C == ctx->coarse;
        if (C->arena) (void)hipFree(C->arena);
        delete C;
    }
    ctx->coarse_pool.clear();
    if (ctx->coarse && !cur_in_pool) {
        if (ctx->coarse->arena) (void)hipFree(ctx->coarse->arena);
        delete ctx->coarse;
    }
    ctx->coarse = nullptr;
}

extern "C" {

int km_set_coarse(km_ctx* ctx, int block) {
    if (!ctx) return KM_EINVAL;
    KM_REQUIRE(ctx, block >= 0 && block <= 4096, "aggregate size %d out of range", block);
    kmcoarse::for_each_level(ctx, [&](km_coarse* C) { C->block_req = block; });
    return KM_OK;
}

int km_coarse_correct(km_ctx* ctx, int NX, int NY, const double* coef, const double* r, double* z1, double* r1) {
    if (!ctx) return KM_EINVAL;
    KM_REQUIRE(ctx, coef && r && z1 && r1, "null pointer");
    KM_REQUIRE(ctx, NX >= 3 && NY >= 3 && (int64_t)NX * NY < ((int64_t)1 << 31), "system shape %d x %d", NX, NY);
    if (int rc = kmcoarse::setup(ctx, NX, NY, coef, ctx->stream)) return rc;
    return kmcoarse::apply<double>(ctx, coef, coef, r, z1, r1, nullptr, nullptr, ctx->stream);
}

int km_set_coarse_update(km_ctx* ctx, int every) {
    if (!ctx) return KM_EINVAL;
    KM_REQUIRE(ctx, every >= 0, "update interval %d", every);
    kmcoarse::for_each_level(ctx, [&](km_coarse* C) {
        C->update = every;
        C->since = 0;
    });
    return KM_OK;
}

int km_coarse_info(km_ctx* ctx, int* block, int* unknowns) {
    if (!ctx || !block || !unknowns) return KM_EINVAL;
    KM_REQUIRE(ctx, ctx->coarse != nullptr, "no coarse level");
    *block = ctx->coarse->B;
    *unknowns = ctx->coarse->m;
    return KM_OK;
}

}  // extern "C"
