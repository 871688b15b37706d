// Krylov solvers on the implicit 5-point stencil for gfx950.
//
// The reference runs one BiCGSTAB iteration as 14 kernel launches and 5 host
// round trips (bicgsolver.py:88-136).  Here an iteration is three kernels:
//
//   A  p' = r + beta (p - omega v) on an LDS tile + halo, v' = A p',  <r~, v'>
//   B  s  = r - alpha v'           on an LDS tile + halo, t  = A s,   <t, s>, <t, t>
//   C  x += alpha p' + omega s,  r = s - omega t,                     <r, r>, <r, r~>
//
// Each kernel leaves per-workgroup partial sums; the next kernel reduces them
// in every workgroup (same order everywhere, so every workgroup derives the
// same FP64 scalar) and workgroup 0 records the scalars, the stopping test and
// the iteration count in device memory.  Kernel A of iteration i judges
// iteration i-1 (the reference's test right after update_r).  Kernels return
// at once when the device-side ``done`` flag is set, so the host enqueues
// fixed-size chunks of iterations (replayed from a captured hipGraph) and
// polls between chunks instead of after every reduction.  p and v are
// double-buffered by iteration parity because kernel A reads old values in
// its halo while writing new ones.
//
// T = double (plain FP64 path) or float (inner solves of the mixed-precision
// path); dot products and scalars are always FP64.
#include "km_common.h"
#include "km_coarse.h"

extern __shared__ __align__(16) unsigned char km_smem[];

namespace {

struct TileOrigin {
    int i0, j0;
};

__device__ __forceinline__ TileOrigin tile_origin(const Tile& tl, int tile) {
    const int ti = tile / tl.tiles_j, tj = tile - ti * tl.tiles_j;
    return {ti * tl.TI, tj * tl.TJ};
}

// Stage f(k) for the tile plus a one-cell halo into LDS; zero outside the
// array (the reference's halo reads land on zero coefficients,
// cgsolver.py:59-60).  If ``center`` is non-null the tile's own values are
// also stored there.
template <typename T, typename F>
__device__ __forceinline__ void stage(const Tile& tl, const TileOrigin& o, T* lds, F f, T* center) {
    const int W = tl.TJ + 2, H = tl.TI + 2;
    for (int t = threadIdx.x; t < W * H; t += KM_BLOCK) {
        const int li = t / W, lj = t - li * W;
        const int gi = o.i0 - 1 + li, gj = o.j0 - 1 + lj;
        T val = T(0);
        if (gi >= 0 && gi < tl.NX && gj >= 0 && gj < tl.NY) {
            const int64_t k = (int64_t)gi * tl.NY + gj;
            val = f(k);
            if (center && li >= 1 && li <= tl.TI && lj >= 1 && lj <= tl.TJ) center[k] = val;
        }
        lds[t] = val;
    }
}

// y = c0 x + c1 x_W + c2 x_E + c3 x_N + c4 x_S from the staged tile, summed
// in the reference's order (bicgsolver.py:68-69); g(k, y, x_centre) consumes it.
template <typename T, typename G>
__device__ __forceinline__ void apply(const Tile& tl, const TileOrigin& o, const T* lds,
                                      const T* __restrict__ coef, G g) {
    const int W = tl.TJ + 2;
    const int64_t N = (int64_t)tl.NX * tl.NY;
    for (int t = threadIdx.x; t < tl.TI * tl.TJ; t += KM_BLOCK) {
        const int li = t / tl.TJ, lj = t - li * tl.TJ;
        const int gi = o.i0 + li, gj = o.j0 + lj;
        if (gi < tl.NX && gj < tl.NY) {
            const int64_t k = (int64_t)gi * tl.NY + gj;
            const T* P = lds + (li + 1) * W + (lj + 1);
            const T y = coef[k] * P[0] + coef[N + k] * P[-W] + coef[2 * N + k] * P[W] +
                        coef[3 * N + k] * P[1] + coef[4 * N + k] * P[-1];
            g(k, y, P[0]);
        }
    }
}

// Run stage + apply over this workgroup's tiles (XCD-aware: each XCD owns a
// contiguous range of tiles, so neighbouring rows share its L2).
template <typename T, typename F, typename G>
__device__ __forceinline__ void tile_sweep(const Tile& tl, T* lds, F f, T* center, const T* __restrict__ coef, G g) {
    const int L = xcd_remap(blockIdx.x, tl.G);
    const int t0 = L * tl.tpb;
    const int t1 = t0 + tl.tpb < tl.T ? t0 + tl.tpb : tl.T;
    for (int tile = t0; tile < t1; ++tile) {
        const TileOrigin o = tile_origin(tl, tile);
        stage<T>(tl, o, lds, f, center);
        __syncthreads();
        apply<T>(tl, o, lds, coef, g);
        __syncthreads();
    }
}

template <typename F>
__device__ __forceinline__ void flat_sweep_b(const Flat& fl, int bx, F f) {
    const int64_t k0 = (int64_t)bx * fl.chunk;
    const int64_t k1 = k0 + fl.chunk < fl.n ? k0 + fl.chunk : fl.n;
    for (int64_t k = k0 + threadIdx.x; k < k1; k += KM_BLOCK) f(k);
}

template <typename F>
__device__ __forceinline__ void flat_sweep(const Flat& fl, F f) {
    const int64_t k0 = (int64_t)blockIdx.x * fl.chunk;
    const int64_t k1 = k0 + fl.chunk < fl.n ? k0 + fl.chunk : fl.n;
    for (int64_t k = k0 + threadIdx.x; k < k1; k += KM_BLOCK) f(k);
}

// Partials of the workgroup that runs this code: slot = xcd-remapped index for
// tiled launches, blockIdx for flat ones.
__device__ __forceinline__ void write_partial(double* P, int slot, double v, double* sh) {
    const double s = block_sum_all(v, sh);
    if (threadIdx.x == 0) P[slot] = s;
}

// Two partial sums of the workgroup with one pair of barriers (sh: 8 doubles):
// P[slot] = sum a, P[KM_MAX_RED + slot] = sum b, each in block_sum_all's order.
__device__ __forceinline__ void write_partials2(double* P, int slot, double a, double b, double* sh) {
    a = wave_sum(a);
    b = wave_sum(b);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) {
        sh[w] = a;
        sh[4 + w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        P[slot] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
        P[KM_MAX_RED + slot] = (sh[4] + sh[5]) + (sh[6] + sh[7]);
    }
}

__device__ __forceinline__ bool finite(double x) { return x == x && x - x == 0.0; }

__device__ __forceinline__ bool lead() { return blockIdx.x == 0 && threadIdx.x == 0; }

__device__ __forceinline__ void finish(KState* st, int status, int iter, double rTr) {
    st->status = status;
    st->iter = iter;
    st->rTr = rTr;
    st->done = 1;
}

// ------------------------------------------------------------ plain kernels

template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_stencil(Tile tl, const T* __restrict__ coef, const T* __restrict__ x, T* __restrict__ y) {
    T* lds = reinterpret_cast<T*>(km_smem);
    tile_sweep<T>(tl, lds, [&](int64_t k) { return x[k]; }, (T*)nullptr, coef, [&](int64_t k, T v, T) { y[k] = v; });
}

// r = b - A x (FP64) with per-workgroup partials of ||r||^2
__global__ void __launch_bounds__(KM_BLOCK)
k_residual(Tile tl, const double* __restrict__ coef, const double* __restrict__ b, const double* __restrict__ x,
           double* __restrict__ r, double* __restrict__ P) {
    double* lds = reinterpret_cast<double*>(km_smem);
    __shared__ double sh[4];
    double acc = 0.0;
    tile_sweep<double>(tl, lds, [&](int64_t k) { return x[k]; }, (double*)nullptr, coef,
                       [&](int64_t k, double y, double) {
                           const double rr = b[k] - y;
                           if (r) r[k] = rr;
                           acc += rr * rr;
                       });
    write_partial(P, xcd_remap(blockIdx.x, tl.G), acc, sh);
}

template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_dot(Flat fl, const T* __restrict__ a, const T* __restrict__ b, double* __restrict__ P) {
    __shared__ double sh[4];
    double acc = 0.0;
    flat_sweep(fl, [&](int64_t k) { acc += (double)a[k] * (double)b[k]; });
    write_partial(P, blockIdx.x, acc, sh);
}

// One workgroup: out[q] = sum of P[q * KM_MAX_RED + k], k < G.
template <int NV>
__global__ void __launch_bounds__(KM_BLOCK) k_finish(const double* __restrict__ P, int G, double* __restrict__ out) {
    __shared__ double sh[4];
    double s[NV];
    sum_partials<NV>(P, G, s, sh);
    if (threadIdx.x == 0)
        for (int q = 0; q < NV; ++q) out[q] = s[q];
}

__global__ void __launch_bounds__(KM_BLOCK)
k_to_float(int64_t n, const double* __restrict__ src, float* __restrict__ dst, double scale) {
    for (int64_t k = (int64_t)blockIdx.x * KM_BLOCK + threadIdx.x; k < n; k += (int64_t)gridDim.x * KM_BLOCK)
        dst[k] = (float)(src[k] * scale);
}

__global__ void __launch_bounds__(KM_BLOCK)
k_add_scaled(int64_t n, double* __restrict__ x, const float* __restrict__ d, double scale) {
    for (int64_t k = (int64_t)blockIdx.x * KM_BLOCK + threadIdx.x; k < n; k += (int64_t)gridDim.x * KM_BLOCK)
        x[k] = x[k] + scale * (double)d[k];
}

// ------------------------------------------------------------ BiCGSTAB
//
// Three kernels per iteration i (bicgsolver.py:94-135), with rho formed as
// the reference forms it, <r_i, r~> (bicgsolver.py:95):
//
//   P(i)  from R(i-1)'s partials <r_i, r_i> and <r_i, r~>: the stopping test
//         on ||r_i|| (the reference's test at the end of iteration i-1,
//         :128-131), the omega == 0 test (karman_taichi.py:451-454), the
//         max-iteration cap, the rho == 0 breakdown (:96-98);
//         beta = (rho_i / rho_{i-1}) (alpha / omega) (:102);
//         on the tile + halo: p_i = r_i + beta (p - omega v) (:43-46);
//         v_i = A p_i (:72-76).  Partials <r~, v_i> (:107).
//   S(i)  alpha_i = rho_i / <r~, v_i> (:108); on the tile + halo:
//         s = r_i - alpha v_i (:60-63); t = A s (:65-69).  Partials <t, s>,
//         <t, t> (:116-117).
//   R(i)  omega_i = <t,s> / <t,t> (:118); x += alpha p + omega s (:78-81);
//         r_{i+1} = s - omega t (:83-86); flat, no halo.  Partials
//         <r, r> (:127) and <r, r~>.
//
// The init kernel writes R's partials for r_0.  Ring slot of the iteration's
// phase: rho_i, beta_i and the iteration index (P), alpha_i (S), omega_i
// (R); a launch reads earlier slots and the fields other launches wrote.

// bicgsolver.py:21-35: x = 0, r = r~ = b (or r precomputed from x0), the
// parity-0 p and v buffers zeroed; partials of <r, r> and <r, r~> = <r, r>
// in R's planes.
template <typename T>
__device__ __forceinline__ void bicg_init_body(const Flat& fl, int bx, const T* __restrict__ b, T* __restrict__ x,
                                               T* __restrict__ r, T* __restrict__ rt, T* __restrict__ p0,
                                               T* __restrict__ v0, int from_r, double eps, int max_iter,
                                               double ref_norm2, KState* st, double* __restrict__ P) {
    __shared__ double sh[4];
    double rr = 0.0;
    flat_sweep_b(fl, bx, [&](int64_t k) {
        T rk;
        if (from_r) {
            rk = r[k];
        } else {
            rk = b[k];
            r[k] = rk;
            x[k] = T(0);
        }
        rt[k] = rk;
        p0[k] = T(0);
        v0[k] = T(0);
        rr += (double)rk * (double)rk;
    });
    write_partial(P + KM_PC(0), bx, rr, sh);
    write_partial(P + KM_PC(0) + KM_MAX_RED, bx, rr, sh);
    if (bx == 0 && threadIdx.x == 0) {
        st->eps = eps;
        st->max_iter = max_iter;
        st->ref = ref_norm2;
        st->done = 0;
        st->status = KM_MAXITER;
        st->iter = 0;
        st->ring[KM_RING - 1].iter = -1;
    }
}

template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_bicg_init(Flat fl, const T* __restrict__ b, T* __restrict__ x, T* __restrict__ r, T* __restrict__ rt,
            T* __restrict__ p0, T* __restrict__ v0, int from_r, double eps, int max_iter, double ref_norm2,
            KState* st, double* __restrict__ P) {
    bicg_init_body<T>(fl, blockIdx.x, b, x, r, rt, p0, v0, from_r, eps, max_iter, ref_norm2, st, P);
}

// The iteration kernels issue every global load of a tile (the halo inputs of
// the staged vector, the five coefficient planes and r~) into registers
// before they reduce the previous kernel's partials and before the LDS
// barrier: the scalars are only needed for the arithmetic, so the partial
// reduction, the halo loads and the coefficient loads overlap in one memory
// round trip.  CPT = cells per thread of a tile (compile time), staged halo
// positions per thread <= CPT + 2 (tile_ok).
template <typename T, int CPT>
struct TileLoads {
    static constexpr int SPT = CPT + 2;
    int64_t sk[SPT];
    int slds[SPT];
    bool sval[SPT], scen[SPT];
    int64_t ck[CPT];
    int cl[CPT];
    bool cval[CPT];
    T c[5][CPT];

    __device__ __forceinline__ void index(const Tile& tl, const TileOrigin& o) {
        const int W = tl.TJ + 2, H = tl.TI + 2;
#pragma unroll
        for (int q = 0; q < SPT; ++q) {
            const int t = threadIdx.x + q * KM_BLOCK;
            const int li = t / W, lj = t - li * W;
            const int gi = o.i0 - 1 + li, gj = o.j0 - 1 + lj;
            slds[q] = t < W * H ? t : -1;
            sval[q] = t < W * H && gi >= 0 && gi < tl.NX && gj >= 0 && gj < tl.NY;
            scen[q] = li >= 1 && li <= tl.TI && lj >= 1 && lj <= tl.TJ;
            sk[q] = sval[q] ? (int64_t)gi * tl.NY + gj : 0;
        }
#pragma unroll
        for (int q = 0; q < CPT; ++q) {
            const int t = threadIdx.x + q * KM_BLOCK;
            const int li = t / tl.TJ, lj = t - li * tl.TJ;
            const int gi = o.i0 + li, gj = o.j0 + lj;
            cval[q] = t < tl.TI * tl.TJ && gi < tl.NX && gj < tl.NY;
            cl[q] = (li + 1) * W + (lj + 1);
            ck[q] = cval[q] ? (int64_t)gi * tl.NY + gj : 0;
        }
    }
    __device__ __forceinline__ void load_coef(const T* __restrict__ coef, int64_t N) {
#pragma unroll
        for (int q = 0; q < CPT; ++q)
#pragma unroll
            for (int m = 0; m < 5; ++m) c[m][q] = cval[q] ? coef[m * N + ck[q]] : T(0);
    }
    // y = c0 x + c1 x_W + c2 x_E + c3 x_N + c4 x_S (bicgsolver.py:68-69 order)
    __device__ __forceinline__ T apply(const T* lds, int W, int q) const {
        const T* P = lds + cl[q];
        return c[0][q] * P[0] + c[1][q] * P[-W] + c[2][q] * P[W] + c[3][q] * P[1] + c[4][q] * P[-1];
    }
};

// L: the workgroup's logical index within its system (0 .. tl.G-1); the
// workgroup with L == 0 records the iteration's scalars.  Gr: R's workgroups.
// tot: R's totals summed once (k_bicg_total at P + KM_PT), or null: every
// workgroup sums R's partials itself (the same bits either way).
template <typename T, int CPT>
__device__ __forceinline__ void bicg_p_body(const Tile& tl, int L, int Gr, int phase, const T* __restrict__ coef,
                                            const T* __restrict__ r, const T* __restrict__ rt, T* pbuf, T* vbuf,
                                            KState* st, double* __restrict__ P, const double* tot = nullptr) {
    T* lds = reinterpret_cast<T*>(km_smem);
    __shared__ double sh[4];
    const int64_t N = (int64_t)tl.NX * tl.NY;
    const int par = phase & 1;
    const T* __restrict__ p_old = pbuf + par * N;
    const T* __restrict__ v_old = vbuf + par * N;
    T* __restrict__ p_new = pbuf + (1 - par) * N;
    T* __restrict__ v_new = vbuf + (1 - par) * N;
    const int W = tl.TJ + 2;
    const bool lead0 = L == 0 && threadIdx.x == 0;
    const int t0 = L * tl.tpb, t1 = t0 + tl.tpb < tl.T ? t0 + tl.tpb : tl.T;
    const KScal sp = st->ring[(phase + KM_RING - 1) % KM_RING];
    const int i = sp.iter + 1;
    T bt = T(0), om = T(0);
    double rv = 0.0;
    for (int tile = t0; tile < t1; ++tile) {
        TileLoads<T, CPT> ld;
        ld.index(tl, tile_origin(tl, tile));
        T ra[CPT + 2], pa[CPT + 2], va[CPT + 2], rtv[CPT];
#pragma unroll
        for (int q = 0; q < CPT + 2; ++q) {
            const bool v = ld.sval[q];
            ra[q] = v ? r[ld.sk[q]] : T(0);
            pa[q] = v && i > 0 ? p_old[ld.sk[q]] : T(0);
            va[q] = v && i > 0 ? v_old[ld.sk[q]] : T(0);
        }
        ld.load_coef(coef, N);
#pragma unroll
        for (int q = 0; q < CPT; ++q) rtv[q] = ld.cval[q] ? rt[ld.ck[q]] : T(0);
        if (tile == t0) {
            double pr[2];
            if (tot) {
                pr[0] = tot[3];
                pr[1] = tot[4];
            } else {
                sum_partials<2>(P + KM_PC(0), Gr, pr, sh);
            }
            const double rTr = pr[0], rho = pr[1];   // <r_i, r_i>, <r_i, r~> (bicgsolver.py:95)
            const double r0 = i == 0 ? (st->ref > 0.0 ? st->ref : rTr) : st->r0;
            int status = -1;
            if (!finite(rTr)) status = KM_NONFINITE;
            else if (i == 0 && rTr == 0.0) status = KM_OK;
            else if (i > 0 && sqrt(rTr / r0) < st->eps) status = KM_OK;   // bicgsolver.py:130
            else if (i > 0 && sp.omega == 0.0) status = KM_BREAKDOWN;     // karman_taichi.py:451
            else if (i >= st->max_iter) status = KM_MAXITER;
            else if (!finite(rho)) status = KM_NONFINITE;
            else if (rho == 0.0) status = KM_BREAKDOWN;                   // bicgsolver.py:96-98
            if (lead0 && i == 0) st->r0 = r0;
            if (status >= 0) {
                if (lead0) finish(st, status, i, rTr);
                return;
            }
            const double beta = i == 0 ? 0.0 : (rho / sp.rho) * (sp.alpha / sp.omega);   // :102
            if (lead0) {
                st->ring[phase].rho = rho;
                st->ring[phase].beta = beta;
                st->ring[phase].iter = i;
                st->rTr = rTr;
                st->iter = i;
            }
            bt = (T)beta;
            om = i == 0 ? T(0) : (T)sp.omega;
        }
        // update_p (bicgsolver.py:43-46; p = r at i = 0) on the tile and its halo
#pragma unroll
        for (int q = 0; q < CPT + 2; ++q) {
            const bool v = ld.sval[q];
            const T pn = v ? ra[q] + bt * (pa[q] - om * va[q]) : T(0);
            if (ld.slds[q] >= 0) lds[ld.slds[q]] = pn;
            if (v && ld.scen[q]) p_new[ld.sk[q]] = pn;
        }
        __syncthreads();
        // compute_Ap (bicgsolver.py:72-76) and reduce(r~, v) (:107)
#pragma unroll
        for (int q = 0; q < CPT; ++q)
            if (ld.cval[q]) {
                const T y = ld.apply(lds, W, q);
                v_new[ld.ck[q]] = y;
                rv += (double)rtv[q] * (double)y;
            }
        __syncthreads();
    }
    write_partial(P + KM_PA, L, rv, sh);
}

template <typename T, int CPT>
__device__ __forceinline__ void bicg_s_body(const Tile& tl, int L, int phase, const T* __restrict__ coef,
                                            const T* __restrict__ r, const T* vbuf, T* __restrict__ s,
                                            T* __restrict__ t, KState* st, double* __restrict__ P,
                                            const double* tot = nullptr) {
    T* lds = reinterpret_cast<T*>(km_smem);
    __shared__ double sh[8];
    const int64_t N = (int64_t)tl.NX * tl.NY;
    const T* __restrict__ v = vbuf + (1 - (phase & 1)) * N;
    const int W = tl.TJ + 2;
    const bool lead0 = L == 0 && threadIdx.x == 0;
    const int t0 = L * tl.tpb, t1 = t0 + tl.tpb < tl.T ? t0 + tl.tpb : tl.T;
    T al = T(0);
    double ts = 0.0, tt = 0.0;
    for (int tile = t0; tile < t1; ++tile) {
        TileLoads<T, CPT> ld;
        ld.index(tl, tile_origin(tl, tile));
        T ra[CPT + 2], va[CPT + 2];
#pragma unroll
        for (int q = 0; q < CPT + 2; ++q) {
            ra[q] = ld.sval[q] ? r[ld.sk[q]] : T(0);
            va[q] = ld.sval[q] ? v[ld.sk[q]] : T(0);
        }
        ld.load_coef(coef, N);
        if (tile == t0) {
            double pa[1];
            if (tot) pa[0] = tot[0];
            else sum_partials<1>(P + KM_PA, tl.G, pa, sh);
            const double rv = pa[0];
            const double rho = st->ring[phase].rho;
            int status = -1;
            if (!finite(rv)) status = KM_NONFINITE;
            else if (rv == 0.0) status = KM_BREAKDOWN;
            if (status >= 0) {
                if (lead0) finish(st, status, st->ring[phase].iter, st->rTr);
                return;
            }
            const double alpha = rho / rv;   // bicgsolver.py:108
            if (lead0) st->ring[phase].alpha = alpha;
            al = (T)alpha;
        }
        // update_s (bicgsolver.py:60-63) on the tile and its halo
#pragma unroll
        for (int q = 0; q < CPT + 2; ++q) {
            const T sn = ld.sval[q] ? ra[q] - al * va[q] : T(0);
            if (ld.slds[q] >= 0) lds[ld.slds[q]] = sn;
            if (ld.sval[q] && ld.scen[q]) s[ld.sk[q]] = sn;
        }
        __syncthreads();
        // compute_As (:65-69); reduce(t, s), reduce(t, t) (:116-117)
#pragma unroll
        for (int q = 0; q < CPT; ++q)
            if (ld.cval[q]) {
                const T y = ld.apply(lds, W, q);
                const T sk = lds[ld.cl[q]];
                t[ld.ck[q]] = y;
                ts += (double)y * (double)sk;
                tt += (double)y * (double)y;
            }
        __syncthreads();
    }
    write_partials2(P + KM_PB, L, ts, tt, sh);
}

// R: bx = the workgroup's index in the flat geometry, Gs = S's workgroups.
// Each thread keeps kRPer cells' loads in flight.
constexpr int kRPer = 4;
template <typename T>
__device__ __forceinline__ void bicg_r_body(const Flat& fl, int bx, int Gs, int phase, const T* __restrict__ rt,
                                            const T* __restrict__ s, const T* __restrict__ t, const T* pbuf,
                                            T* __restrict__ x, T* __restrict__ r, KState* st,
                                            double* __restrict__ P, const double* tot = nullptr) {
    __shared__ double sh[8];
    const T* __restrict__ p = pbuf + (1 - (phase & 1)) * fl.n;   // p_i, written by P(i)
    const int64_t k0 = (int64_t)bx * fl.chunk;
    const int64_t k1 = k0 + fl.chunk < fl.n ? k0 + fl.chunk : fl.n;
    // the first loads go out before the partial sums are reduced
    T sv[kRPer], tv[kRPer], pv[kRPer], xv[kRPer], rtv[kRPer];
    int64_t kk[kRPer];
    auto load = [&](int64_t base) {
#pragma unroll
        for (int u = 0; u < kRPer; ++u) {
            kk[u] = base + threadIdx.x + (int64_t)u * KM_BLOCK;
            const bool ok = kk[u] < k1;
            sv[u] = ok ? s[kk[u]] : T(0);
            tv[u] = ok ? t[kk[u]] : T(0);
            pv[u] = ok ? p[kk[u]] : T(0);
            xv[u] = ok ? x[kk[u]] : T(0);
            rtv[u] = ok ? rt[kk[u]] : T(0);
        }
    };
    load(k0);
    double pb[2];
    if (tot) {
        pb[0] = tot[1];
        pb[1] = tot[2];
    } else {
        sum_partials<2>(P + KM_PB, Gs, pb, sh);
    }
    const double omega = pb[1] != 0.0 ? pb[0] / pb[1] : 0.0;   // bicgsolver.py:118
    const double alpha = st->ring[phase].alpha;
    if (bx == 0 && threadIdx.x == 0) st->ring[phase].omega = omega;
    const T al = (T)alpha, om = (T)omega;
    double rr = 0.0, rrt = 0.0;
    for (int64_t base = k0; base < k1; base += (int64_t)kRPer * KM_BLOCK) {
        if (base != k0) load(base);
#pragma unroll
        for (int u = 0; u < kRPer; ++u) {
            if (kk[u] >= k1) continue;
            x[kk[u]] = xv[u] + (al * pv[u] + om * sv[u]);   // update_x (bicgsolver.py:78-81)
            const T rk = sv[u] - om * tv[u];                 // update_r (:83-86)
            r[kk[u]] = rk;
            rr += (double)rk * (double)rk;
            rrt += (double)rk * (double)rtv[u];
        }
    }
    write_partials2(P + KM_PC(0), bx, rr, rrt, sh);
}

template <typename T, int CPT>
__global__ void __launch_bounds__(KM_BLOCK)
k_bicg_p(Tile tl, int Gr, int phase, const T* __restrict__ coef, const T* __restrict__ r, const T* __restrict__ rt,
         T* pbuf, T* vbuf, KState* st, double* __restrict__ P, const double* tot) {
    if (st->done) return;
    bicg_p_body<T, CPT>(tl, xcd_remap(blockIdx.x, tl.G), Gr, phase, coef, r, rt, pbuf, vbuf, st, P, tot);
}

template <typename T, int CPT>
__global__ void __launch_bounds__(KM_BLOCK)
k_bicg_s(Tile tl, int phase, const T* __restrict__ coef, const T* __restrict__ r, const T* vbuf, T* __restrict__ s,
         T* __restrict__ t, KState* st, double* __restrict__ P, const double* tot) {
    if (st->done) return;
    bicg_s_body<T, CPT>(tl, xcd_remap(blockIdx.x, tl.G), phase, coef, r, vbuf, s, t, st, P, tot);
}

template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_bicg_r(Flat fl, int Gs, int phase, const T* __restrict__ rt, const T* __restrict__ s, const T* __restrict__ t,
         const T* pbuf, T* __restrict__ x, T* __restrict__ r, KState* st, double* __restrict__ P, const double* tot) {
    if (st->done) return;
    bicg_r_body<T>(fl, blockIdx.x, Gs, phase, rt, s, t, pbuf, x, r, st, P, tot);
}

// One launch's NV partial planes (G workgroups each, plane stride KM_MAX_RED)
// summed once, in sum_partials' order, for the next launch's workgroups
// (which would each sum them otherwise).  A finished solve skips it.
template <int NV>
__global__ void __launch_bounds__(KM_BLOCK)
k_bicg_total(const double* __restrict__ P, int G, double* __restrict__ out, const KState* __restrict__ st) {
    if (st->done) return;
    __shared__ double sh[4];
    double t[NV];
    sum_partials<NV>(P, G, t, sh);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < NV; ++q) out[q] = t[q];
}

// ------------------------------------------------------------ batched systems
// Workgroup bx of a batched launch runs system find_case(starts, n, bx) with
// local index bx - starts[c]; each system keeps its own KState, ring and
// partial region, so the arithmetic is that of the single-system kernels.

__global__ void __launch_bounds__(KM_BLOCK)
kb_bicg_init(const BSys* __restrict__ sys, const int* __restrict__ starts, int n, double eps, int max_iter,
             KState* agg) {
    const int c = find_case(starts, n, blockIdx.x);
    const BSys& S = sys[c];
    const int64_t N = S.fl.n;
    double* ws = S.ws;
    bicg_init_body<double>(S.fl, blockIdx.x - starts[c], S.b, S.x, ws, ws + N, ws + 2 * N, ws + 4 * N, 0, eps,
                           max_iter, 0.0, S.st, S.P);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        agg->done = 0;
        agg->status = KM_MAXITER;
        agg->iter = 0;
        agg->max_iter = max_iter;
    }
}

// tstarts: prefix sums of the systems' tile workgroups; fstarts: of their flat ones
template <int CPT>
__global__ void __launch_bounds__(KM_BLOCK)
kb_bicg_p(const BSys* __restrict__ sys, const int* __restrict__ starts, int n, int phase) {
    const int c = find_case(starts, n, blockIdx.x);
    const BSys& S = sys[c];
    if (S.st->done) return;
    const Tile tl = S.tl;
    const int64_t N = (int64_t)tl.NX * tl.NY;
    double* ws = S.ws;
    // the XCD-aware tile order of the single-system kernel, within the system
    bicg_p_body<double, CPT>(tl, xcd_remap(blockIdx.x - starts[c], tl.G), S.fl.G, phase, S.coef, ws, ws + N,
                             ws + 2 * N, ws + 4 * N, S.st, S.P);
}

template <int CPT>
__global__ void __launch_bounds__(KM_BLOCK)
kb_bicg_s(const BSys* __restrict__ sys, const int* __restrict__ starts, int n, int phase) {
    const int c = find_case(starts, n, blockIdx.x);
    const BSys& S = sys[c];
    if (S.st->done) return;
    const Tile tl = S.tl;
    const int64_t N = (int64_t)tl.NX * tl.NY;
    double* ws = S.ws;
    bicg_s_body<double, CPT>(tl, xcd_remap(blockIdx.x - starts[c], tl.G), phase, S.coef, ws, ws + 4 * N, ws + 6 * N,
                             ws + 7 * N, S.st, S.P);
}

__global__ void __launch_bounds__(KM_BLOCK)
kb_bicg_r(const BSys* __restrict__ sys, const int* __restrict__ fstarts, int n, int phase) {
    const int c = find_case(fstarts, n, blockIdx.x);
    const BSys& S = sys[c];
    if (S.st->done) return;
    const int64_t N = S.fl.n;
    double* ws = S.ws;
    bicg_r_body<double>(S.fl, blockIdx.x - fstarts[c], S.tl.G, phase, ws + N, ws + 6 * N, ws + 7 * N, ws + 2 * N,
                        S.x, ws, S.st, S.P);
}

// agg->done = every system done (one workgroup)
__global__ void __launch_bounds__(KM_BLOCK) kb_all_done(const BSys* __restrict__ sys, int n, KState* agg) {
    __shared__ int open;
    if (threadIdx.x == 0) open = 0;
    __syncthreads();
    for (int c = threadIdx.x; c < n; c += KM_BLOCK)
        if (!sys[c].st->done) open = 1;
    __syncthreads();
    if (threadIdx.x == 0) agg->done = open ? 0 : 1;
}

// Cells per thread of a tile, rounded up to an instantiated value.
inline int tile_cpt(const Tile& tl) {
    const int c = (tl.TI * tl.TJ + KM_BLOCK - 1) / KM_BLOCK;
    return c <= 1 ? 1 : (c <= 2 ? 2 : 4);
}

// Launch helpers: pick the CPT instantiation; with events, the launch is
// bracketed by the dispatch's own start/stop timestamps (hipExtLaunchKernel),
// which add no marker packets to the stream.
#define KM_LAUNCH(kern, g, l, s, e0, e1, ...)                                                        \
    do {                                                                                             \
        if (e0) hipExtLaunchKernelGGL(kern, dim3(g), dim3(KM_BLOCK), l, s, e0, e1, 0, __VA_ARGS__);  \
        else hipLaunchKernelGGL(kern, dim3(g), dim3(KM_BLOCK), l, s, __VA_ARGS__);                    \
    } while (0)

// Partial sets of the plain iteration summed once by k_bicg_total (P's for
// S, S's for R, R's and the init kernel's for P) instead of by every consumer
// workgroup.  mode: km_set_plain_totals (-1 by size, 0 never, 1 all three).
// By size only R's for P, where P's workgroups x 2 R planes exceed
// kPlainTotalWork: at 4096 x 820 (1645 P workgroups, 2 x 2048 R partials
// each) P takes 56.6 instead of 62.3 us and the extra launch ~2.8 us; S and R
// each gain ~1.3 us, less than a launch.
constexpr int64_t kPlainTotalWork = 4000000;
constexpr int PT_P = 1, PT_S = 2, PT_R = 4;

int plain_fin(int mode, const Tile& tl, const Flat& fl) {
    if (mode == 0) return 0;
    if (mode > 0) return PT_P | PT_S | PT_R;
    return (int64_t)tl.G * 2 * fl.G > kPlainTotalWork ? PT_R : 0;
}

// R's (or the init kernel's) partials totalled for the next P
void launch_total_r(const Flat& fl, int fin, const KState* st, double* P, hipStream_t s) {
    if (fin & PT_R)
        hipLaunchKernelGGL(k_bicg_total<2>, dim3(1), dim3(KM_BLOCK), 0, s, (const double*)(P + KM_PC(0)), fl.G,
                           P + KM_PT + 3, st);
}

// One BiCGSTAB iteration (P, S, R, and the totals fin asks for).  `timed` (a
// KM_KERNEL_BICG_* id, or -1) selects the launch bracketed by e0 / e1.
template <typename T>
void launch_bicg_iter(const Tile& tl, const Flat& fl, int phase, const T* coef, T* r, const T* rt, T* sv, T* t, T* x,
                      T* pbuf, T* vbuf, KState* st, double* P, size_t lds, hipStream_t s, int fin, int timed = -1,
                      hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    const double* tp = fin & PT_R ? P + KM_PT : nullptr;
    const double* ts = fin & PT_P ? P + KM_PT : nullptr;
    const double* tr = fin & PT_S ? P + KM_PT : nullptr;
    hipEvent_t a0 = timed == KM_KERNEL_BICG_A ? e0 : nullptr, a1 = timed == KM_KERNEL_BICG_A ? e1 : nullptr;
    switch (tile_cpt(tl)) {
        case 1: KM_LAUNCH((k_bicg_p<T, 1>), tl.G, lds, s, a0, a1, tl, fl.G, phase, coef, r, rt, pbuf, vbuf, st, P, tp); break;
        case 2: KM_LAUNCH((k_bicg_p<T, 2>), tl.G, lds, s, a0, a1, tl, fl.G, phase, coef, r, rt, pbuf, vbuf, st, P, tp); break;
        default: KM_LAUNCH((k_bicg_p<T, 4>), tl.G, lds, s, a0, a1, tl, fl.G, phase, coef, r, rt, pbuf, vbuf, st, P, tp);
    }
    if (fin & PT_P)
        hipLaunchKernelGGL(k_bicg_total<1>, dim3(1), dim3(KM_BLOCK), 0, s, (const double*)(P + KM_PA), tl.G,
                           P + KM_PT, (const KState*)st);
    a0 = timed == KM_KERNEL_BICG_B ? e0 : nullptr;
    a1 = timed == KM_KERNEL_BICG_B ? e1 : nullptr;
    switch (tile_cpt(tl)) {
        case 1: KM_LAUNCH((k_bicg_s<T, 1>), tl.G, lds, s, a0, a1, tl, phase, coef, r, vbuf, sv, t, st, P, ts); break;
        case 2: KM_LAUNCH((k_bicg_s<T, 2>), tl.G, lds, s, a0, a1, tl, phase, coef, r, vbuf, sv, t, st, P, ts); break;
        default: KM_LAUNCH((k_bicg_s<T, 4>), tl.G, lds, s, a0, a1, tl, phase, coef, r, vbuf, sv, t, st, P, ts);
    }
    if (fin & PT_S)
        hipLaunchKernelGGL(k_bicg_total<2>, dim3(1), dim3(KM_BLOCK), 0, s, (const double*)(P + KM_PB), tl.G,
                           P + KM_PT + 1, (const KState*)st);
    a0 = timed == KM_KERNEL_BICG_R ? e0 : nullptr;
    a1 = timed == KM_KERNEL_BICG_R ? e1 : nullptr;
    KM_LAUNCH((k_bicg_r<T>), fl.G, 0, s, a0, a1, fl, tl.G, phase, rt, (const T*)sv, (const T*)t, (const T*)pbuf, x, r,
              st, P, tr);
    launch_total_r(fl, fin, st, P, s);
}

// A single kernel of an iteration (back-to-back timing).
template <typename T>
void launch_bicg_one(int which, const Tile& tl, const Flat& fl, int phase, const T* coef, T* r, const T* rt, T* sv,
                     T* t, T* x, T* pbuf, T* vbuf, KState* st, double* P, size_t lds, hipStream_t s, int fin) {
    const double* tp = fin & PT_R ? P + KM_PT : nullptr;
    const double* ts = fin & PT_P ? P + KM_PT : nullptr;
    const double* tr = fin & PT_S ? P + KM_PT : nullptr;
    if (which == KM_KERNEL_BICG_A) {
        switch (tile_cpt(tl)) {
            case 1: hipLaunchKernelGGL((k_bicg_p<T, 1>), tl.G, KM_BLOCK, lds, s, tl, fl.G, phase, coef, r, rt, pbuf, vbuf, st, P, tp); break;
            case 2: hipLaunchKernelGGL((k_bicg_p<T, 2>), tl.G, KM_BLOCK, lds, s, tl, fl.G, phase, coef, r, rt, pbuf, vbuf, st, P, tp); break;
            default: hipLaunchKernelGGL((k_bicg_p<T, 4>), tl.G, KM_BLOCK, lds, s, tl, fl.G, phase, coef, r, rt, pbuf, vbuf, st, P, tp);
        }
    } else if (which == KM_KERNEL_BICG_B) {
        switch (tile_cpt(tl)) {
            case 1: hipLaunchKernelGGL((k_bicg_s<T, 1>), tl.G, KM_BLOCK, lds, s, tl, phase, coef, r, vbuf, sv, t, st, P, ts); break;
            case 2: hipLaunchKernelGGL((k_bicg_s<T, 2>), tl.G, KM_BLOCK, lds, s, tl, phase, coef, r, vbuf, sv, t, st, P, ts); break;
            default: hipLaunchKernelGGL((k_bicg_s<T, 4>), tl.G, KM_BLOCK, lds, s, tl, phase, coef, r, vbuf, sv, t, st, P, ts);
        }
    } else {
        hipLaunchKernelGGL((k_bicg_r<T>), fl.G, KM_BLOCK, 0, s, fl, tl.G, phase, rt, (const T*)sv, (const T*)t,
                           (const T*)pbuf, x, r, st, P, tr);
    }
}

// ------------------------------------------------------------ CG (cgsolver.py:83-114)

template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_cg_init(Flat fl, const T* __restrict__ b, T* __restrict__ x, T* __restrict__ r, T* __restrict__ p0, double eps,
          int max_iter, KState* st, double* __restrict__ P) {
    __shared__ double sh[4];
    double acc = 0.0;
    flat_sweep(fl, [&](int64_t k) {
        const T rk = b[k];
        r[k] = rk;
        x[k] = T(0);
        p0[k] = T(0);
        acc += (double)rk * (double)rk;
    });
    write_partial(P + KM_PC(1), blockIdx.x, acc, sh);
    if (lead()) {
        st->eps = eps;
        st->max_iter = max_iter;
        st->ref = 0.0;
        st->done = 0;
        st->status = KM_MAXITER;
        st->iter = 0;
        st->ring[KM_RING - 1].iter = 0;
    }
}

template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_cg_a(Tile tl, Flat fc, int phase, const T* __restrict__ coef, const T* __restrict__ r, T* pbuf, T* __restrict__ Ap,
       KState* st, double* __restrict__ P) {
    if (st->done) return;
    T* lds = reinterpret_cast<T*>(km_smem);
    __shared__ double sh[4];
    double nw[1];
    sum_partials<1>(P + KM_PC((phase + 1) & 1), fc.G, nw, sh);
    const double rTr = nw[0];
    const KScal sp = st->ring[(phase + KM_RING - 1) % KM_RING];
    const int i = sp.iter;
    int status = -1;
    const double r0 = i == 0 ? rTr : st->r0;
    if (!finite(rTr)) status = KM_NONFINITE;
    else if (i == 0 && rTr == 0.0) status = KM_OK;
    else if (i > 0 && sqrt(rTr) < sqrt(r0) * st->eps) status = KM_OK;   // cgsolver.py:240
    else if (i >= st->max_iter) status = KM_MAXITER;
    if (status >= 0) {
        if (lead()) {
            if (i == 0) st->r0 = r0;
            finish(st, status, i, rTr);
        }
        return;
    }
    const double beta = i == 0 ? 0.0 : rTr / sp.rho;   // cgsolver.py:245
    if (lead()) {
        if (i == 0) st->r0 = r0;
        st->ring[phase].rho = rTr;   // old_rTr of this iteration
        st->ring[phase].beta = beta;
        st->rTr = rTr;
        st->iter = i;
    }
    const int64_t N = (int64_t)tl.NX * tl.NY;
    const int par = phase & 1;
    const T* __restrict__ p_old = pbuf + par * N;
    T* __restrict__ p_new = pbuf + (1 - par) * N;
    const T bt = (T)beta;
    double acc = 0.0;
    // update_p (cgsolver.py:216-219), compute_Ap (:194-198), reduce(p, Ap) (:233)
    tile_sweep<T>(tl, lds, [&](int64_t k) { return r[k] + bt * p_old[k]; }, p_new, coef, [&](int64_t k, T y, T pk) {
        Ap[k] = y;
        acc += (double)pk * (double)y;
    });
    write_partial(P + KM_PA, xcd_remap(blockIdx.x, tl.G), acc, sh);
}

template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_cg_b(Flat fl, int Ga, int phase, T* __restrict__ x, T* __restrict__ r, const T* pbuf, const T* __restrict__ Ap,
       KState* st, double* __restrict__ P) {
    if (st->done) return;
    __shared__ double sh[4];
    double pap[1];
    sum_partials<1>(P + KM_PA, Ga, pap, sh);
    const int i = st->ring[(phase + KM_RING - 1) % KM_RING].iter;
    if (!finite(pap[0]) || pap[0] == 0.0) {
        if (lead()) finish(st, pap[0] == 0.0 ? KM_BREAKDOWN : KM_NONFINITE, i, st->rTr);
        return;
    }
    const double alpha = st->ring[phase].rho / pap[0];   // cgsolver.py:234
    if (lead()) {
        st->ring[phase].alpha = alpha;
        st->ring[phase].iter = i + 1;
    }
    const T* __restrict__ p = pbuf + (1 - (phase & 1)) * fl.n;
    const T al = (T)alpha;
    double acc = 0.0;
    flat_sweep(fl, [&](int64_t k) {
        x[k] = x[k] + al * p[k];          // cgsolver.py:206-209
        const T rk = r[k] - al * Ap[k];   // cgsolver.py:211-214
        r[k] = rk;
        acc += (double)rk * (double)rk;
    });
    write_partial(P + KM_PC(phase & 1), blockIdx.x, acc, sh);
}

// ------------------------------------------------------------ host helpers

int grid_stride_blocks(int64_t n) {
    int64_t g = (n + KM_BLOCK - 1) / KM_BLOCK;
    if (g > 2048) g = 2048;
    if (g < 1) g = 1;
    return (int)g;
}

template <typename T>
size_t tile_lds(const Tile& tl) { return (size_t)(tl.TI + 2) * (tl.TJ + 2) * sizeof(T); }

int check_system(km_ctx* ctx, int NX, int NY) {
    KM_REQUIRE(ctx, NX >= 1 && NY >= 1, "system shape %d x %d is invalid", NX, NY);
    KM_REQUIRE(ctx, (int64_t)NX * NY < ((int64_t)1 << 31), "system too large");
    return KM_OK;
}

int get_tile(km_ctx* ctx, int NX, int NY, int prec, Tile* out) {
    int TI = 0, TJ = 0;
    if (int rc = km_choose_tile(ctx, NX, NY, prec, &TI, &TJ)) return rc;
    KM_REQUIRE(ctx, tile_ok(TI, TJ), "tile %dx%d is not valid for the iteration kernels", TI, TJ);
    *out = make_tile(NX, NY, TI, TJ);
    return KM_OK;
}

const int kChunk = 8;   // iterations per captured chunk (a multiple of KM_RING)

// Plain BiCGSTAB on typed buffers.  ws holds 8n elements of T.
template <typename T>
int bicgstab_typed(km_ctx* ctx, const Tile& tl, const T* coef, const T* b, T* x, T* ws, bool from_r, double eps,
                   int max_iter, bool use_graph, KState* st, KState* out, double ref_norm2 = 0.0,
                   double* P = nullptr) {
    const int64_t n = (int64_t)tl.NX * tl.NY;
    T* r = ws;
    T* rt = ws + n;
    T* pbuf = ws + 2 * n;   // 2n
    T* vbuf = ws + 4 * n;   // 2n
    T* s = ws + 6 * n;
    T* t = ws + 7 * n;
    const Flat fl = make_flat(n);
    if (!P) P = ctx->partials;
    hipLaunchKernelGGL(k_bicg_init<T>, dim3(fl.G), dim3(KM_BLOCK), 0, ctx->stream, fl, b, x, r, rt, pbuf, vbuf,
                       from_r ? 1 : 0, eps, max_iter, ref_norm2, st, P);
    const int fin = plain_fin(ctx->plain_totals, tl, fl);
    launch_total_r(fl, fin, st, P, ctx->stream);
    KM_CHECK_LAUNCH(ctx);
    const size_t lds = tile_lds<T>(tl);
    auto enqueue = [=](hipStream_t s_, int phase) {
        launch_bicg_iter<T>(tl, fl, phase, coef, r, rt, s, t, x, pbuf, vbuf, st, P, lds, s_, fin);
    };
    GraphKey key;
    key.v = {1, sizeof(T), (uintptr_t)tl.NX, (uintptr_t)tl.NY, (uintptr_t)tl.TI, (uintptr_t)tl.TJ,
             (uintptr_t)coef, (uintptr_t)x, (uintptr_t)ws, (uintptr_t)st, (uintptr_t)P, (uintptr_t)fin};
    return run_chunks(ctx, st, key, use_graph, kChunk, max_iter, enqueue, out);
}

template <typename T>
int cg_typed(km_ctx* ctx, const Tile& tl, const T* coef, const T* b, T* x, T* ws, double eps, int max_iter,
             bool use_graph, KState* st, KState* out) {
    const int64_t n = (int64_t)tl.NX * tl.NY;
    T* r = ws;
    T* pbuf = ws + n;   // 2n
    T* Ap = ws + 3 * n;
    const Flat fl = make_flat(n);
    double* P = ctx->partials;
    hipLaunchKernelGGL(k_cg_init<T>, dim3(fl.G), dim3(KM_BLOCK), 0, ctx->stream, fl, b, x, r, pbuf, eps, max_iter,
                       st, P);
    KM_CHECK_LAUNCH(ctx);
    const size_t lds = tile_lds<T>(tl);
    auto enqueue = [=](hipStream_t s_, int phase) {
        hipLaunchKernelGGL(k_cg_a<T>, dim3(tl.G), dim3(KM_BLOCK), lds, s_, tl, fl, phase, coef, r, pbuf, Ap, st, P);
        hipLaunchKernelGGL(k_cg_b<T>, dim3(fl.G), dim3(KM_BLOCK), 0, s_, fl, tl.G, phase, x, r, pbuf, Ap, st, P);
    };
    GraphKey key;
    key.v = {2, sizeof(T), (uintptr_t)tl.NX, (uintptr_t)tl.NY, (uintptr_t)tl.TI, (uintptr_t)tl.TJ,
             (uintptr_t)coef, (uintptr_t)x, (uintptr_t)ws, (uintptr_t)st};
    return run_chunks(ctx, st, key, use_graph, kChunk, max_iter, enqueue, out);
}

// partials / dres (may be null: the context's workspace) receive the
// per-workgroup partial sums and the reduced result on the device.
int residual_norm2(km_ctx* ctx, const Tile& tl, const double* coef, const double* b, const double* x, double* r,
                   bool sync, double* out, double* partials = nullptr, double* dres = nullptr) {
    double* P = (partials ? partials : ctx->partials) + KM_PX;
    if (!dres) dres = ctx->dres;
    hipLaunchKernelGGL(k_residual, dim3(tl.G), dim3(KM_BLOCK), tile_lds<double>(tl), ctx->stream, tl, coef, b, x, r, P);
    hipLaunchKernelGGL(k_finish<1>, dim3(1), dim3(KM_BLOCK), 0, ctx->stream, (const double*)P, tl.G, dres);
    KM_CHECK_LAUNCH(ctx);
    if (sync) {
        KM_HIP(ctx, hipMemcpyAsync(ctx->h_res, dres, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        KM_HIP_SYNC(ctx, hipStreamSynchronize(ctx->stream));
        *out = ctx->h_res[0];
    }
    return KM_OK;
}

int dot_sync(km_ctx* ctx, int64_t n, int prec, const void* a, const void* b, double* out, double* partials = nullptr,
             double* dres = nullptr) {
    const Flat fl = make_flat(n);
    double* P = (partials ? partials : ctx->partials) + KM_PX;
    if (!dres) dres = ctx->dres;
    if (prec == KM_F64)
        hipLaunchKernelGGL(k_dot<double>, dim3(fl.G), dim3(KM_BLOCK), 0, ctx->stream, fl, (const double*)a,
                           (const double*)b, P);
    else
        hipLaunchKernelGGL(k_dot<float>, dim3(fl.G), dim3(KM_BLOCK), 0, ctx->stream, fl, (const float*)a,
                           (const float*)b, P);
    hipLaunchKernelGGL(k_finish<1>, dim3(1), dim3(KM_BLOCK), 0, ctx->stream, (const double*)P, fl.G, dres + 1);
    KM_CHECK_LAUNCH(ctx);
    KM_HIP(ctx, hipMemcpyAsync(ctx->h_res + 1, dres + 1, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    KM_HIP_SYNC(ctx, hipStreamSynchronize(ctx->stream));
    *out = ctx->h_res[1];
    return KM_OK;
}

template <int CPT>
void batch_iter(const BSys* d_sys, const int* tstarts, const int* fstarts, int n, int64_t tblocks, int64_t fblocks,
                int phase, size_t lds, hipStream_t q) {
    hipLaunchKernelGGL(kb_bicg_p<CPT>, dim3((unsigned)tblocks), dim3(KM_BLOCK), lds, q, d_sys, tstarts, n, phase);
    hipLaunchKernelGGL(kb_bicg_s<CPT>, dim3((unsigned)tblocks), dim3(KM_BLOCK), lds, q, d_sys, tstarts, n, phase);
    hipLaunchKernelGGL(kb_bicg_r, dim3((unsigned)fblocks), dim3(KM_BLOCK), 0, q, d_sys, fstarts, n, phase);
}

void fill_info(km_solve_info* info, const KState& s) {
    if (!info) return;
    info->iterations = s.iter;
    info->status = s.status;
    info->rel_resid = s.r0 > 0 ? sqrt(s.rTr / s.r0) : 0.0;
}

// Tile autotuner: on the first solve of an NX x NY system in precision T,
// time a short BiCGSTAB iteration sequence (init, 4 warm-up iterations, 16
// timed iterations, stopping test off) for each valid tile candidate on the
// caller's own system and keep the fastest.  ~15 candidates x 20 iterations:
// a few milliseconds, once per shape.
template <typename T>
int autotune_tiles(km_ctx* ctx, int NX, int NY, int prec, const T* coef, const T* b, T* ws) {
    const int64_t n = (int64_t)NX * NY;
    std::vector<std::pair<int, int>> cand;
    for (int segs = 1; segs <= 8; ++segs) {
        const int tj = (NY + segs - 1) / segs;
        if (tj > 1024 || (segs > 1 && (NY + segs - 2) / (segs - 1) == tj)) continue;
        for (int ti : {1, 2, 3, 4, 6, 8, 12, 16}) {
            if (ti > NX || !tile_ok(ti, tj) || ti * tj < 96) continue;
            cand.push_back({ti, tj});
        }
    }
    if (cand.empty()) return KM_OK;
    T *r = ws, *rt = ws + n, *pbuf = ws + 2 * n, *vbuf = ws + 4 * n, *sv = ws + 6 * n, *t = ws + 7 * n, *x = ws + 8 * n;
    KState* st = &ctx->st[3];
    double* P = ctx->partials;
    const Flat fl = make_flat(n);
    hipEvent_t e0, e1;
    KM_HIP_SYNC(ctx, hipEventCreate(&e0));
    KM_HIP_SYNC(ctx, hipEventCreate(&e1));
    double best = 1e300;
    std::pair<int, int> pick = cand[0];
    for (const auto& c : cand) {
        const Tile tl = make_tile(NX, NY, c.first, c.second);
        const size_t lds = tile_lds<T>(tl);
        const int fin = plain_fin(ctx->plain_totals, tl, fl);
        hipLaunchKernelGGL(k_bicg_init<T>, dim3(fl.G), dim3(KM_BLOCK), 0, ctx->stream, fl, b, x, r, rt, pbuf, vbuf, 0,
                           0.0, 1 << 30, 0.0, st, P);
        launch_total_r(fl, fin, st, P, ctx->stream);
        for (int it = 0; it < 20; ++it) {
            if (it == 4) (void)hipEventRecord(e0, ctx->stream);
            const int ph = it % KM_RING;
            launch_bicg_iter<T>(tl, fl, ph, coef, r, rt, sv, t, x, pbuf, vbuf, st, P, lds, ctx->stream, fin);
        }
        (void)hipEventRecord(e1, ctx->stream);
        KM_HIP_SYNC(ctx, hipEventSynchronize(e1));
        KM_CHECK_LAUNCH(ctx);
        float ms = 0.f;
        KM_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
        KM_HIP_SYNC(ctx, hipMemcpy(&ctx->h_st[3], st, sizeof(KState), hipMemcpyDeviceToHost));
        if (!ctx->h_st[3].done && ms < best) {   // a candidate whose run stopped early is not comparable
            best = ms;
            pick = c;
        }
    }
    { KmHostLock lk_(km_host_mutex()); (void)hipEventDestroy(e0); }
    { KmHostLock lk_(km_host_mutex()); (void)hipEventDestroy(e1); }
    ctx->tiles[std::make_tuple(NX, NY, prec)] = pick;
    return KM_OK;
}

template <typename T>
int ensure_tuned(km_ctx* ctx, int NX, int NY, int prec, const T* coef, const T* b, T* ws) {
    auto key = std::make_tuple(NX, NY, prec);
    if (!ctx->autotune || ctx->tile_override.count(key) || ctx->tuned.count(key)) return KM_OK;
    ctx->tuned.insert(key);
    return autotune_tiles<T>(ctx, NX, NY, prec, coef, b, ws);
}

// Mixed precision: FP32 BiCGSTAB inner solves on a rounded copy of A with
// FP64 scalars, wrapped in FP64 iterative refinement judged by the true
// residual ||b - A x|| / ||b||.
int bicgstab_mixed(km_ctx* ctx, int NX, int NY, const double* coef, const double* b, double* x, double eps,
                   int max_iter, int flags, km_solve_info* info) {
    const int64_t n = (int64_t)NX * NY;
    // FP64: r (n); FP32: coef (5n), rhs (n), d (n), inner workspace (9n)
    const size_t bytes = (size_t)n * sizeof(double) + (size_t)16 * n * sizeof(float) + 1024;
    if (int rc = km_ensure_workspace(ctx, bytes)) return rc;
    double* r = static_cast<double*>(km_ws_vectors(ctx));
    float* c32 = reinterpret_cast<float*>(r + n);
    float* rhs = c32 + 5 * n;
    float* d = rhs + n;
    float* w32 = d + n;
    const bool use_graph = !(flags & KM_FLAG_NO_GRAPH);
    hipLaunchKernelGGL(k_to_float, dim3(grid_stride_blocks(5 * n)), dim3(KM_BLOCK), 0, ctx->stream, 5 * n, coef, c32,
                       1.0);
    hipLaunchKernelGGL(k_to_float, dim3(grid_stride_blocks(n)), dim3(KM_BLOCK), 0, ctx->stream, n, b, rhs, 1.0);
    KM_CHECK_LAUNCH(ctx);
    if (int rc = ensure_tuned<float>(ctx, NX, NY, KM_F32, c32, rhs, w32)) return rc;
    Tile t64, t32;
    if (int rc = get_tile(ctx, NX, NY, KM_F64, &t64)) return rc;
    if (int rc = get_tile(ctx, NX, NY, KM_F32, &t32)) return rc;
    double bb = 0.0;
    if (int rc = dot_sync(ctx, n, KM_F64, b, b, &bb)) return rc;
    if (!(flags & KM_FLAG_X0)) KM_HIP(ctx, hipMemsetAsync(x, 0, n * sizeof(double), ctx->stream));
    double rr = 0.0;
    if (int rc = residual_norm2(ctx, t64, coef, b, x, r, true, &rr)) return rc;
    const double bnorm = sqrt(bb);
    if (info) { info->bnorm = bnorm; info->refinements = 0; info->fallback = 0; }
    if (bb == 0.0) {
        if (info) { info->iterations = 0; info->status = KM_OK; info->rel_resid = 0; info->rel_resid_true = 0; }
        return KM_OK;
    }
    int total = 0, status = KM_MAXITER, sweeps = 0;
    double rel = sqrt(rr) / bnorm;
    const double inner_floor = 1e-6;   // what an FP32 Krylov solve reliably reaches
    while (true) {
        if (!(rel == rel)) { status = KM_NONFINITE; break; }
        if (rel < eps) { status = KM_OK; break; }
        if (total >= max_iter || sweeps >= 64) { status = KM_MAXITER; break; }
        const double rn = sqrt(rr);
        hipLaunchKernelGGL(k_to_float, dim3(grid_stride_blocks(n)), dim3(KM_BLOCK), 0, ctx->stream, n,
                           (const double*)r, rhs, 1.0 / rn);
        KM_CHECK_LAUNCH(ctx);
        double inner_eps = eps / rel;
        if (inner_eps < inner_floor) inner_eps = inner_floor;
        if (inner_eps > 1e-2) inner_eps = 1e-2;
        KState s;
        if (int rc = bicgstab_typed<float>(ctx, t32, c32, rhs, d, w32, false, inner_eps, max_iter - total,
                                           use_graph, &ctx->st[0], &s))
            return rc;
        total += s.iter;
        sweeps++;
        hipLaunchKernelGGL(k_add_scaled, dim3(grid_stride_blocks(n)), dim3(KM_BLOCK), 0, ctx->stream, n, x,
                           (const float*)d, rn);
        KM_CHECK_LAUNCH(ctx);
        const double prev = rel;
        if (int rc = residual_norm2(ctx, t64, coef, b, x, r, true, &rr)) return rc;
        rel = sqrt(rr) / bnorm;
        if (!(rel < 0.5 * prev)) {
            // FP32 recurrences lost accuracy: take the sweep back (or restart
            // from zero if it produced non-finite values)
            if (rel == rel && rel < 1e300) {
                hipLaunchKernelGGL(k_add_scaled, dim3(grid_stride_blocks(n)), dim3(KM_BLOCK), 0, ctx->stream, n, x,
                                   (const float*)d, -rn);
                KM_CHECK_LAUNCH(ctx);
            } else {
                KM_HIP(ctx, hipMemsetAsync(x, 0, n * sizeof(double), ctx->stream));
            }
            status = KM_STALL;
            break;
        }
    }
    if (status == KM_STALL || status == KM_NONFINITE) {
        // The FP32 iteration broke down or stopped reducing the FP64 residual
        // (ill-conditioned, unpreconditioned systems): finish with FP64
        // BiCGSTAB from the last good iterate.
        km_solve_info sub;
        const int rc = km_bicgstab(ctx, NX, NY, coef, b, x, eps, max_iter > total ? max_iter - total : 1,
                                   (flags & ~KM_FLAG_MIXED) | KM_FLAG_X0, &sub);
        if (rc < 0) return rc;
        total += sub.iterations;
        status = sub.status;
        rel = sub.rel_resid_true;
        if (info) info->fallback = 1;
    }
    if (info) {
        info->iterations = total;
        info->status = status;
        info->refinements = sweeps;
        info->rel_resid = rel;
        info->rel_resid_true = rel;
    }
    return KM_OK;
}

}  // namespace

int batch_bicgstab(km_ctx* ctx, int n, const BSys* sys, const BSys* d_sys, const int* tstarts, const int* fstarts,
                   int64_t tblocks, int64_t fblocks, int TI, int TJ, double eps, int max_iter, bool use_graph,
                   KState* agg, KState* states) {
    KM_REQUIRE(ctx, n >= 1 && tblocks > 0 && fblocks > 0, "empty batch");
    KM_REQUIRE(ctx, tile_ok(TI, TJ), "tile %dx%d is not valid for the iteration kernels", TI, TJ);
    KM_REQUIRE(ctx, tblocks < ((int64_t)1 << 31) && fblocks < ((int64_t)1 << 31), "batch too large");
    const size_t lds = (size_t)(TI + 2) * (TJ + 2) * sizeof(double);
    const int cpt = tile_cpt(make_tile(TI, TJ, TI, TJ));
    hipLaunchKernelGGL(kb_bicg_init, dim3((unsigned)fblocks), dim3(KM_BLOCK), 0, ctx->stream, d_sys, fstarts, n, eps,
                       max_iter, agg);
    KM_CHECK_LAUNCH(ctx);
    auto enqueue = [=](hipStream_t q, int phase) {
        if (cpt == 1) batch_iter<1>(d_sys, tstarts, fstarts, n, tblocks, fblocks, phase, lds, q);
        else if (cpt == 2) batch_iter<2>(d_sys, tstarts, fstarts, n, tblocks, fblocks, phase, lds, q);
        else batch_iter<4>(d_sys, tstarts, fstarts, n, tblocks, fblocks, phase, lds, q);
        if (phase == KM_RING - 1) hipLaunchKernelGGL(kb_all_done, dim3(1), dim3(KM_BLOCK), 0, q, d_sys, n, agg);
    };
    GraphKey key;
    key.v = {9, (uintptr_t)n, (uintptr_t)d_sys, (uintptr_t)tstarts, (uintptr_t)tblocks, (uintptr_t)TI, (uintptr_t)TJ,
             (uintptr_t)sys[0].coef, (uintptr_t)agg};
    for (int c = 0; c < n; ++c) key.v.push_back((uintptr_t)sys[c].coef);
    KState fin;
    if (int rc = run_chunks(ctx, agg, key, use_graph, kChunk, max_iter, enqueue, &fin)) return rc;
    KM_HIP(ctx, hipMemcpyAsync(states, sys[0].st, n * sizeof(KState), hipMemcpyDeviceToHost, ctx->stream));
    KM_HIP_SYNC(ctx, hipStreamSynchronize(ctx->stream));
    return KM_OK;
}

int km_set_plain_totals(km_ctx* ctx, int mode) {
    if (!ctx) return KM_EINVAL;
    ctx->plain_totals = mode < 0 ? -1 : (mode > 0 ? 1 : 0);
    return KM_OK;
}

// Tile choice: an explicit override, else the cached choice, else a default
// that splits each row into near-equal segments of at most 256 columns and
// stacks rows until a tile covers about 256 cells.
int km_choose_tile(km_ctx* ctx, int NX, int NY, int prec, int* TI, int* TJ) {
    auto key = std::make_tuple(NX, NY, prec);
    auto ov = ctx->tile_override.find(key);
    if (ov != ctx->tile_override.end()) { *TI = ov->second.first; *TJ = ov->second.second; return KM_OK; }
    auto it = ctx->tiles.find(key);
    if (it != ctx->tiles.end()) { *TI = it->second.first; *TJ = it->second.second; return KM_OK; }
    const int segs = (NY + 255) / 256;
    int tj = (NY + segs - 1) / segs;
    int ti = (256 + tj - 1) / tj;
    if (ti > NX) ti = NX;
    ctx->tiles[key] = {ti, tj};
    *TI = ti;
    *TJ = tj;
    return KM_OK;
}

extern "C" {

int km_stencil_apply(km_ctx* ctx, int NX, int NY, int prec, const void* coef, const void* x, void* y) {
    if (!ctx) return KM_EINVAL;
    if (int rc = check_system(ctx, NX, NY)) return rc;
    KM_REQUIRE(ctx, coef && x && y, "km_stencil_apply: NULL pointer");
    KM_REQUIRE(ctx, prec == KM_F64 || prec == KM_F32, "unknown precision %d", prec);
    Tile tl;
    if (int rc = get_tile(ctx, NX, NY, prec, &tl)) return rc;
    if (prec == KM_F64)
        hipLaunchKernelGGL(k_stencil<double>, dim3(tl.G), dim3(KM_BLOCK), tile_lds<double>(tl), ctx->stream, tl,
                           (const double*)coef, (const double*)x, (double*)y);
    else
        hipLaunchKernelGGL(k_stencil<float>, dim3(tl.G), dim3(KM_BLOCK), tile_lds<float>(tl), ctx->stream, tl,
                           (const float*)coef, (const float*)x, (float*)y);
    KM_CHECK_LAUNCH(ctx);
    return KM_OK;
}

// Average device duration of one launch of a solver kernel, timed inside a
// solve-like sequence: BiCGSTAB iterations on (coef, b) run back to back with
// the stopping test disabled (eps = 0), and every launch of the chosen kernel
// carries start/stop events taken from its own dispatch.  KM_KERNEL_STENCIL
// times plain stencil launches.
int km_time_kernel(km_ctx* ctx, int which, int NX, int NY, const double* coef, const double* b, int reps,
                   double* avg_us) {
    if (!ctx) return KM_EINVAL;
    if (int rc = check_system(ctx, NX, NY)) return rc;
    KM_REQUIRE(ctx, coef && b && avg_us && reps > 0 && reps <= 4096, "km_time_kernel: invalid arguments");
    const bool b2b = (which & KM_KERNEL_BACK_TO_BACK) != 0;
    const bool sym = (which & KM_KERNEL_SYMMETRIC) != 0;
    which &= ~(KM_KERNEL_BACK_TO_BACK | KM_KERNEL_SYMMETRIC);
    if (which >= KM_KERNEL_TL_PV && which <= KM_KERNEL_TL_CT)
        return kmcoarse::time_kernel(ctx, which, NX, NY, coef, b, reps, b2b, avg_us, sym);
    KM_REQUIRE(ctx, !sym, "KM_KERNEL_SYMMETRIC applies to the two-level kernels");
    KM_REQUIRE(ctx, which == KM_KERNEL_BICG_A || which == KM_KERNEL_BICG_B || which == KM_KERNEL_BICG_R ||
                        which == KM_KERNEL_STENCIL,
               "unknown kernel id %d", which);
    const int64_t n = (int64_t)NX * NY;
    if (int rc = km_ensure_workspace(ctx, (size_t)9 * n * sizeof(double))) return rc;
    double* ws = static_cast<double*>(km_ws_vectors(ctx));
    if (int rc = ensure_tuned<double>(ctx, NX, NY, KM_F64, coef, b, ws)) return rc;
    Tile tl;
    if (int rc = get_tile(ctx, NX, NY, KM_F64, &tl)) return rc;
    double *r = ws, *rt = ws + n, *pbuf = ws + 2 * n, *vbuf = ws + 4 * n, *sv = ws + 6 * n, *t = ws + 7 * n;
    double* x = ws + 8 * n;
    KState* st = &ctx->st[2];
    const Flat fl = make_flat(n);
    double* P = ctx->partials;
    const size_t lds = tile_lds<double>(tl);
    std::vector<hipEvent_t> ev(2 * (size_t)reps, nullptr);
    for (auto& e : ev) KM_HIP_SYNC(ctx, hipEventCreate(&e));
    auto cleanup = [&]() {
        for (auto& e : ev)
            if (e) { KmHostLock lk_(km_host_mutex()); (void)hipEventDestroy(e); }
    };
    const int fin = plain_fin(ctx->plain_totals, tl, fl);
    hipLaunchKernelGGL(k_bicg_init<double>, dim3(fl.G), dim3(KM_BLOCK), 0, ctx->stream, fl, b, x, r, rt, pbuf, vbuf,
                       0, 0.0, 1 << 30, 0.0, st, P);
    launch_total_r(fl, fin, st, P, ctx->stream);
    const int warm = 8;
    if (b2b) {
        // reach a mid-solve state, then time `reps` identical launches of the
        // kernel between two events: P of the next iteration, or S / R of
        // the last one (inputs unchanged from launch to launch)
        for (int it = 0; it < warm; ++it)
            launch_bicg_iter<double>(tl, fl, it % KM_RING, coef, r, rt, sv, t, x, pbuf, vbuf, st, P, lds, ctx->stream,
                                     fin);
        KM_HIP(ctx, hipEventRecord(ev[0], ctx->stream));
        for (int k = 0; k < reps; ++k) {
            if (which == KM_KERNEL_STENCIL)
                hipLaunchKernelGGL(k_stencil<double>, dim3(tl.G), dim3(KM_BLOCK), lds, ctx->stream, tl, coef,
                                   (const double*)r, t);
            else
                launch_bicg_one<double>(which, tl, fl, which == KM_KERNEL_BICG_A ? warm % KM_RING : (warm - 1) % KM_RING,
                                        coef, r, rt, sv, t, x, pbuf, vbuf, st, P, lds, ctx->stream, fin);
        }
        KM_HIP(ctx, hipEventRecord(ev[1], ctx->stream));
    }
    for (int it = 0; it < (b2b ? 0 : warm + reps); ++it) {
        const int ph = it % KM_RING;
        const int k = it - warm;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (k >= 0) { e0 = ev[2 * k]; e1 = ev[2 * k + 1]; }
        if (which == KM_KERNEL_STENCIL) {
            KM_LAUNCH((k_stencil<double>), tl.G, lds, ctx->stream, e0, e1, tl, coef, (const double*)r, t);
            continue;
        }
        launch_bicg_iter<double>(tl, fl, ph, coef, r, rt, sv, t, x, pbuf, vbuf, st, P, lds, ctx->stream, fin, which, e0,
                                 e1);
    }
    hipError_t le = hipGetLastError();
    hipError_t se;
    {
        KmHostLock lk(km_host_mutex());
        se = hipStreamSynchronize(ctx->stream);
    }
    if (le != hipSuccess || se != hipSuccess) {
        cleanup();
        return km_fail(ctx, KM_EHIP, "km_time_kernel: %s / %s", hipGetErrorString(le), hipGetErrorString(se));
    }
    double total_ms = 0.0;
    for (int k = 0; k < (b2b ? 1 : reps); ++k) {
        float ms = 0.f;
        hipError_t e = hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]);
        if (e != hipSuccess) {
            cleanup();
            return km_fail(ctx, KM_EHIP, "hipEventElapsedTime: %s", hipGetErrorString(e));
        }
        total_ms += ms;
    }
    cleanup();
    KM_HIP_SYNC(ctx, hipMemcpy(&ctx->h_st[3], st, sizeof(KState), hipMemcpyDeviceToHost));
    KM_REQUIRE(ctx, !ctx->h_st[3].done, "km_time_kernel: the iteration stopped during timing (status %d)",
               ctx->h_st[3].status);
    *avg_us = 1e3 * total_ms / reps;
    return KM_OK;
}

int km_dot(km_ctx* ctx, int64_t n, int prec, const void* a, const void* b, double* out) {
    if (!ctx) return KM_EINVAL;
    KM_REQUIRE(ctx, n >= 0 && a && b && out, "km_dot: invalid arguments");
    KM_REQUIRE(ctx, prec == KM_F64 || prec == KM_F32, "unknown precision %d", prec);
    if (int rc = km_ensure_workspace(ctx, 0)) return rc;
    if (n == 0) { *out = 0.0; return KM_OK; }
    return dot_sync(ctx, n, prec, a, b, out);
}

int km_residual(km_ctx* ctx, int NX, int NY, const double* coef, const double* b, const double* x, double* r,
                double* rnorm2) {
    if (!ctx) return KM_EINVAL;
    if (int rc = check_system(ctx, NX, NY)) return rc;
    KM_REQUIRE(ctx, coef && b && x, "km_residual: NULL pointer");
    if (int rc = km_ensure_workspace(ctx, 0)) return rc;
    Tile tl;
    if (int rc = get_tile(ctx, NX, NY, KM_F64, &tl)) return rc;
    double dummy = 0.0;
    return residual_norm2(ctx, tl, coef, b, x, r, rnorm2 != nullptr, rnorm2 ? rnorm2 : &dummy);
}

}  // extern "C"

int km_residual_async(km_ctx* ctx, int NX, int NY, const double* coef, const double* b, const double* x, double* r,
                      int slot) {
    Tile tl;
    if (int rc = get_tile(ctx, NX, NY, KM_F64, &tl)) return rc;
    double* P = ctx->partials + KM_PX;
    hipLaunchKernelGGL(k_residual, dim3(tl.G), dim3(KM_BLOCK), tile_lds<double>(tl), ctx->stream, tl, coef, b, x, r, P);
    KM_REQUIRE(ctx, slot >= 0 && slot < 16, "result slot %d", slot);
    double* out = reinterpret_cast<double*>(ctx->d_flag + KM_HW_RES) + slot;   // host memory: no copy back
    hipLaunchKernelGGL(k_finish<1>, dim3(1), dim3(KM_BLOCK), 0, ctx->stream, (const double*)P, tl.G, out);
    KM_CHECK_LAUNCH(ctx);
    return KM_OK;
}

// Where a plain solve keeps its scalar state, partial sums, reduced results
// and its eight Krylov vectors (km_bicgstab: the context's workspace;
// km_bicgstab_noalloc: the library's scratch array).
struct PlainMem {
    KState* st;
    double* P;
    double* dres;
    double* ws;
};

// Plain BiCGSTAB (bicgsolver.py:88-136) on the memory of m, with the true
// residual after the solve when info is given.  KM_FLAG_TRUE: restarts from
// x while the true residual stays above eps (see include/karman.h).
int bicgstab_plain(km_ctx* ctx, const PlainMem& m, const Tile& tl, const double* coef, const double* b, double* x,
                   double eps, int max_iter, int flags, bool use_graph, km_solve_info* info) {
    const int64_t n = (int64_t)tl.NX * tl.NY;
    const bool from_r = (flags & KM_FLAG_X0) != 0;
    const bool check_true = (flags & KM_FLAG_TRUE) != 0;
    double bb = 0.0;
    if (from_r || check_true)
        if (int rc = dot_sync(ctx, n, KM_F64, b, b, &bb, m.P, m.dres)) return rc;
    if (from_r) {
        // r = b - A x0; the stopping test stays relative to ||b||
        if (int rc = residual_norm2(ctx, tl, coef, b, x, m.ws /* r */, false, nullptr, m.P, m.dres)) return rc;
    }
    KState s;
    if (int rc = bicgstab_typed<double>(ctx, tl, coef, b, x, m.ws, from_r, eps, max_iter, use_graph, m.st, &s,
                                        from_r ? bb : 0.0, m.P))
        return rc;
    int total = s.iter, restarts = 0;
    double rel_true = -1.0;
    if (check_true && bb > 0.0) {
        double rr = 0.0;
        if (int rc = residual_norm2(ctx, tl, coef, b, x, m.ws /* r */, true, &rr, m.P, m.dres)) return rc;
        rel_true = sqrt(rr / bb);
        while (s.status == KM_OK && rel_true >= eps && restarts < 8 && total < max_iter) {
            // r = b - A x is in place: restart from x, still relative to ||b||
            const double prev = rel_true;
            if (int rc = bicgstab_typed<double>(ctx, tl, coef, b, x, m.ws, true, eps, max_iter - total, use_graph,
                                                m.st, &s, bb, m.P))
                return rc;
            total += s.iter;
            restarts++;
            if (int rc = residual_norm2(ctx, tl, coef, b, x, m.ws, true, &rr, m.P, m.dres)) return rc;
            rel_true = sqrt(rr / bb);
            if (!(rel_true < 0.5 * prev)) break;
        }
    }
    fill_info(info, s);
    if (info) {
        if (rel_true < 0.0) {
            double rr = 0.0;
            if (!from_r && !check_true)
                if (int rc = dot_sync(ctx, n, KM_F64, b, b, &bb, m.P, m.dres)) return rc;
            if (int rc = residual_norm2(ctx, tl, coef, b, x, nullptr, true, &rr, m.P, m.dres)) return rc;
            rel_true = bb > 0 ? sqrt(rr / bb) : 0.0;
        }
        info->iterations = total;
        info->bnorm = sqrt(bb);
        info->rel_resid_true = rel_true;
        info->refinements = restarts;
        info->fallback = 0;
    }
    return KM_OK;
}

// The no-allocation entry point's scratch: one file-scope device array per
// GPU (each device holds its own copy from the library's load until the
// process ends).  Layout: KState slots | partial sums (the plain solve's
// regions, up to KM_PX + 2 KM_MAX_RED doubles) | 64 reduced results | the
// eight Krylov vectors (8 n doubles).  In host code `scratch` only names the
// symbol; its device address comes from hipGetSymbolAddress (ctx->scratch).
constexpr size_t kScratchStateBytes = 4096;
constexpr size_t kScratchPartialBytes = (size_t)(KM_PX + 2 * KM_MAX_RED) * sizeof(double);
constexpr size_t kScratchHeaderBytes = kScratchStateBytes + kScratchPartialBytes + 64 * sizeof(double);
static_assert(8 * sizeof(KState) <= kScratchStateBytes, "KState slots exceed their scratch region");
static_assert(kScratchHeaderBytes % 256 == 0, "the scratch vectors must stay 256-byte aligned");
static __device__ __align__(256) unsigned char scratch[KM_NOALLOC_SCRATCH_BYTES];

static_assert((KM_NOALLOC_SCRATCH_BYTES - kScratchHeaderBytes) / (8 * sizeof(double)) == KM_NOALLOC_MAX_CELLS,
              "KM_NOALLOC_MAX_CELLS does not match the scratch layout");

size_t scratch_need(int64_t n) { return kScratchHeaderBytes + (size_t)n * 8 * sizeof(double); }
int64_t scratch_max_cells() { return (int64_t)((KM_NOALLOC_SCRATCH_BYTES - kScratchHeaderBytes) / (8 * sizeof(double))); }

// The tile of an n-cell system without timing candidates or caching a choice
// (the autotuner's timing needs workspace): an override, else the cached
// choice, else the default split of km_choose_tile.
int peek_tile(km_ctx* ctx, int NX, int NY, Tile* out) {
    auto key = std::make_tuple(NX, NY, (int)KM_F64);
    int TI, TJ;
    auto ov = ctx->tile_override.find(key);
    auto it = ctx->tiles.find(key);
    if (ov != ctx->tile_override.end()) {
        TI = ov->second.first;
        TJ = ov->second.second;
    } else if (it != ctx->tiles.end()) {
        TI = it->second.first;
        TJ = it->second.second;
    } else {
        const int segs = (NY + 255) / 256;
        TJ = (NY + segs - 1) / segs;
        TI = (256 + TJ - 1) / TJ;
        if (TI > NX) TI = NX;
    }
    KM_REQUIRE(ctx, tile_ok(TI, TJ), "tile %dx%d is not valid for the iteration kernels", TI, TJ);
    *out = make_tile(NX, NY, TI, TJ);
    return KM_OK;
}

int km_resolve_scratch(km_ctx* ctx) {
    static std::mutex mu;
    static void* addr[64] = {};
    std::lock_guard<std::mutex> lk(mu);
    KM_REQUIRE(ctx, ctx->device >= 0 && ctx->device < 64, "device %d", ctx->device);
    if (!addr[ctx->device]) {
        void* p = nullptr;
        KM_HIP(ctx, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        addr[ctx->device] = p;
    }
    ctx->scratch = addr[ctx->device];
    return KM_OK;
}

extern "C" {

int km_bicgstab(km_ctx* ctx, int NX, int NY, const double* coef, const double* b, double* x, double eps,
                int max_iter, int flags, km_solve_info* info) {
    if (!ctx) return KM_EINVAL;
    if (int rc = check_system(ctx, NX, NY)) return rc;
    KM_REQUIRE(ctx, coef && b && x, "km_bicgstab: NULL pointer");
    KM_REQUIRE(ctx, eps >= 0.0, "eps must be non-negative");
    const int64_t n = (int64_t)NX * NY;
    if (max_iter <= 0) max_iter = (int)n;   // cgsolver.py:14
    if (flags & KM_FLAG_MIXED) return bicgstab_mixed(ctx, NX, NY, coef, b, x, eps, max_iter, flags, info);
    if (int rc = km_ensure_workspace(ctx, (size_t)9 * n * sizeof(double))) return rc;
    double* ws = static_cast<double*>(km_ws_vectors(ctx));
    if (int rc = ensure_tuned<double>(ctx, NX, NY, KM_F64, coef, b, ws)) return rc;
    Tile tl;
    if (int rc = get_tile(ctx, NX, NY, KM_F64, &tl)) return rc;
    const PlainMem m{&ctx->st[0], ctx->partials, ctx->dres, ws};
    return bicgstab_plain(ctx, m, tl, coef, b, x, eps, max_iter, flags, !(flags & KM_FLAG_NO_GRAPH), info);
}

int km_bicgstab_noalloc(km_ctx* ctx, int NX, int NY, const double* coef, const double* b, double* x, double eps,
                        int max_iter, int flags, km_solve_info* info) {
    if (!ctx) return KM_EINVAL;
    if (int rc = check_system(ctx, NX, NY)) return rc;
    KM_REQUIRE(ctx, coef && b && x, "km_bicgstab_noalloc: NULL pointer");
    KM_REQUIRE(ctx, eps >= 0.0, "eps must be non-negative");
    KM_REQUIRE(ctx, !(flags & KM_FLAG_MIXED), "km_bicgstab_noalloc has no mixed-precision path");
    const int64_t n = (int64_t)NX * NY;
    const size_t need = scratch_need(n);
    KM_REQUIRE(ctx, need <= KM_NOALLOC_SCRATCH_BYTES,
               "km_bicgstab_noalloc: a %d x %d system needs %zu bytes of scratch; the library reserves %zu "
               "(at most %lld cells)",
               NX, NY, need, (size_t)KM_NOALLOC_SCRATCH_BYTES, (long long)scratch_max_cells());
    KM_REQUIRE(ctx, ctx->scratch, "km_bicgstab_noalloc: the scratch address of device %d was not resolved",
               ctx->device);
    if (max_iter <= 0) max_iter = (int)n;   // cgsolver.py:14
    Tile tl;
    if (int rc = peek_tile(ctx, NX, NY, &tl)) return rc;
    char* base = static_cast<char*>(ctx->scratch);
    const PlainMem m{reinterpret_cast<KState*>(base), reinterpret_cast<double*>(base + kScratchStateBytes),
                     reinterpret_cast<double*>(base + kScratchStateBytes + kScratchPartialBytes),
                     reinterpret_cast<double*>(base + kScratchHeaderBytes)};
    // direct launches only: instantiating a graph allocates
    return bicgstab_plain(ctx, m, tl, coef, b, x, eps, max_iter, flags, false, info);
}

int km_cg(km_ctx* ctx, int NX, int NY, const double* coef, const double* b, double* x, double eps, int max_iter,
          int flags, km_solve_info* info) {
    if (!ctx) return KM_EINVAL;
    if (int rc = check_system(ctx, NX, NY)) return rc;
    KM_REQUIRE(ctx, coef && b && x, "km_cg: NULL pointer");
    KM_REQUIRE(ctx, !(flags & (KM_FLAG_X0 | KM_FLAG_MIXED)), "km_cg supports neither x0 nor mixed precision");
    const int64_t n = (int64_t)NX * NY;
    if (max_iter <= 0) max_iter = (int)n;
    if (int rc = km_ensure_workspace(ctx, (size_t)9 * n * sizeof(double))) return rc;
    double* ws = static_cast<double*>(km_ws_vectors(ctx));
    if (int rc = ensure_tuned<double>(ctx, NX, NY, KM_F64, coef, b, ws)) return rc;
    Tile tl;
    if (int rc = get_tile(ctx, NX, NY, KM_F64, &tl)) return rc;
    KState s;
    if (int rc = cg_typed<double>(ctx, tl, coef, b, x, ws, eps, max_iter, !(flags & KM_FLAG_NO_GRAPH), &ctx->st[0], &s))
        return rc;
    fill_info(info, s);
    if (info) {
        double bb = 0.0, rr = 0.0;
        if (int rc = dot_sync(ctx, n, KM_F64, b, b, &bb)) return rc;
        if (int rc = residual_norm2(ctx, tl, coef, b, x, nullptr, true, &rr)) return rc;
        info->bnorm = sqrt(bb);
        info->rel_resid_true = bb > 0 ? sqrt(rr / bb) : 0.0;
        info->refinements = 0;
        info->fallback = 0;
    }
    return KM_OK;
}

}  // extern "C"
