// Krylov solvers on the implicit 5-point stencil for gfx950.
//
// The reference runs one BiCGSTAB iteration as 14 kernel launches and 5 host
// round trips (bicgsolver.py:88-136).  Here an iteration is three kernels:
//
//   A  p' = r + beta (p - omega v) on an LDS tile + halo, v' = A p',  <r~, v'>
//   B  s  = r - alpha v'           on an LDS tile + halo, t  = A s,   <t, s>, <t, t>
//   C  x += alpha p' + omega s,  r = s - omega t,                     <r, r>, <r, r~>
//
// Each reduction is finished by the last workgroup to arrive, which sums the
// per-workgroup partials in a fixed order (bitwise reproducible) and updates
// the FP64 scalars (rho, alpha, omega, beta), the convergence test and the
// iteration counter in device memory.  Kernels return immediately once the
// device-side ``done`` flag is set, so the host enqueues fixed-size chunks of
// iterations (replayed from a captured hipGraph) and polls the flag between
// chunks instead of after every reduction.  p and v are double-buffered by
// iteration parity because kernel A reads the old values in its halo while
// writing the new ones.
//
// T = double (plain FP64 path) or float (inner solves of the mixed-precision
// path); dot products and scalars are always FP64.
#include "km_common.h"

extern __shared__ __align__(16) unsigned char km_smem[];

void* km_ws_vectors(km_ctx* ctx);

namespace {

struct TileOrigin {
    int i0, j0;
};

__device__ __forceinline__ TileOrigin tile_origin(const Tile& tl) {
    const int L = xcd_remap(blockIdx.x, tl.G);
    const int ti = L / tl.tiles_j, tj = L - ti * tl.tiles_j;
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

__device__ __forceinline__ bool finite(double x) { return x == x && x - x == 0.0; }

__device__ __forceinline__ void set_done(KState* st, int status) {
    st->status = status;
    st->done = 1;
}

// ------------------------------------------------------------ plain kernels

template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_stencil(Tile tl, const T* __restrict__ coef, const T* __restrict__ x, T* __restrict__ y) {
    T* lds = reinterpret_cast<T*>(km_smem);
    const TileOrigin o = tile_origin(tl);
    stage<T>(tl, o, lds, [&](int64_t k) { return x[k]; }, (T*)nullptr);
    __syncthreads();
    apply<T>(tl, o, lds, coef, [&](int64_t k, T v, T) { y[k] = v; });
}

// r = b - A x (FP64) and ||r||^2 into res[0]
__global__ void __launch_bounds__(KM_BLOCK)
k_residual(Tile tl, const double* __restrict__ coef, const double* __restrict__ b, const double* __restrict__ x,
           double* __restrict__ r, double* __restrict__ partials, uint32_t* __restrict__ counter,
           double* __restrict__ res) {
    double* lds = reinterpret_cast<double*>(km_smem);
    __shared__ double sh[4];
    const TileOrigin o = tile_origin(tl);
    stage<double>(tl, o, lds, [&](int64_t k) { return x[k]; }, (double*)nullptr);
    __syncthreads();
    double acc = 0.0;
    apply<double>(tl, o, lds, coef, [&](int64_t k, double y, double) {
        const double rr = b[k] - y;
        if (r) r[k] = rr;
        acc += rr * rr;
    });
    double vals[1] = {block_sum(acc, sh)};
    if (publish_partials<1>(vals, blockIdx.x, tl.G, partials, counter)) {
        double out[1];
        finish_partials<1>(partials, tl.G, out, sh);
        if (threadIdx.x == 0) res[0] = out[0];
    }
}

template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_dot(int64_t n, const T* __restrict__ a, const T* __restrict__ b, double* __restrict__ partials,
      uint32_t* __restrict__ counter, double* __restrict__ res) {
    __shared__ double sh[4];
    double acc = 0.0;
    for (int64_t k = (int64_t)blockIdx.x * KM_BLOCK + threadIdx.x; k < n; k += (int64_t)gridDim.x * KM_BLOCK)
        acc += (double)a[k] * (double)b[k];
    double vals[1] = {block_sum(acc, sh)};
    if (publish_partials<1>(vals, blockIdx.x, gridDim.x, partials, counter)) {
        double out[1];
        finish_partials<1>(partials, gridDim.x, out, sh);
        if (threadIdx.x == 0) res[0] = out[0];
    }
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

// bicgsolver.py:21-35: x = 0, r = r~ = b (or r precomputed from x0), the
// parity-0 p and v buffers zeroed, rho_1 = alpha = omega = 1; r0 = <r, r>.
template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_bicg_init(int64_t n, const T* __restrict__ b, T* __restrict__ x, T* __restrict__ r, T* __restrict__ rt,
            T* __restrict__ p0, T* __restrict__ v0, int from_r, double eps, int max_iter, double ref_norm2,
            KState* st, double* __restrict__ partials) {
    __shared__ double sh[4];
    double acc = 0.0;
    for (int64_t k = (int64_t)blockIdx.x * KM_BLOCK + threadIdx.x; k < n; k += (int64_t)gridDim.x * KM_BLOCK) {
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
        acc += (double)rk * (double)rk;
    }
    double vals[1] = {block_sum(acc, sh)};
    if (publish_partials<1>(vals, blockIdx.x, gridDim.x, partials, &st->counter[0])) {
        double out[1];
        finish_partials<1>(partials, gridDim.x, out, sh);
        if (threadIdx.x == 0) {
            const double r0 = out[0];
            // stopping test denominator: r0.r0 (bicgsolver.py:90), or ||b||^2
            // for a warm start so that x0 is not credited through it
            st->r0 = ref_norm2 > 0.0 ? ref_norm2 : r0;
            st->rTr = r0;
            st->rho = r0;          // rho of iteration 0 is <r, r~> = <r, r>
            st->rho_1 = 1.0;
            st->alpha = 1.0;
            st->omega = 1.0;
            st->beta = 0.0;        // p' = r + 0 (0 - omega 0) = r on iteration 0
            st->eps = eps;
            st->max_iter = max_iter;
            st->iter = 0;
            st->status = KM_MAXITER;
            st->done = 0;
            if (!finite(r0)) set_done(st, KM_NONFINITE);
            else if (r0 == 0.0) set_done(st, KM_OK);
            else if (max_iter <= 0) set_done(st, KM_MAXITER);
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_bicg_a(Tile tl, const T* __restrict__ coef, const T* __restrict__ r, const T* __restrict__ rt, T* pbuf, T* vbuf,
         KState* st, double* __restrict__ partials) {
    if (st->done) return;
    T* lds = reinterpret_cast<T*>(km_smem);
    __shared__ double sh[4];
    const int64_t N = (int64_t)tl.NX * tl.NY;
    const int par = st->iter & 1;
    const T* __restrict__ p_old = pbuf + par * N;
    const T* __restrict__ v_old = vbuf + par * N;
    T* __restrict__ p_new = pbuf + (1 - par) * N;
    T* __restrict__ v_new = vbuf + (1 - par) * N;
    const T beta = (T)st->beta, omega = (T)st->omega;
    const TileOrigin o = tile_origin(tl);
    // update_p (bicgsolver.py:43-46): p = r + beta (p - omega v)
    stage<T>(tl, o, lds, [&](int64_t k) { return r[k] + beta * (p_old[k] - omega * v_old[k]); }, p_new);
    __syncthreads();
    double acc = 0.0;
    // compute_Ap (bicgsolver.py:72-76) and reduce(r~, v) (bicgsolver.py:107)
    apply<T>(tl, o, lds, coef, [&](int64_t k, T y, T) {
        v_new[k] = y;
        acc += (double)rt[k] * (double)y;
    });
    double vals[1] = {block_sum(acc, sh)};
    if (publish_partials<1>(vals, blockIdx.x, tl.G, partials, &st->counter[1])) {
        double out[1];
        finish_partials<1>(partials, tl.G, out, sh);
        if (threadIdx.x == 0) {
            const double rv = out[0];
            if (!finite(rv)) set_done(st, KM_NONFINITE);
            else if (rv == 0.0) set_done(st, KM_BREAKDOWN);
            else st->alpha = st->rho / rv;   // bicgsolver.py:108
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_bicg_b(Tile tl, const T* __restrict__ coef, const T* __restrict__ r, const T* vbuf, T* __restrict__ s,
         T* __restrict__ t, KState* st, double* __restrict__ partials) {
    if (st->done) return;
    T* lds = reinterpret_cast<T*>(km_smem);
    __shared__ double sh[4];
    const int64_t N = (int64_t)tl.NX * tl.NY;
    const int par = st->iter & 1;
    const T* __restrict__ v = vbuf + (1 - par) * N;
    const T alpha = (T)st->alpha;
    const TileOrigin o = tile_origin(tl);
    // update_s (bicgsolver.py:60-63): s = r - alpha v
    stage<T>(tl, o, lds, [&](int64_t k) { return r[k] - alpha * v[k]; }, s);
    __syncthreads();
    double ts = 0.0, tt = 0.0;
    // compute_As (bicgsolver.py:65-69), reduce(t, s), reduce(t, t) (:116-117)
    apply<T>(tl, o, lds, coef, [&](int64_t k, T y, T sk) {
        t[k] = y;
        ts += (double)y * (double)sk;
        tt += (double)y * (double)y;
    });
    double vals[2] = {block_sum(ts, sh), block_sum(tt, sh)};
    if (publish_partials<2>(vals, blockIdx.x, tl.G, partials, &st->counter[2])) {
        double out[2];
        finish_partials<2>(partials, tl.G, out, sh);
        if (threadIdx.x == 0) {
            if (!finite(out[0]) || !finite(out[1])) set_done(st, KM_NONFINITE);
            else st->omega = out[1] != 0.0 ? out[0] / out[1] : 0.0;   // bicgsolver.py:118
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_bicg_c(int64_t n, T* __restrict__ x, const T* pbuf, const T* __restrict__ s, const T* __restrict__ t,
         T* __restrict__ r, const T* __restrict__ rt, KState* st, double* __restrict__ partials) {
    if (st->done) return;
    __shared__ double sh[4];
    const int par = st->iter & 1;
    const T* __restrict__ p = pbuf + (1 - par) * n;
    const T alpha = (T)st->alpha, omega = (T)st->omega;
    double rr = 0.0, rrt = 0.0;
    for (int64_t k = (int64_t)blockIdx.x * KM_BLOCK + threadIdx.x; k < n; k += (int64_t)gridDim.x * KM_BLOCK) {
        // update_x, update_r (bicgsolver.py:78-86)
        x[k] = x[k] + (alpha * p[k] + omega * s[k]);
        const T rk = s[k] - omega * t[k];
        r[k] = rk;
        rr += (double)rk * (double)rk;
        rrt += (double)rk * (double)rt[k];
    }
    double vals[2] = {block_sum(rr, sh), block_sum(rrt, sh)};
    if (publish_partials<2>(vals, blockIdx.x, gridDim.x, partials, &st->counter[3])) {
        double out[2];
        finish_partials<2>(partials, gridDim.x, out, sh);
        if (threadIdx.x == 0) {
            const double rTr = out[0], rho_new = out[1];
            st->rTr = rTr;
            st->iter = st->iter + 1;
            if (!finite(rTr)) set_done(st, KM_NONFINITE);
            else if (sqrt(rTr / st->r0) < st->eps) set_done(st, KM_OK);       // bicgsolver.py:130
            else if (st->omega == 0.0) set_done(st, KM_BREAKDOWN);           // karman_taichi.py:451
            else if (st->iter >= st->max_iter) set_done(st, KM_MAXITER);
            else {
                st->rho_1 = st->rho;                                           // bicgsolver.py:135
                st->rho = rho_new;                                             // bicgsolver.py:95
                if (!finite(rho_new)) set_done(st, KM_NONFINITE);
                else if (rho_new == 0.0) set_done(st, KM_BREAKDOWN);           // bicgsolver.py:96-98
                else st->beta = (rho_new / st->rho_1) * (st->alpha / st->omega);   // bicgsolver.py:102
            }
        }
    }
}

// ------------------------------------------------------------ CG (cgsolver.py:83-114)

template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_cg_init(int64_t n, const T* __restrict__ b, T* __restrict__ x, T* __restrict__ r, T* __restrict__ p0,
          double eps, int max_iter, KState* st, double* __restrict__ partials) {
    __shared__ double sh[4];
    double acc = 0.0;
    for (int64_t k = (int64_t)blockIdx.x * KM_BLOCK + threadIdx.x; k < n; k += (int64_t)gridDim.x * KM_BLOCK) {
        const T rk = b[k];
        r[k] = rk;
        x[k] = T(0);
        p0[k] = T(0);
        acc += (double)rk * (double)rk;
    }
    double vals[1] = {block_sum(acc, sh)};
    if (publish_partials<1>(vals, blockIdx.x, gridDim.x, partials, &st->counter[0])) {
        double out[1];
        finish_partials<1>(partials, gridDim.x, out, sh);
        if (threadIdx.x == 0) {
            const double r0 = out[0];
            st->r0 = r0;
            st->rTr = r0;
            st->rho = r0;     // old_rTr
            st->beta = 0.0;   // p = r + 0 p (cgsolver.py:228)
            st->alpha = 0.0;
            st->eps = eps;
            st->max_iter = max_iter;
            st->iter = 0;
            st->status = KM_MAXITER;
            st->done = 0;
            if (!finite(r0)) set_done(st, KM_NONFINITE);
            else if (r0 == 0.0) set_done(st, KM_OK);
            else if (max_iter <= 0) set_done(st, KM_MAXITER);
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_cg_a(Tile tl, const T* __restrict__ coef, const T* __restrict__ r, T* pbuf, T* __restrict__ Ap, KState* st,
       double* __restrict__ partials) {
    if (st->done) return;
    T* lds = reinterpret_cast<T*>(km_smem);
    __shared__ double sh[4];
    const int64_t N = (int64_t)tl.NX * tl.NY;
    const int par = st->iter & 1;
    const T* __restrict__ p_old = pbuf + par * N;
    T* __restrict__ p_new = pbuf + (1 - par) * N;
    const T beta = (T)st->beta;
    const TileOrigin o = tile_origin(tl);
    // update_p (cgsolver.py:216-219): p = r + beta p
    stage<T>(tl, o, lds, [&](int64_t k) { return r[k] + beta * p_old[k]; }, p_new);
    __syncthreads();
    double acc = 0.0;
    apply<T>(tl, o, lds, coef, [&](int64_t k, T y, T pk) {
        Ap[k] = y;
        acc += (double)pk * (double)y;
    });
    double vals[1] = {block_sum(acc, sh)};
    if (publish_partials<1>(vals, blockIdx.x, tl.G, partials, &st->counter[1])) {
        double out[1];
        finish_partials<1>(partials, tl.G, out, sh);
        if (threadIdx.x == 0) {
            const double pAp = out[0];
            if (!finite(pAp)) set_done(st, KM_NONFINITE);
            else if (pAp == 0.0) set_done(st, KM_BREAKDOWN);
            else st->alpha = st->rho / pAp;   // cgsolver.py:233-234
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_cg_b(int64_t n, T* __restrict__ x, T* __restrict__ r, const T* pbuf, const T* __restrict__ Ap, KState* st,
       double* __restrict__ partials) {
    if (st->done) return;
    __shared__ double sh[4];
    const int par = st->iter & 1;
    const T* __restrict__ p = pbuf + (1 - par) * n;
    const T alpha = (T)st->alpha;
    double acc = 0.0;
    for (int64_t k = (int64_t)blockIdx.x * KM_BLOCK + threadIdx.x; k < n; k += (int64_t)gridDim.x * KM_BLOCK) {
        x[k] = x[k] + alpha * p[k];          // cgsolver.py:206-209
        const T rk = r[k] - alpha * Ap[k];   // cgsolver.py:211-214
        r[k] = rk;
        acc += (double)rk * (double)rk;
    }
    double vals[1] = {block_sum(acc, sh)};
    if (publish_partials<1>(vals, blockIdx.x, gridDim.x, partials, &st->counter[2])) {
        double out[1];
        finish_partials<1>(partials, gridDim.x, out, sh);
        if (threadIdx.x == 0) {
            const double nw = out[0];
            st->rTr = nw;
            st->iter = st->iter + 1;
            if (!finite(nw)) set_done(st, KM_NONFINITE);
            else if (sqrt(nw) < sqrt(st->r0) * st->eps) set_done(st, KM_OK);   // cgsolver.py:240
            else if (st->iter >= st->max_iter) set_done(st, KM_MAXITER);
            else {
                st->beta = nw / st->rho;   // cgsolver.py:245
                st->rho = nw;
            }
        }
    }
}

// ------------------------------------------------------------ host helpers

int flat_grid(int64_t n) {
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
    *out = make_tile(NX, NY, TI, TJ);
    KM_REQUIRE(ctx, 2 * (int64_t)out->G <= KM_MAX_PARTIALS, "tile grid of %d workgroups exceeds the partial buffer", out->G);
    return KM_OK;
}

// Poll loop: enqueue chunks of iterations (a captured graph, or direct
// launches), two chunks in flight, until the device sets ``done``.
template <typename F>
int run_chunks(km_ctx* ctx, KState* st, GraphKey key, bool use_graph, int chunk, int max_iter, F enqueue,
               KState* final_state) {
    hipGraphExec_t exec = nullptr;
    if (use_graph) {
        key.v.push_back((uintptr_t)chunk);
        key.v.push_back((uintptr_t)ctx->ws_gen);
        auto it = ctx->graphs.find(key);
        if (it != ctx->graphs.end()) {
            exec = it->second;
        } else {
            hipGraph_t graph = nullptr;
            KM_HIP(ctx, hipStreamBeginCapture(ctx->cap_stream, hipStreamCaptureModeThreadLocal));
            for (int c = 0; c < chunk; ++c) enqueue(ctx->cap_stream);
            hipError_t le = hipGetLastError();
            hipError_t ce = hipStreamEndCapture(ctx->cap_stream, &graph);
            if (le != hipSuccess || ce != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                return km_fail(ctx, KM_EHIP, "graph capture failed: %s / %s", hipGetErrorString(le),
                               hipGetErrorString(ce));
            }
            hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ie != hipSuccess) return km_fail(ctx, KM_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(ie));
            ctx->graphs[key] = exec;
        }
    }
    auto launch_chunk = [&]() -> int {
        if (exec) {
            KM_HIP(ctx, hipGraphLaunch(exec, ctx->stream));
        } else {
            for (int c = 0; c < chunk; ++c) enqueue(ctx->stream);
            KM_CHECK_LAUNCH(ctx);
        }
        return KM_OK;
    };
    // a device-side cap ends every solve; the host cap only guards against a
    // state that never reports done because a launch failed
    const int64_t host_cap = (int64_t)max_iter / chunk + 4;
    if (int rc = launch_chunk()) return rc;
    KM_HIP(ctx, hipMemcpyAsync(&ctx->h_st[0], st, sizeof(KState), hipMemcpyDeviceToHost, ctx->stream));
    KM_HIP(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    for (int64_t c = 1;; ++c) {
        if (c > host_cap) return km_fail(ctx, KM_EHIP, "solver did not report completion after %lld chunks", (long long)c);
        if (int rc = launch_chunk()) return rc;
        KM_HIP(ctx, hipMemcpyAsync(&ctx->h_st[c & 1], st, sizeof(KState), hipMemcpyDeviceToHost, ctx->stream));
        KM_HIP(ctx, hipEventRecord(ctx->ev[c & 1], ctx->stream));
        KM_HIP(ctx, hipEventSynchronize(ctx->ev[(c - 1) & 1]));
        if (ctx->h_st[(c - 1) & 1].done) break;
    }
    KM_HIP(ctx, hipMemcpyAsync(&ctx->h_st[2], st, sizeof(KState), hipMemcpyDeviceToHost, ctx->stream));
    KM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *final_state = ctx->h_st[2];
    return KM_OK;
}

const int kChunk = 8;

// Plain BiCGSTAB on typed buffers.  ws holds 8n elements of T.
template <typename T>
int bicgstab_typed(km_ctx* ctx, const Tile& tl, const T* coef, const T* b, T* x, T* ws, bool from_r, double eps,
                   int max_iter, bool use_graph, KState* st, KState* out, double ref_norm2 = 0.0) {
    const int64_t n = (int64_t)tl.NX * tl.NY;
    T* r = ws;
    T* rt = ws + n;
    T* pbuf = ws + 2 * n;   // 2n
    T* vbuf = ws + 4 * n;   // 2n
    T* s = ws + 6 * n;
    T* t = ws + 7 * n;
    const int fg = flat_grid(n);
    hipLaunchKernelGGL(k_bicg_init<T>, dim3(fg), dim3(KM_BLOCK), 0, ctx->stream, n, b, x, r, rt, pbuf, vbuf,
                       from_r ? 1 : 0, eps, max_iter, ref_norm2, st, ctx->partials);
    KM_CHECK_LAUNCH(ctx);
    const size_t lds = tile_lds<T>(tl);
    double* partials = ctx->partials;
    auto enqueue = [=](hipStream_t s_) {
        hipLaunchKernelGGL(k_bicg_a<T>, dim3(tl.G), dim3(KM_BLOCK), lds, s_, tl, coef, r, rt, pbuf, vbuf, st, partials);
        hipLaunchKernelGGL(k_bicg_b<T>, dim3(tl.G), dim3(KM_BLOCK), lds, s_, tl, coef, r, vbuf, s, t, st, partials);
        hipLaunchKernelGGL(k_bicg_c<T>, dim3(fg), dim3(KM_BLOCK), 0, s_, n, x, pbuf, s, t, r, rt, st, partials);
    };
    GraphKey key;
    key.v = {1, sizeof(T), (uintptr_t)tl.NX, (uintptr_t)tl.NY, (uintptr_t)tl.TI, (uintptr_t)tl.TJ,
             (uintptr_t)coef, (uintptr_t)x, (uintptr_t)ws, (uintptr_t)st};
    return run_chunks(ctx, st, key, use_graph, kChunk, max_iter, enqueue, out);
}

template <typename T>
int cg_typed(km_ctx* ctx, const Tile& tl, const T* coef, const T* b, T* x, T* ws, double eps, int max_iter,
             bool use_graph, KState* st, KState* out) {
    const int64_t n = (int64_t)tl.NX * tl.NY;
    T* r = ws;
    T* pbuf = ws + n;   // 2n
    T* Ap = ws + 3 * n;
    const int fg = flat_grid(n);
    hipLaunchKernelGGL(k_cg_init<T>, dim3(fg), dim3(KM_BLOCK), 0, ctx->stream, n, b, x, r, pbuf, eps, max_iter, st,
                       ctx->partials);
    KM_CHECK_LAUNCH(ctx);
    const size_t lds = tile_lds<T>(tl);
    double* partials = ctx->partials;
    auto enqueue = [=](hipStream_t s_) {
        hipLaunchKernelGGL(k_cg_a<T>, dim3(tl.G), dim3(KM_BLOCK), lds, s_, tl, coef, r, pbuf, Ap, st, partials);
        hipLaunchKernelGGL(k_cg_b<T>, dim3(fg), dim3(KM_BLOCK), 0, s_, n, x, r, pbuf, Ap, st, partials);
    };
    GraphKey key;
    key.v = {2, sizeof(T), (uintptr_t)tl.NX, (uintptr_t)tl.NY, (uintptr_t)tl.TI, (uintptr_t)tl.TJ,
             (uintptr_t)coef, (uintptr_t)x, (uintptr_t)ws, (uintptr_t)st};
    return run_chunks(ctx, st, key, use_graph, kChunk, max_iter, enqueue, out);
}

int residual_norm2(km_ctx* ctx, const Tile& tl, const double* coef, const double* b, const double* x, double* r,
                   bool sync, double* out) {
    hipLaunchKernelGGL(k_residual, dim3(tl.G), dim3(KM_BLOCK), tile_lds<double>(tl), ctx->stream, tl, coef, b, x, r,
                       ctx->partials, &ctx->st[1].counter[0], ctx->dres);
    KM_CHECK_LAUNCH(ctx);
    if (sync) {
        KM_HIP(ctx, hipMemcpyAsync(ctx->h_res, ctx->dres, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        KM_HIP(ctx, hipStreamSynchronize(ctx->stream));
        *out = ctx->h_res[0];
    }
    return KM_OK;
}

int dot_sync(km_ctx* ctx, int64_t n, int prec, const void* a, const void* b, double* out) {
    const int g = flat_grid(n);
    if (prec == KM_F64)
        hipLaunchKernelGGL(k_dot<double>, dim3(g), dim3(KM_BLOCK), 0, ctx->stream, n, (const double*)a,
                           (const double*)b, ctx->partials, &ctx->st[1].counter[1], ctx->dres + 1);
    else
        hipLaunchKernelGGL(k_dot<float>, dim3(g), dim3(KM_BLOCK), 0, ctx->stream, n, (const float*)a,
                           (const float*)b, ctx->partials, &ctx->st[1].counter[1], ctx->dres + 1);
    KM_CHECK_LAUNCH(ctx);
    KM_HIP(ctx, hipMemcpyAsync(ctx->h_res + 1, ctx->dres + 1, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    KM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *out = ctx->h_res[1];
    return KM_OK;
}

void fill_info(km_solve_info* info, const KState& s) {
    if (!info) return;
    info->iterations = s.iter;
    info->status = s.status;
    info->rel_resid = s.r0 > 0 ? sqrt(s.rTr / s.r0) : 0.0;
}

// Mixed precision: FP32 BiCGSTAB inner solves on a rounded copy of A with
// FP64 scalars, wrapped in FP64 iterative refinement judged by the true
// residual ||b - A x|| / ||b||.
int bicgstab_mixed(km_ctx* ctx, int NX, int NY, const double* coef, const double* b, double* x, double eps,
                   int max_iter, int flags, km_solve_info* info) {
    const int64_t n = (int64_t)NX * NY;
    Tile t64, t32;
    if (int rc = get_tile(ctx, NX, NY, KM_F64, &t64)) return rc;
    if (int rc = get_tile(ctx, NX, NY, KM_F32, &t32)) return rc;
    // FP64: r (n); FP32: coef (5n), rhs (n), d (n), inner workspace (8n)
    const size_t bytes = (size_t)n * sizeof(double) + (size_t)15 * n * sizeof(float) + 1024;
    if (int rc = km_ensure_workspace(ctx, bytes)) return rc;
    double* r = static_cast<double*>(km_ws_vectors(ctx));
    float* c32 = reinterpret_cast<float*>(r + n);
    float* rhs = c32 + 5 * n;
    float* d = rhs + n;
    float* w32 = d + n;
    const bool use_graph = !(flags & KM_FLAG_NO_GRAPH);
    const int fg = flat_grid(5 * n);
    hipLaunchKernelGGL(k_to_float, dim3(fg), dim3(KM_BLOCK), 0, ctx->stream, 5 * n, coef, c32, 1.0);
    KM_CHECK_LAUNCH(ctx);
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
        hipLaunchKernelGGL(k_to_float, dim3(flat_grid(n)), dim3(KM_BLOCK), 0, ctx->stream, n, (const double*)r, rhs,
                           1.0 / rn);
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
        hipLaunchKernelGGL(k_add_scaled, dim3(flat_grid(n)), dim3(KM_BLOCK), 0, ctx->stream, n, x, (const float*)d, rn);
        KM_CHECK_LAUNCH(ctx);
        const double prev = rel;
        if (int rc = residual_norm2(ctx, t64, coef, b, x, r, true, &rr)) return rc;
        rel = sqrt(rr) / bnorm;
        if (!(rel < 0.5 * prev)) {
            // FP32 recurrences lost accuracy: take the sweep back (or restart
            // from zero if it produced non-finite values)
            if (rel == rel && rel < 1e300) {
                hipLaunchKernelGGL(k_add_scaled, dim3(flat_grid(n)), dim3(KM_BLOCK), 0, ctx->stream, n, x,
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

// Tile choice: an explicit override, else the cached choice, else a default
// that splits each row into near-equal segments of at most 256 columns and
// stacks rows until a workgroup covers about 256 cells.
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

// Average device duration of one launch of a solver kernel, timed with HIP
// events around `reps` back-to-back launches on the context's stream.  The
// workspace is set up as for a solve of (coef, b); the kernels run with
// ``done`` clear so every launch does its full work.
int km_time_kernel(km_ctx* ctx, int which, int NX, int NY, const double* coef, const double* b, int reps,
                   double* avg_us) {
    if (!ctx) return KM_EINVAL;
    if (int rc = check_system(ctx, NX, NY)) return rc;
    KM_REQUIRE(ctx, coef && b && avg_us && reps > 0, "km_time_kernel: invalid arguments");
    KM_REQUIRE(ctx, which >= KM_KERNEL_BICG_A && which <= KM_KERNEL_STENCIL, "unknown kernel id %d", which);
    const int64_t n = (int64_t)NX * NY;
    Tile tl;
    if (int rc = get_tile(ctx, NX, NY, KM_F64, &tl)) return rc;
    if (int rc = km_ensure_workspace(ctx, (size_t)9 * n * sizeof(double))) return rc;
    double* ws = static_cast<double*>(km_ws_vectors(ctx));
    double *r = ws, *rt = ws + n, *pbuf = ws + 2 * n, *vbuf = ws + 4 * n, *sv = ws + 6 * n, *t = ws + 7 * n;
    double* x = ws + 8 * n;
    KState* st = &ctx->st[2];
    const int fg = flat_grid(n);
    hipLaunchKernelGGL(k_bicg_init<double>, dim3(fg), dim3(KM_BLOCK), 0, ctx->stream, n, b, x, r, rt, pbuf, vbuf, 0,
                       0.0, 1 << 30, 0.0, st, ctx->partials);
    KM_CHECK_LAUNCH(ctx);
    const size_t lds = tile_lds<double>(tl);
    auto launch = [&]() {
        switch (which) {
            case KM_KERNEL_BICG_A:
                hipLaunchKernelGGL(k_bicg_a<double>, dim3(tl.G), dim3(KM_BLOCK), lds, ctx->stream, tl, coef, r, rt,
                                   pbuf, vbuf, st, ctx->partials);
                break;
            case KM_KERNEL_BICG_B:
                hipLaunchKernelGGL(k_bicg_b<double>, dim3(tl.G), dim3(KM_BLOCK), lds, ctx->stream, tl, coef, r, vbuf,
                                   sv, t, st, ctx->partials);
                break;
            case KM_KERNEL_BICG_C:
                hipLaunchKernelGGL(k_bicg_c<double>, dim3(fg), dim3(KM_BLOCK), 0, ctx->stream, n, x, pbuf, sv, t, r,
                                   rt, st, ctx->partials);
                break;
            default:
                hipLaunchKernelGGL(k_stencil<double>, dim3(tl.G), dim3(KM_BLOCK), lds, ctx->stream, tl, coef, r, t);
        }
    };
    launch();   // warm-up
    hipEvent_t e0, e1;
    KM_HIP(ctx, hipEventCreate(&e0));
    KM_HIP(ctx, hipEventCreate(&e1));
    KM_HIP(ctx, hipEventRecord(e0, ctx->stream));
    for (int i = 0; i < reps; ++i) launch();
    KM_HIP(ctx, hipEventRecord(e1, ctx->stream));
    KM_HIP(ctx, hipEventSynchronize(e1));
    float ms = 0.f;
    KM_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    KM_CHECK_LAUNCH(ctx);
    *avg_us = 1e3 * (double)ms / reps;
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

int km_bicgstab(km_ctx* ctx, int NX, int NY, const double* coef, const double* b, double* x, double eps,
                int max_iter, int flags, km_solve_info* info) {
    if (!ctx) return KM_EINVAL;
    if (int rc = check_system(ctx, NX, NY)) return rc;
    KM_REQUIRE(ctx, coef && b && x, "km_bicgstab: NULL pointer");
    KM_REQUIRE(ctx, eps >= 0.0, "eps must be non-negative");
    const int64_t n = (int64_t)NX * NY;
    if (max_iter <= 0) max_iter = (int)n;   // cgsolver.py:14
    if (flags & KM_FLAG_MIXED) return bicgstab_mixed(ctx, NX, NY, coef, b, x, eps, max_iter, flags, info);
    Tile tl;
    if (int rc = get_tile(ctx, NX, NY, KM_F64, &tl)) return rc;
    if (int rc = km_ensure_workspace(ctx, (size_t)8 * n * sizeof(double))) return rc;
    double* ws = static_cast<double*>(km_ws_vectors(ctx));
    const bool from_r = (flags & KM_FLAG_X0) != 0;
    double bb = 0.0;
    if (from_r) {
        // r = b - A x0; the stopping test stays relative to ||b||
        if (int rc = dot_sync(ctx, n, KM_F64, b, b, &bb)) return rc;
        if (int rc = residual_norm2(ctx, tl, coef, b, x, ws /* r */, false, nullptr)) return rc;
    }
    KState s;
    if (int rc = bicgstab_typed<double>(ctx, tl, coef, b, x, ws, from_r, eps, max_iter,
                                        !(flags & KM_FLAG_NO_GRAPH), &ctx->st[0], &s, from_r ? bb : 0.0))
        return rc;
    fill_info(info, s);
    if (info) {
        double rr = 0.0;
        if (!from_r)
            if (int rc = dot_sync(ctx, n, KM_F64, b, b, &bb)) return rc;
        if (int rc = residual_norm2(ctx, tl, coef, b, x, nullptr, true, &rr)) return rc;
        info->bnorm = sqrt(bb);
        info->rel_resid_true = bb > 0 ? sqrt(rr / bb) : 0.0;
        info->refinements = 0;
        info->fallback = 0;
    }
    return KM_OK;
}

int km_cg(km_ctx* ctx, int NX, int NY, const double* coef, const double* b, double* x, double eps, int max_iter,
          int flags, km_solve_info* info) {
    if (!ctx) return KM_EINVAL;
    if (int rc = check_system(ctx, NX, NY)) return rc;
    KM_REQUIRE(ctx, coef && b && x, "km_cg: NULL pointer");
    KM_REQUIRE(ctx, !(flags & (KM_FLAG_X0 | KM_FLAG_MIXED)), "km_cg supports neither x0 nor mixed precision");
    const int64_t n = (int64_t)NX * NY;
    if (max_iter <= 0) max_iter = (int)n;
    Tile tl;
    if (int rc = get_tile(ctx, NX, NY, KM_F64, &tl)) return rc;
    if (int rc = km_ensure_workspace(ctx, (size_t)4 * n * sizeof(double))) return rc;
    double* ws = static_cast<double*>(km_ws_vectors(ctx));
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
