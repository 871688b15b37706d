// Flexible BiCGSTAB with the learned right preconditioner (km_unet.hip).
//
// Outer loop (FP64): r = b - A x, judged only by ||r|| / ||b||.  Each sweep
// solves A d = r / ||r|| with an inner flexible BiCGSTAB in FP32 vectors and
// FP64 scalars, where p^ = M(p) and s^ = M(s) are network outputs, so the
// preconditioner may differ between applications: x advances along the
// vectors actually produced (x += alpha p^ + omega s^) and v = A p^,
// t = A s^ are formed from them.  A sweep that does not halve the true
// residual is taken back and plain FP64 BiCGSTAB finishes the solve.
// Without a network (params == null) the preconditioner is Jacobi, D^-1, or
// with the coarse level the two-level preconditioner with red-black SOR
// sweeps, or with KM_FLAG_SMOOTH one or two red-black Gauss-Seidel sweeps (both the
// fused iterations of km_twolevel.hip).
//
// Inner iteration i (ring slot ph = i mod 4, as in the plain solver):
//   K_p  stopping test on <r, r>; p = r + beta (p - omega v)
//   M    p^ = M(p)
//   K_v  v = A p^, partial <r~, v>
//   K_s  alpha = rho / <r~, v>; s = r - alpha v
//   M    s^ = M(s)
//   K_t  t = A s^, partials <t, s>, <t, t>
//   K_x  omega = <t, s> / <t, t>; x += alpha p^ + omega s^; r = s - omega t; partials <r~, r>, <r, r>
#include "km_unet.h"
#include "km_coarse.h"
#include "km_mg.h"

#include <algorithm>
#include <cmath>

namespace {

constexpr int FB_PC = 0;                 // [2][KM_MAX_RED] <r~, r>, <r, r>
constexpr int FB_PV = 2 * KM_MAX_RED;    // [1] <r~, v>
constexpr int FB_PT = 3 * KM_MAX_RED;    // [2] <t, s>, <t, t>
constexpr int FB_PN = 5 * KM_MAX_RED;    // [1] p.p (the network's input scale)
constexpr int FB_PS = 6 * KM_MAX_RED;    // [1] s.s
constexpr int FB_PR = 7 * KM_MAX_RED;    // [1] r1.r1 after the coarse correction

__device__ __forceinline__ bool fin(double v) { return v == v && v - v == 0.0; }

template <typename T>
__device__ __forceinline__ T sten5(const T* __restrict__ c, const T* __restrict__ x, int64_t k, int64_t n, int NX,
                                   int NY) {
    const int i = (int)(k / NY), j = (int)(k % NY);
    T y = c[k] * x[k];
    y = y + c[n + k] * (i > 0 ? x[k - NY] : T(0));
    y = y + c[2 * n + k] * (i < NX - 1 ? x[k + NY] : T(0));
    y = y + c[3 * n + k] * (j < NY - 1 ? x[k + 1] : T(0));
    y = y + c[4 * n + k] * (j > 0 ? x[k - 1] : T(0));
    return y;
}

template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_fb_init(Flat fl, const T* __restrict__ rhs, T* __restrict__ x, T* __restrict__ r, T* __restrict__ rt,
          T* __restrict__ p, T* __restrict__ v, double eps, int max_iter, KState* __restrict__ st,
          double* __restrict__ P) {
    __shared__ double sh[4];
    const int64_t lo = (int64_t)blockIdx.x * fl.chunk;
    const int64_t hi = lo + fl.chunk < fl.n ? lo + fl.chunk : fl.n;
    double a = 0.0;
    for (int64_t k = lo + threadIdx.x; k < hi; k += KM_BLOCK) {
        const T b = rhs[k];
        x[k] = T(0);
        r[k] = b;
        rt[k] = b;
        p[k] = T(0);
        v[k] = T(0);
        a += (double)b * (double)b;
    }
    a = block_sum(a, sh);
    if (threadIdx.x == 0) {
        P[FB_PC + blockIdx.x] = a;
        P[FB_PC + KM_MAX_RED + blockIdx.x] = a;
        if (blockIdx.x == 0) {
            st->eps = eps;
            st->max_iter = max_iter;
            st->done = 0;
            st->status = KM_MAXITER;
            st->iter = 0;
            st->r0 = 0.0;
            st->rTr = 0.0;
            st->ref = 0.0;
            st->ring[KM_RING - 1].iter = -1;
            st->ring[KM_RING - 1].best = 1e300;
            st->ring[KM_RING - 1].best_it = 0;
        }
    }
}

// Stall: the recursive residual norm has not dropped below 0.9 x its best
// value for kStallWindow iterations.
constexpr int kStallWindow = 300;

template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_fb_p(Flat fl, int ph, const T* __restrict__ r, T* __restrict__ p, const T* __restrict__ v,
       KState* __restrict__ st, const double* __restrict__ P, double* __restrict__ Pw) {
    __shared__ double sh[8];
    if (st->done) return;
    const int prev = (ph + KM_RING - 1) % KM_RING;
    const KScal sp = st->ring[prev];
    const int i = sp.iter + 1;
    double tot[2];
    sum_partials<2>(P + FB_PC, fl.G, tot, sh);
    const double rho = tot[0], rr = tot[1];
    const double r0 = i == 0 ? rr : st->r0;
    int status = -1;
    if (!fin(rr)) status = KM_NONFINITE;
    else if (i == 0 && rr == 0.0) status = KM_OK;
    else if (sqrt(rr / r0) < st->eps) status = KM_OK;
    else if (i > 0 && sp.omega == 0.0) status = KM_BREAKDOWN;
    else if (i >= st->max_iter) status = KM_MAXITER;
    else if (rho == 0.0 || !fin(rho)) status = KM_BREAKDOWN;
    const bool better = rr < 0.81 * sp.best;
    if (status < 0 && !better && i - sp.best_it >= kStallWindow) status = KM_STALL;
    if (status >= 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            st->status = status;
            st->iter = i;
            st->rTr = rr;
            if (i == 0) st->r0 = rr;
            st->done = 1;
        }
        return;
    }
    const int64_t lo = (int64_t)blockIdx.x * fl.chunk;
    const int64_t hi = lo + fl.chunk < fl.n ? lo + fl.chunk : fl.n;
    double pp = 0.0;
    if (i == 0) {
        for (int64_t k = lo + threadIdx.x; k < hi; k += KM_BLOCK) {
            const T pk = r[k];
            p[k] = pk;
            pp += (double)pk * (double)pk;
        }
    } else {
        const double beta = (rho / sp.rho) * (sp.alpha / sp.omega);
        const T fb = (T)beta, fo = (T)sp.omega;
        for (int64_t k = lo + threadIdx.x; k < hi; k += KM_BLOCK) {
            const T pk = r[k] + fb * (p[k] - fo * v[k]);
            p[k] = pk;
            pp += (double)pk * (double)pk;
        }
    }
    pp = block_sum(pp, sh);
    if (threadIdx.x == 0) Pw[FB_PN + blockIdx.x] = pp;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->ring[ph].rho = rho;
        st->ring[ph].iter = i;
        st->ring[ph].best = better ? rr : sp.best;
        st->ring[ph].best_it = better ? i : sp.best_it;
        st->iter = i;
        st->rTr = rr;
        if (i == 0) st->r0 = rr;
    }
}

template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_fb_v(Flat fl, int NX, int NY, const T* __restrict__ c, const T* __restrict__ ph_, const T* __restrict__ rt,
       T* __restrict__ v, const KState* __restrict__ st, double* __restrict__ P) {
    __shared__ double sh[4];
    if (st->done) return;
    const int64_t lo = (int64_t)blockIdx.x * fl.chunk;
    const int64_t hi = lo + fl.chunk < fl.n ? lo + fl.chunk : fl.n;
    double a = 0.0;
    for (int64_t k = lo + threadIdx.x; k < hi; k += KM_BLOCK) {
        const T y = sten5<T>(c, ph_, k, fl.n, NX, NY);
        v[k] = y;
        a += (double)rt[k] * (double)y;
    }
    a = block_sum(a, sh);
    if (threadIdx.x == 0) P[FB_PV + blockIdx.x] = a;
}

template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_fb_s(Flat fl, int ph, const T* __restrict__ r, const T* __restrict__ v, T* __restrict__ s,
       KState* __restrict__ st, const double* __restrict__ P, double* __restrict__ Pw) {
    __shared__ double sh[8];
    if (st->done) return;
    double tot[1];
    sum_partials<1>(P + FB_PV, fl.G, tot, sh);
    const double rv = tot[0];
    if (rv == 0.0 || !fin(rv)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            st->status = KM_BREAKDOWN;
            st->iter = st->ring[ph].iter;
            st->done = 1;
        }
        return;
    }
    const double alpha = st->ring[ph].rho / rv;
    const T fa = (T)alpha;
    const int64_t lo = (int64_t)blockIdx.x * fl.chunk;
    const int64_t hi = lo + fl.chunk < fl.n ? lo + fl.chunk : fl.n;
    double ss = 0.0;
    for (int64_t k = lo + threadIdx.x; k < hi; k += KM_BLOCK) {
        const T sk = r[k] - fa * v[k];
        s[k] = sk;
        ss += (double)sk * (double)sk;
    }
    ss = block_sum(ss, sh);
    if (threadIdx.x == 0) {
        Pw[FB_PS + blockIdx.x] = ss;
        if (blockIdx.x == 0) st->ring[ph].alpha = alpha;
    }
}

template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_fb_t(Flat fl, int NX, int NY, const T* __restrict__ c, const T* __restrict__ sh_, const T* __restrict__ s,
       T* __restrict__ t, const KState* __restrict__ st, double* __restrict__ P) {
    __shared__ double sh[4];
    if (st->done) return;
    const int64_t lo = (int64_t)blockIdx.x * fl.chunk;
    const int64_t hi = lo + fl.chunk < fl.n ? lo + fl.chunk : fl.n;
    double a = 0.0, b = 0.0;
    for (int64_t k = lo + threadIdx.x; k < hi; k += KM_BLOCK) {
        const T y = sten5<T>(c, sh_, k, fl.n, NX, NY);
        t[k] = y;
        a += (double)y * (double)s[k];
        b += (double)y * (double)y;
    }
    a = block_sum(a, sh);
    b = block_sum(b, sh);
    if (threadIdx.x == 0) {
        P[FB_PT + blockIdx.x] = a;
        P[FB_PT + KM_MAX_RED + blockIdx.x] = b;
    }
}

template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_fb_x(Flat fl, int ph, const T* __restrict__ s, const T* __restrict__ t, const T* __restrict__ ph_,
       const T* __restrict__ sh_, const T* __restrict__ rt, T* __restrict__ x, T* __restrict__ r,
       KState* __restrict__ st, double* __restrict__ P) {
    __shared__ double sh[8];
    if (st->done) return;
    double tot[2];
    sum_partials<2>(P + FB_PT, fl.G, tot, sh);
    const double omega = tot[1] != 0.0 ? tot[0] / tot[1] : 0.0;
    const double alpha = st->ring[ph].alpha;
    const int64_t lo = (int64_t)blockIdx.x * fl.chunk;
    const int64_t hi = lo + fl.chunk < fl.n ? lo + fl.chunk : fl.n;
    double a = 0.0, b = 0.0;
    for (int64_t k = lo + threadIdx.x; k < hi; k += KM_BLOCK) {
        x[k] = x[k] + (T)(alpha * (double)ph_[k] + omega * (double)sh_[k]);
        const T rn = s[k] - (T)omega * t[k];
        r[k] = rn;
        a += (double)rt[k] * (double)rn;
        b += (double)rn * (double)rn;
    }
    a = block_sum(a, sh);
    b = block_sum(b, sh);
    if (threadIdx.x == 0) {
        P[FB_PC + blockIdx.x] = a;
        P[FB_PC + KM_MAX_RED + blockIdx.x] = b;
        if (blockIdx.x == 0) st->ring[ph].omega = omega;
    }
}

// z = D^-1 r (rows with a zero diagonal map to zero): the Jacobi hook of
// karman_taichi.py:399,419, used when no network is given.
template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_jacobi(int64_t n, const T* __restrict__ c, const T* __restrict__ r, T* __restrict__ z, const int* __restrict__ skip) {
    if (*skip) return;
    for (int64_t i = (int64_t)blockIdx.x * KM_BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * KM_BLOCK) {
        const T d = c[i];
        z[i] = d != T(0) ? r[i] / d : T(0);
    }
}

template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_scale_to(int64_t n, const double* __restrict__ a, T* __restrict__ out, double scale) {
    for (int64_t i = (int64_t)blockIdx.x * KM_BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * KM_BLOCK)
        out[i] = (T)(a[i] * scale);
}

template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_axpy_to(int64_t n, double* __restrict__ x, const T* __restrict__ d, double scale) {
    for (int64_t i = (int64_t)blockIdx.x * KM_BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * KM_BLOCK)
        x[i] = x[i] + scale * (double)d[i];
}

template <typename F>
__device__ __forceinline__ void flat_sweep_fg(const Flat& fl, F f) {
    const int64_t k0 = (int64_t)blockIdx.x * fl.chunk;
    const int64_t k1 = k0 + fl.chunk < fl.n ? k0 + fl.chunk : fl.n;
    for (int64_t k = k0 + threadIdx.x; k < k1; k += KM_BLOCK) f(k);
}

int blocks(int64_t n) {
    int64_t g = (n + KM_BLOCK - 1) / KM_BLOCK;
    if (g > 2048) g = 2048;
    return g < 1 ? 1 : (int)g;
}

const int kFbChunk = 8;

// Inner flexible solve of A d = rhs with vectors of type T, d returned in xd.
// coefT: the coefficients in T (the matvecs); c32: FP32 coefficients (the
// network's features); coef64: the FP64 system (coarse level).  ws holds 10n
// values of T.  The preconditioner: the network (params) or Jacobi, after
// the coarse correction when coarse is set.  Without a network the coarse
// level (and KM_FLAG_SMOOTH) run as the fused iterations of km_twolevel.hip.
template <typename T>
int fb_inner(km_ctx* ctx, int NX, int NY, const T* coefT, const float* c32, const double* coef64, const T* rhs, T* xd,
             T* ws, double eps, int max_iter, bool use_graph, const float* params, bool coarse, KState* out,
             bool mg = false) {
    const int64_t n = (int64_t)NX * NY;
    T* r = ws;
    T* rt = ws + n;
    T* p = ws + 2 * n;
    T* v = ws + 3 * n;
    T* s = ws + 4 * n;
    T* t = ws + 5 * n;
    T* ph = ws + 6 * n;
    T* sh = ws + 7 * n;
    T* r1 = ws + 8 * n;      // residual after the coarse correction
    KState* st = &ctx->st[4];
    double* P = ctx->partials;
    const Flat fl = make_flat(n);
    hipLaunchKernelGGL(k_fb_init<T>, dim3(fl.G), dim3(KM_BLOCK), 0, ctx->stream, fl, rhs, xd, r, rt, p, v, eps,
                       max_iter, st, P);
    KM_CHECK_LAUNCH(ctx);
    const int* skip = &st->done;
    const int jb = blocks(n);
    auto precond = [=](hipStream_t q, const T* src, T* dst, const double* norm_partials) {
        if (coarse && params) {
            (void)kmcoarse::apply<T>(ctx, coef64, coefT, src, dst, r1, P + FB_PR, skip, q);
            (void)kmnet::apply_pre<T>(ctx, NX, NY, c32, r1, dst, skip, q, params, P + FB_PR, fl.G, true);
        } else if (params) {
            (void)kmnet::apply_pre<T>(ctx, NX, NY, c32, src, dst, skip, q, params, norm_partials, fl.G, false);
        } else if (mg) {
            // FP64 vectors only (fbicg_solve rejects KM_FLAG_MG with KM_FLAG_MIXED)
            (void)kmmg::vcycle(ctx, reinterpret_cast<const double*>(src), reinterpret_cast<double*>(dst), q, skip);
        } else {
            hipLaunchKernelGGL(k_jacobi<T>, dim3(jb), dim3(KM_BLOCK), 0, q, n, coefT, src, dst, skip);
        }
    };
    auto enqueue = [=](hipStream_t q, int phs) {
        hipLaunchKernelGGL(k_fb_p<T>, dim3(fl.G), dim3(KM_BLOCK), 0, q, fl, phs, (const T*)r, p, (const T*)v, st,
                           (const double*)P, P);
        precond(q, p, ph, P + FB_PN);
        hipLaunchKernelGGL(k_fb_v<T>, dim3(fl.G), dim3(KM_BLOCK), 0, q, fl, NX, NY, coefT, (const T*)ph, (const T*)rt,
                           v, (const KState*)st, P);
        hipLaunchKernelGGL(k_fb_s<T>, dim3(fl.G), dim3(KM_BLOCK), 0, q, fl, phs, (const T*)r, (const T*)v, s, st,
                           (const double*)P, P);
        precond(q, s, sh, P + FB_PS);
        hipLaunchKernelGGL(k_fb_t<T>, dim3(fl.G), dim3(KM_BLOCK), 0, q, fl, NX, NY, coefT, (const T*)sh, (const T*)s, t,
                           (const KState*)st, P);
        hipLaunchKernelGGL(k_fb_x<T>, dim3(fl.G), dim3(KM_BLOCK), 0, q, fl, phs, (const T*)s, (const T*)t, (const T*)ph,
                           (const T*)sh, (const T*)rt, xd, r, st, P);
    };
    GraphKey key;
    key.v = {7, sizeof(T), (uintptr_t)NX, (uintptr_t)NY, (uintptr_t)coefT, (uintptr_t)c32, (uintptr_t)rhs,
             (uintptr_t)xd, (uintptr_t)ws, (uintptr_t)params, (uintptr_t)ctx->net, (uintptr_t)st,
             (uintptr_t)coarse, (uintptr_t)ctx->coarse, (uintptr_t)mg,
             (uintptr_t)(mg && ctx->mg ? kmmg::arena(ctx) : nullptr)};
    return run_chunks(ctx, st, key, use_graph, kFbChunk, max_iter, enqueue, out);
}

int residual64(km_ctx* ctx, int NX, int NY, const double* coef, const double* b, const double* x, double* r,
               double* rr) {
    const int rc = km_residual(ctx, NX, NY, coef, b, x, r, rr);
    return rc < 0 ? rc : KM_OK;
}

// The first FP64 sweep of a fused solve aims at kInnerMargin eps in its
// recursive residual, so that the true residual, which decides, is below
// eps without a second sweep.
constexpr double kInnerMargin = 0.75;

// Workspace bytes of fbicg_solve<T> on n cells.
template <typename T>
size_t fbicg_bytes(int64_t n) {
    const bool f64 = sizeof(T) == sizeof(double);
    // FP64 r (n); FP32 coefficients (5n); T: coefficients unless FP64 (5n), rhs (n), d (n), inner (8n).
    // At least what the plain fallback needs, so the arena is not regrown mid-solve.
    size_t bytes = (size_t)n * sizeof(double) + (size_t)5 * n * sizeof(float) +
                   (size_t)(f64 ? 12 : 17) * n * sizeof(T) + 2048;
    if (bytes < (size_t)9 * n * sizeof(double)) bytes = (size_t)9 * n * sizeof(double);
    return bytes;
}

// Outer FP64 refinement around inner flexible solves in T.  FP64 inner
// vectors aim each sweep at the requested tolerance; FP32 inner vectors run
// sweeps to at most 1e-5 each (what FP32 recurrences reliably reach, at most
// 500 iterations).  A sweep also ends when its recursive residual stalls.
// pre (may be null): the state and true <r,r> of a first sweep already run
// into x (the direct case below, by the pair solve); the sweep is then not
// run again.
template <typename T>
int fbicg_solve(km_ctx* ctx, int NX, int NY, const double* coef, const double* b, double* x, double eps,
                int max_iter, int flags, const float* params, km_solve_info* info, const KState* pre = nullptr,
                double pre_rr = 0.0) {
    const int64_t n = (int64_t)NX * NY;
    const bool use_graph = !(flags & KM_FLAG_NO_GRAPH);
    const bool f64 = sizeof(T) == sizeof(double);
    if (int rc = km_ensure_workspace(ctx, fbicg_bytes<T>(n))) return rc;
    double* r64 = static_cast<double*>(km_ws_vectors(ctx));
    float* c32 = reinterpret_cast<float*>(r64 + n);
    // T vectors start at the next 256-byte boundary after the FP32 coefficients
    T* base = reinterpret_cast<T*>((reinterpret_cast<uintptr_t>(c32 + 5 * n) + 255) & ~uintptr_t(255));
    const T* coefT;
    T* rhs;
    if (f64) {
        coefT = reinterpret_cast<const T*>(coef);
        rhs = base;
    } else {
        coefT = reinterpret_cast<const T*>(c32);
        rhs = base;
    }
    T* d = rhs + n;
    T* wT = d + n;
    // FP32 coefficients: the network's features and the FP32 vectors' matvecs
    if (params || !f64) {
        hipLaunchKernelGGL(k_scale_to<float>, dim3(blocks(5 * n)), dim3(KM_BLOCK), 0, ctx->stream, 5 * n, coef, c32,
                           1.0);
        KM_CHECK_LAUNCH(ctx);
    }
    const bool coarse = (flags & KM_FLAG_COARSE) != 0;
    const bool smooth = (flags & KM_FLAG_SMOOTH) != 0;
    const bool mg = (flags & KM_FLAG_MG) != 0;
    const bool sym = (flags & KM_FLAG_SYMMETRIC) != 0;
    KM_REQUIRE(ctx, !(smooth && (coarse || params)), "KM_FLAG_SMOOTH takes neither a coarse level nor a network");
    KM_REQUIRE(ctx, !(mg && (coarse || smooth || params || !f64)),
               "KM_FLAG_MG takes FP64 vectors and no other preconditioner");
    if (mg)
        if (int rc = kmmg::setup(ctx, NX, NY, coef, ctx->stream)) return rc;
    if (params) {
        if (int rc = kmnet::prepare(ctx, 1, NX, NY, params)) return rc;
        if (int rc = kmnet::solve_stats(ctx, NX, NY, c32, ctx->stream)) return rc;
    }
    // a first sweep run by a group solve (pre) already set the level up and
    // left it in ctx->coarse
    if (coarse && !pre)
        if (int rc = kmcoarse::setup(ctx, NX, NY, coef, ctx->stream, true)) return rc;
    double bb = 0.0, bnorm = 0.0, rr = 0.0, rel = 0.0;
    int total = 0, sweeps = 0, status = KM_MAXITER;
    // FP64 vectors from x = 0 with a fused inner solve: the first sweep solves
    // A x = b straight into x (x = 0 is its start); b.b is the solve's own
    // first <r,r>, so no round trip to the host precedes it
    const bool direct = f64 && !params && !(flags & KM_FLAG_X0) && (coarse || smooth);
    KM_REQUIRE(ctx, direct || !pre, "a first sweep was run for a solve that takes none");
    if (direct) {
        KState st;
        double rr_first = pre_rr;
        if (pre) {
            st = *pre;
        } else {
            T* xT = reinterpret_cast<T*>(x);
            const T* bT = reinterpret_cast<const T*>(b);
            // the true residual is enqueued behind the solve and read with its state
            const std::function<int()> tail = [&]() { return km_residual_async(ctx, NX, NY, coef, b, x, r64); };
            const int rc = smooth ? kmcoarse::smooth_inner<T>(ctx, NX, NY, coefT, bT, xT, wT, kInnerMargin * eps, max_iter,
                                                              use_graph, &st, &tail)
                                  : kmcoarse::fused_inner<T>(ctx, NX, NY, coefT, bT, xT, wT, kInnerMargin * eps, max_iter,
                                                             use_graph, &st, &tail, sym, coef);
            if (rc) return rc;
            rr_first = km_mapped_res(ctx, 0);
        }
        bb = st.r0;
        bnorm = sqrt(bb);
        if (info) {
            info->bnorm = bnorm;
            info->refinements = 0;
            info->fallback = 0;
        }
        if (!(bb > 0.0) && bb == bb) {   // b = 0: x stays 0
            if (info) {
                info->iterations = 0;
                info->status = KM_OK;
                info->rel_resid = 0.0;
                info->rel_resid_true = 0.0;
            }
            return KM_OK;
        }
        total = st.iter;
        sweeps = 1;
        rr = rr_first;
        rel = sqrt(rr) / bnorm;
        if (!(rel < 0.5)) {
            // the sweep stalled or diverged: back to x = 0, plain BiCGSTAB finishes
            KM_HIP(ctx, hipMemsetAsync(x, 0, n * sizeof(double), ctx->stream));
            status = KM_STALL;
            rel = 1.0;
        } else if (pre && !(rel < eps)) {
            // more sweeps follow: they start from r = b - A x in r64
            if (int rc = residual64(ctx, NX, NY, coef, b, x, r64, &rr)) return rc;
            rel = sqrt(rr) / bnorm;
        }
    } else {
        if (int rc = km_dot(ctx, n, KM_F64, b, b, &bb)) return rc < 0 ? rc : KM_EHIP;
        bnorm = sqrt(bb);
        if (info) {
            info->bnorm = bnorm;
            info->refinements = 0;
            info->fallback = 0;
        }
        if (!(flags & KM_FLAG_X0)) KM_HIP(ctx, hipMemsetAsync(x, 0, n * sizeof(double), ctx->stream));
        if (bb == 0.0) {
            KM_HIP(ctx, hipMemsetAsync(x, 0, n * sizeof(double), ctx->stream));
            KM_HIP_SYNC(ctx, hipStreamSynchronize(ctx->stream));
            if (info) {
                info->iterations = 0;
                info->status = KM_OK;
                info->rel_resid = 0.0;
                info->rel_resid_true = 0.0;
            }
            return KM_OK;
        }
        if (flags & KM_FLAG_X0) {
            if (int rc = residual64(ctx, NX, NY, coef, b, x, r64, &rr)) return rc;
        } else {
            // x = 0: r = b and r.r = b.b, no second round trip to the host
            KM_HIP(ctx, hipMemcpyAsync(r64, b, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
            rr = bb;
        }
        rel = sqrt(rr) / bnorm;
    }
    // warm start x += ||r|| M(r / ||r||), kept only if it lowers the residual
    if (params && rel >= eps && rel == rel) {
        const double rn = sqrt(rr);
        hipLaunchKernelGGL(k_scale_to<T>, dim3(blocks(n)), dim3(KM_BLOCK), 0, ctx->stream, n, (const double*)r64, rhs,
                           1.0 / rn);
        KM_CHECK_LAUNCH(ctx);
        if (coarse) {
            T* r1w = wT + 8 * n;
            double* Pw = ctx->partials + FB_PR;
            if (int rc = kmcoarse::apply<T>(ctx, coef, coefT, rhs, d, r1w, Pw, nullptr, ctx->stream)) return rc;
            if (int rc = kmnet::apply_pre<T>(ctx, NX, NY, c32, r1w, d, nullptr, ctx->stream, params, Pw,
                                             make_flat(n).G, true))
                return rc;
        } else if (int rc = kmnet::apply<T>(ctx, NX, NY, c32, rhs, d, nullptr, ctx->stream, params)) {
            return rc;
        }
        hipLaunchKernelGGL(k_axpy_to<T>, dim3(blocks(n)), dim3(KM_BLOCK), 0, ctx->stream, n, x, (const T*)d, rn);
        KM_CHECK_LAUNCH(ctx);
        double rr2 = 0.0;
        if (int rc = residual64(ctx, NX, NY, coef, b, x, r64, &rr2)) return rc;
        if (rr2 < rr) {
            rr = rr2;
        } else {
            hipLaunchKernelGGL(k_axpy_to<T>, dim3(blocks(n)), dim3(KM_BLOCK), 0, ctx->stream, n, x, (const T*)d, -rn);
            KM_CHECK_LAUNCH(ctx);
            if (int rc = residual64(ctx, NX, NY, coef, b, x, r64, &rr)) return rc;
        }
        rel = sqrt(rr) / bnorm;
    }
    const double inner_floor = f64 ? 0.0 : 1e-5;
    const int inner_cap = f64 ? max_iter : 500;
    while (status != KM_STALL) {
        if (!(rel == rel)) { status = KM_NONFINITE; break; }
        if (rel < eps) { status = KM_OK; break; }
        if (total >= max_iter || sweeps >= 64) { status = KM_MAXITER; break; }
        const double rn = sqrt(rr);
        hipLaunchKernelGGL(k_scale_to<T>, dim3(blocks(n)), dim3(KM_BLOCK), 0, ctx->stream, n, (const double*)r64, rhs,
                           1.0 / rn);
        KM_CHECK_LAUNCH(ctx);
        // FP64: aim a little below eps so rounding between the recursive and
        // the true residual does not cost a second sweep
        double inner_eps = f64 ? kInnerMargin * eps / rel : eps / rel;
        if (inner_eps < inner_floor) inner_eps = inner_floor;
        if (inner_eps > 1e-1) inner_eps = 1e-1;
        int cap = max_iter - total;
        if (cap > inner_cap) cap = inner_cap;
        KState st;
        // the fused inner solves take x += rn d and the FP64 true residual as
        // their tail, enqueued behind the last iterations and read with the
        // state: one host round trip per sweep instead of two
        const bool tailed = smooth || (coarse && !params);
        const std::function<int()> tail = [&, rn]() {
            hipLaunchKernelGGL(k_axpy_to<T>, dim3(blocks(n)), dim3(KM_BLOCK), 0, ctx->stream, n, x, (const T*)d, rn);
            return km_residual_async(ctx, NX, NY, coef, b, x, r64, 0);
        };
        const int rc =
            smooth ? kmcoarse::smooth_inner<T>(ctx, NX, NY, coefT, rhs, d, wT, inner_eps, cap, use_graph, &st, &tail)
            : coarse && !params
                ? kmcoarse::fused_inner<T>(ctx, NX, NY, coefT, rhs, d, wT, inner_eps, cap, use_graph, &st, &tail,
                                           sym, coef)
                : fb_inner<T>(ctx, NX, NY, coefT, c32, coef, rhs, d, wT, inner_eps, cap, use_graph, params, coarse,
                              &st, mg);
        if (rc) return rc;
        total += st.iter;
        sweeps++;
        const double prev = rel;
        if (tailed) {
            rr = km_mapped_res(ctx, 0);
        } else {
            hipLaunchKernelGGL(k_axpy_to<T>, dim3(blocks(n)), dim3(KM_BLOCK), 0, ctx->stream, n, x, (const T*)d, rn);
            KM_CHECK_LAUNCH(ctx);
            if (int rc = residual64(ctx, NX, NY, coef, b, x, r64, &rr)) return rc;
        }
        rel = sqrt(rr) / bnorm;
        if (!(rel < 0.5 * prev)) {
            // the preconditioned sweep stalled or diverged: take it back
            if (rel == rel && rel < 1e300) {
                hipLaunchKernelGGL(k_axpy_to<T>, dim3(blocks(n)), dim3(KM_BLOCK), 0, ctx->stream, n, x, (const T*)d,
                                   -rn);
                KM_CHECK_LAUNCH(ctx);
            } else {
                KM_HIP(ctx, hipMemsetAsync(x, 0, n * sizeof(double), ctx->stream));
            }
            if (int rc = residual64(ctx, NX, NY, coef, b, x, r64, &rr)) return rc;
            rel = sqrt(rr) / bnorm;
            status = KM_STALL;
            break;
        }
    }
    if (status == KM_STALL || status == KM_NONFINITE || (status == KM_MAXITER && total < max_iter)) {
        km_solve_info sub;
        const int left = max_iter > total ? max_iter - total : 1;
        // FP32 vectors with a classical preconditioner: the same solve with
        // FP64 vectors from x (a few dozen iterations where the FP32 sweep
        // stalled), which itself falls back to plain BiCGSTAB; otherwise
        // plain FP64 BiCGSTAB from x
        const int rc = !f64 && !params && (coarse || smooth)
                           ? fbicg_solve<double>(ctx, NX, NY, coef, b, x, eps, left,
                                                 (flags & ~KM_FLAG_MIXED) | KM_FLAG_X0, nullptr, &sub)
                           : km_bicgstab(ctx, NX, NY, coef, b, x, eps, left, (flags & ~KM_FLAG_MIXED) | KM_FLAG_X0,
                                         &sub);
        if (rc < 0) return rc;
        total += sub.iterations;
        sweeps += sub.refinements;
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

// ------------------------------------------------------------ flexible GCR
//
// The outer method for a preconditioner that changes from one application
// to the next (KM_FLAG_GCR): flexible GCR, restarted every kFgMax
// directions.  Iteration k: z = M(r / ||r||), q = A z; q and z orthogonalised
// against the stored q_j twice (classical Gram-Schmidt), scaled to ||q|| = 1
// and stored; alpha = <r, q>, x += alpha z, r -= alpha q.  ||r|| never
// increases, whatever M does, so a non-linear network cannot break the
// recurrences as it breaks flexible BiCGSTAB's.  The residual-space vectors
// (q and the stored q_j) are in T (FP32 with KM_FLAG_MIXED); the
// solution-space ones (z, the stored z_j), x, r, A z and every dot product in
// FP64: A amplifies the rounding noise of an FP32 z by up to ~1e5 on smooth
// directions, which would leave q = A z and z inconsistent.  At each restart, and when the recursive residual reaches eps, r
// is replaced by the true residual b - A x (FP64), which alone decides.  A
// restart cycle that does not lower the true residual by 10% ends the solve
// (stall), and plain BiCGSTAB finishes from x.
constexpr int kFgMax = 16;
constexpr int FG_PD = 16 * KM_MAX_RED;   // [kFgMax][KM_MAX_RED] partials of <q, q_j>
constexpr int FG_PN = 32 * KM_MAX_RED;   // [2] <q, q>, <r, q>; [1] <r, r> at +2
static_assert(FG_PN + 3 * KM_MAX_RED <= KM_MAX_PARTIALS, "FGCR partials");
// device scalars in ctx->dres: beta[kFgMax], then <q,q>, <r,q>, <r,r>
constexpr int FG_SB = 0, FG_SN = kFgMax, FG_SR = kFgMax + 2;

// q = A z with the FP64 coefficients and FP64 products and sums, rounded to
// T once: on a smooth z (the coarse correction is piecewise constant) the
// five terms of a p' row cancel to a small fraction of aP z, which FP32
// arithmetic would leave as rounding noise.
template <typename T>
__device__ __forceinline__ double sten5_64(const double* __restrict__ c, const T* __restrict__ x, int64_t k, int64_t n,
                                           int NX, int NY) {
    const int i = (int)(k / NY), j = (int)(k % NY);
    double y = c[k] * (double)x[k];
    y = y + c[n + k] * (i > 0 ? (double)x[k - NY] : 0.0);
    y = y + c[2 * n + k] * (i < NX - 1 ? (double)x[k + NY] : 0.0);
    y = y + c[3 * n + k] * (j < NY - 1 ? (double)x[k + 1] : 0.0);
    y = y + c[4 * n + k] * (j > 0 ? (double)x[k - 1] : 0.0);
    return y;
}

template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_fg_stencil(int64_t n, int NX, int NY, const double* __restrict__ c, const double* __restrict__ z,
             T* __restrict__ q) {
    for (int64_t k = (int64_t)blockIdx.x * KM_BLOCK + threadIdx.x; k < n; k += (int64_t)gridDim.x * KM_BLOCK)
        q[k] = (T)sten5_64(c, z, k, n, NX, NY);
}

// zd = z (widened to FP64)
template <typename T>
__global__ void __launch_bounds__(KM_BLOCK) k_fg_widen(int64_t n, const T* __restrict__ z, double* __restrict__ zd) {
    for (int64_t k = (int64_t)blockIdx.x * KM_BLOCK + threadIdx.x; k < n; k += (int64_t)gridDim.x * KM_BLOCK)
        zd[k] = (double)z[k];
}

// r1 = r - A z1 in FP64 arithmetic (see k_fg_stencil), with the partials of
// r1.r1 in the flat geometry fl (the network's input scale)
template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_fg_resid1(Flat fl, int NX, int NY, const double* __restrict__ c, const T* __restrict__ r, const T* __restrict__ z1,
            T* __restrict__ r1, double* __restrict__ P) {
    __shared__ double sh[4];
    double a = 0.0;
    flat_sweep_fg(fl, [&](int64_t k) {
        const double v = (double)r[k] - sten5_64(c, z1, k, fl.n, NX, NY);
        const T vt = (T)v;
        r1[k] = vt;
        a += (double)vt * (double)vt;
    });
    a = block_sum(a, sh);
    if (threadIdx.x == 0 && P) P[blockIdx.x] = a;
}

// zd = z1 + D^-1 r1 in FP64 (a Jacobi step after the coarse correction; aP = 0 rows: z1)
template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_fg_jacobi_add(int64_t n, const double* __restrict__ c, const T* __restrict__ z1, const T* __restrict__ r1,
                double* __restrict__ zd) {
    for (int64_t k = (int64_t)blockIdx.x * KM_BLOCK + threadIdx.x; k < n; k += (int64_t)gridDim.x * KM_BLOCK)
        zd[k] = (double)z1[k] + (c[k] != 0.0 ? (double)r1[k] / c[k] : 0.0);
}

// partials of <q, Q_j>, j < m
template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_fg_dots(Flat fl, int m, const T* __restrict__ q, const T* __restrict__ Q, double* __restrict__ P) {
    __shared__ double sh[4];
    double acc[kFgMax];
#pragma unroll
    for (int j = 0; j < kFgMax; ++j) acc[j] = 0.0;
    flat_sweep_fg(fl, [&](int64_t k) {
        const double qk = (double)q[k];
#pragma unroll
        for (int j = 0; j < kFgMax; ++j)
            if (j < m) acc[j] += qk * (double)Q[(int64_t)j * fl.n + k];
    });
    for (int j = 0; j < m; ++j) {
        const double v = block_sum(acc[j], sh);
        if (threadIdx.x == 0) P[(int64_t)j * KM_MAX_RED + blockIdx.x] = v;
        __syncthreads();
    }
}

// out[j] = the sum of the partials of plane j (one workgroup per plane, fixed order)
__global__ void __launch_bounds__(KM_BLOCK)
k_fg_reduce(int G, const double* __restrict__ P, double* __restrict__ out) {
    __shared__ double sh[4];
    double t[1];
    sum_partials<1>(P + (int64_t)blockIdx.x * KM_MAX_RED, G, t, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = t[0];
}

// q -= sum_j beta_j Q_j, z -= sum_j beta_j Z_j (FP64 arithmetic)
template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_fg_sub(Flat fl, int m, const double* __restrict__ beta, T* __restrict__ q, double* __restrict__ z,
         const T* __restrict__ Q, const double* __restrict__ Z) {
    __shared__ double bs[kFgMax];
    if (threadIdx.x < kFgMax) bs[threadIdx.x] = threadIdx.x < (unsigned)m ? beta[threadIdx.x] : 0.0;
    __syncthreads();
    flat_sweep_fg(fl, [&](int64_t k) {
        double qk = (double)q[k], zk = z[k];
        for (int j = 0; j < m; ++j) {
            const double bj = bs[j];
            qk = qk - bj * (double)Q[(int64_t)j * fl.n + k];
            zk = zk - bj * Z[(int64_t)j * fl.n + k];
        }
        q[k] = (T)qk;
        z[k] = zk;
    });
}

// partials of <q, q> and <r, q>
template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_fg_norms(Flat fl, const T* __restrict__ q, const double* __restrict__ r, double* __restrict__ P) {
    __shared__ double sh[4];
    double a = 0.0, b = 0.0;
    flat_sweep_fg(fl, [&](int64_t k) {
        const double qk = (double)q[k];
        a += qk * qk;
        b += r[k] * qk;
    });
    a = block_sum(a, sh);
    __syncthreads();
    b = block_sum(b, sh);
    if (threadIdx.x == 0) {
        P[blockIdx.x] = a;
        P[KM_MAX_RED + blockIdx.x] = b;
    }
}

// Q_k = q / ||q||, Z_k = z / ||q||; x += alpha Z_k, r -= alpha Q_k with
// alpha = <r, q> / ||q||; partials of <r, r>
template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_fg_step(Flat fl, const double* __restrict__ s, const T* __restrict__ q, const double* __restrict__ z,
          T* __restrict__ Qk, double* __restrict__ Zk, double* __restrict__ x, double* __restrict__ r,
          double* __restrict__ P) {
    __shared__ double sh[4];
    const double nq = sqrt(s[0]);
    const double inv = nq > 0.0 ? 1.0 / nq : 0.0;
    const double alpha = s[1] * inv;
    double a = 0.0;
    flat_sweep_fg(fl, [&](int64_t k) {
        const T qn = (T)((double)q[k] * inv);
        const double zn = z[k] * inv;
        Qk[k] = qn;
        Zk[k] = zn;
        x[k] = x[k] + alpha * zn;
        const double rk = r[k] - alpha * (double)qn;
        r[k] = rk;
        a += rk * rk;
    });
    a = block_sum(a, sh);
    if (threadIdx.x == 0) P[blockIdx.x] = a;
}

enum FgPrecond { FG_NET, FG_TWOLEVEL, FG_COARSE_JACOBI, FG_JACOBI, FG_MG };

template <typename T>
int fgcr_solve(km_ctx* ctx, int NX, int NY, const double* coef, const double* b, double* x, double eps,
               int max_iter, int flags, const float* params, km_solve_info* info) {
    const int64_t n = (int64_t)NX * NY;
    const bool f64 = sizeof(T) == sizeof(double);
    const bool coarse = (flags & KM_FLAG_COARSE) != 0;
    KM_REQUIRE(ctx, !(flags & KM_FLAG_SMOOTH), "KM_FLAG_GCR takes no KM_FLAG_SMOOTH");
    const bool mg = (flags & KM_FLAG_MG) != 0;
    KM_REQUIRE(ctx, !(mg && (coarse || params || !f64)), "KM_FLAG_MG takes FP64 vectors and no other preconditioner");
    const FgPrecond kind = params ? FG_NET
                           : mg   ? FG_MG
                                  : (coarse ? ((flags & KM_FLAG_JACOBI) ? FG_COARSE_JACOBI : FG_TWOLEVEL) : FG_JACOBI);
    // r64 | zd, z2d, Z[kFgMax] (FP64) | c32 | rhs, z, q, r1, z2 | Q[kFgMax] (T)
    const size_t bytes = (size_t)(3 + kFgMax) * n * sizeof(double) + (size_t)5 * n * sizeof(float) + 512 +
                         (size_t)(5 + kFgMax) * n * sizeof(T);
    if (int rc = km_ensure_workspace(ctx, bytes)) return rc;
    double* r64 = static_cast<double*>(km_ws_vectors(ctx));
    double* zd = r64 + n;
    double* z2d = zd + n;
    double* Z = z2d + n;
    float* c32 = reinterpret_cast<float*>(Z + (size_t)kFgMax * n);
    T* base = reinterpret_cast<T*>((reinterpret_cast<uintptr_t>(c32 + 5 * n) + 255) & ~uintptr_t(255));
    T* rhs = base;
    T* z = rhs + n;
    T* q = z + n;
    T* r1 = q + n;
    T* z2 = r1 + n;
    T* Q = z2 + n;
    if (params || !f64) {
        hipLaunchKernelGGL(k_scale_to<float>, dim3(blocks(5 * n)), dim3(KM_BLOCK), 0, ctx->stream, 5 * n, coef, c32,
                           1.0);
        KM_CHECK_LAUNCH(ctx);
    }
    const T* coefT = f64 ? reinterpret_cast<const T*>(coef) : reinterpret_cast<const T*>(c32);
    if (params) {
        if (int rc = kmnet::prepare(ctx, 1, NX, NY, params)) return rc;
        if (int rc = kmnet::solve_stats(ctx, NX, NY, c32, ctx->stream)) return rc;
    }
    if (coarse)
        if (int rc = kmcoarse::setup(ctx, NX, NY, coef, ctx->stream, true)) return rc;
    if (mg)
        if (int rc = kmmg::setup(ctx, NX, NY, coef, ctx->stream)) return rc;
    double bb = 0.0;
    if (int rc = km_dot(ctx, n, KM_F64, b, b, &bb)) return rc < 0 ? rc : KM_EHIP;
    if (info) {
        info->bnorm = sqrt(bb);
        info->refinements = 0;
        info->fallback = 0;
    }
    if (!(flags & KM_FLAG_X0)) KM_HIP(ctx, hipMemsetAsync(x, 0, n * sizeof(double), ctx->stream));
    if (bb == 0.0) {
        KM_HIP(ctx, hipMemsetAsync(x, 0, n * sizeof(double), ctx->stream));
        KM_HIP_SYNC(ctx, hipStreamSynchronize(ctx->stream));
        if (info) {
            info->iterations = 0;
            info->status = KM_OK;
            info->rel_resid = info->rel_resid_true = 0.0;
        }
        return KM_OK;
    }
    const double bn = sqrt(bb);
    auto true_residual = [&](double* rr) -> int {
        if (int rc = km_residual_async(ctx, NX, NY, coef, b, x, r64, 1)) return rc;
        KM_HIP_SYNC(ctx, hipStreamSynchronize(ctx->stream));
        *rr = km_mapped_res(ctx, 1);
        return KM_OK;
    };
    double rr = 0.0;
    if (flags & KM_FLAG_X0) {
        if (int rc = true_residual(&rr)) return rc;
    } else {
        KM_HIP(ctx, hipMemcpyAsync(r64, b, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        rr = bb;
    }
    const Flat fl = make_flat(n);
    double* P = ctx->partials;
    double* S = ctx->dres;
    const int jb = blocks(n);
    int status = KM_MAXITER, total = 0, k = 0, restarts = 0;
    double rel = sqrt(rr) / bn, cycle_start = rel;
    while (true) {
        if (!(rel == rel)) { status = KM_NONFINITE; break; }
        if (total >= max_iter) { status = KM_MAXITER; break; }
        // the directions of this application of M to r / ||r||: one, or with
        // the network after the coarse correction two -- the two-level M(r)
        // (SOR sweeps from z1) and z2 = z1 + net(r1) -- so that GCR minimises
        // over both and the network's direction can only add to the
        // classical one
        hipLaunchKernelGGL(k_scale_to<T>, dim3(jb), dim3(KM_BLOCK), 0, ctx->stream, n, (const double*)r64, rhs,
                           1.0 / sqrt(rr));
        KM_CHECK_LAUNCH(ctx);
        int ndir = 1;
        if (kind == FG_NET && coarse) {
            if (int rc = kmcoarse::apply<T>(ctx, coef, coefT, rhs, z, r1, nullptr, nullptr, ctx->stream)) return rc;
            hipLaunchKernelGGL(k_fg_resid1<T>, dim3(fl.G), dim3(KM_BLOCK), 0, ctx->stream, fl, NX, NY, coef,
                               (const T*)rhs, (const T*)z, r1, P + FB_PR);
            KM_HIP(ctx, hipMemcpyAsync(z2, z, n * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream));
            if (int rc = kmnet::apply_pre<T>(ctx, NX, NY, c32, r1, z2, nullptr, ctx->stream, params, P + FB_PR, fl.G,
                                             true))
                return rc;
            if (flags & KM_FLAG_JACOBI) {
                // one direction: z1 + net(r1), then a Jacobi step on what is left
                hipLaunchKernelGGL(k_fg_resid1<T>, dim3(fl.G), dim3(KM_BLOCK), 0, ctx->stream, fl, NX, NY, coef,
                                   (const T*)rhs, (const T*)z2, r1, (double*)nullptr);
                hipLaunchKernelGGL(k_fg_jacobi_add<T>, dim3(jb), dim3(KM_BLOCK), 0, ctx->stream, n, coef,
                                   (const T*)z2, (const T*)r1, zd);
            } else {
                // two directions: the two-level M(r) (red-black SOR sweeps
                // from z1, yc still holding r's coarse solution) and z1 + net(r1)
                hipLaunchKernelGGL(k_fg_widen<T>, dim3(jb), dim3(KM_BLOCK), 0, ctx->stream, n, (const T*)z2, z2d);
                if (int rc = kmcoarse::smooth<T>(ctx, coefT, rhs, z, ctx->stream)) return rc;
                hipLaunchKernelGGL(k_fg_widen<T>, dim3(jb), dim3(KM_BLOCK), 0, ctx->stream, n, (const T*)z, zd);
                ndir = 2;
            }
        } else if (kind == FG_NET) {
            if (int rc = kmnet::apply<T>(ctx, NX, NY, c32, rhs, z, nullptr, ctx->stream, params)) return rc;
            hipLaunchKernelGGL(k_fg_widen<T>, dim3(jb), dim3(KM_BLOCK), 0, ctx->stream, n, (const T*)z, zd);
        } else if (kind == FG_TWOLEVEL) {
            if (int rc = kmcoarse::restrict_solve<T>(ctx, rhs, nullptr, ctx->stream)) return rc;
            if (int rc = kmcoarse::smooth<T>(ctx, coefT, rhs, z, ctx->stream)) return rc;
            hipLaunchKernelGGL(k_fg_widen<T>, dim3(jb), dim3(KM_BLOCK), 0, ctx->stream, n, (const T*)z, zd);
        } else if (kind == FG_MG) {
            // FP64 vectors (T = double): one V-cycle straight into the direction
            if (int rc = kmmg::vcycle(ctx, reinterpret_cast<const double*>(rhs), zd, ctx->stream, nullptr)) return rc;
        } else if (kind == FG_COARSE_JACOBI) {
            if (int rc = kmcoarse::apply<T>(ctx, coef, coefT, rhs, z, r1, nullptr, nullptr, ctx->stream)) return rc;
            hipLaunchKernelGGL(k_fg_resid1<T>, dim3(fl.G), dim3(KM_BLOCK), 0, ctx->stream, fl, NX, NY, coef,
                               (const T*)rhs, (const T*)z, r1, (double*)nullptr);
            hipLaunchKernelGGL(k_fg_jacobi_add<T>, dim3(jb), dim3(KM_BLOCK), 0, ctx->stream, n, coef, (const T*)z,
                               (const T*)r1, zd);
        } else {
            hipLaunchKernelGGL(k_jacobi<T>, dim3(jb), dim3(KM_BLOCK), 0, ctx->stream, n, coefT, (const T*)rhs, z,
                               (const int*)nullptr);
            hipLaunchKernelGGL(k_fg_widen<T>, dim3(jb), dim3(KM_BLOCK), 0, ctx->stream, n, (const T*)z, zd);
        }
        bool stop = false;
        for (int d = 0; d < ndir && !stop; ++d) {
            double* zdir = d == 0 ? zd : z2d;
            // q = A zd, orthogonalised twice against the stored directions
            hipLaunchKernelGGL(k_fg_stencil<T>, dim3(jb), dim3(KM_BLOCK), 0, ctx->stream, n, NX, NY, coef,
                               (const double*)zdir, q);
            for (int pass = 0; pass < 2 && k > 0; ++pass) {
                hipLaunchKernelGGL(k_fg_dots<T>, dim3(fl.G), dim3(KM_BLOCK), 0, ctx->stream, fl, k, (const T*)q,
                                   (const T*)Q, P + FG_PD);
                hipLaunchKernelGGL(k_fg_reduce, dim3(k), dim3(KM_BLOCK), 0, ctx->stream, fl.G,
                                   (const double*)(P + FG_PD), S + FG_SB);
                hipLaunchKernelGGL(k_fg_sub<T>, dim3(fl.G), dim3(KM_BLOCK), 0, ctx->stream, fl, k,
                                   (const double*)(S + FG_SB), q, zdir, (const T*)Q, (const double*)Z);
            }
            hipLaunchKernelGGL(k_fg_norms<T>, dim3(fl.G), dim3(KM_BLOCK), 0, ctx->stream, fl, (const T*)q,
                               (const double*)r64, P + FG_PN);
            hipLaunchKernelGGL(k_fg_reduce, dim3(2), dim3(KM_BLOCK), 0, ctx->stream, fl.G, (const double*)(P + FG_PN),
                               S + FG_SN);
            hipLaunchKernelGGL(k_fg_step<T>, dim3(fl.G), dim3(KM_BLOCK), 0, ctx->stream, fl,
                               (const double*)(S + FG_SN), (const T*)q, (const double*)zdir, Q + (size_t)k * n,
                               Z + (size_t)k * n, x, r64, P + FG_PN + 2 * KM_MAX_RED);
            hipLaunchKernelGGL(k_fg_reduce, dim3(1), dim3(KM_BLOCK), 0, ctx->stream, fl.G,
                               (const double*)(P + FG_PN + 2 * KM_MAX_RED), S + FG_SR);
            KM_CHECK_LAUNCH(ctx);
            KM_HIP(ctx, hipMemcpyAsync(ctx->h_res + 40, S + FG_SN, 3 * sizeof(double), hipMemcpyDeviceToHost,
                                       ctx->stream));
            KM_HIP_SYNC(ctx, hipStreamSynchronize(ctx->stream));
            const double qq = ctx->h_res[40];
            if (!(qq > 0.0) || !(qq == qq)) {
                // no new direction (zd lies in the stored ones); another direction may still add one
                if (!(qq == qq)) {
                    status = KM_NONFINITE;
                    stop = true;
                } else if (d + 1 == ndir) {
                    status = KM_BREAKDOWN;
                    stop = true;
                }
                continue;
            }
            rr = ctx->h_res[42];
            total++;
            k++;
            rel = sqrt(rr) / bn;
            if (rel < eps || k == kFgMax) {
                if (int rc = true_residual(&rr)) return rc;
                rel = sqrt(rr) / bn;
                if (rel < eps) {
                    status = KM_OK;
                    stop = true;
                } else if (k == kFgMax) {
                    if (!(rel < 0.9 * cycle_start)) {
                        status = KM_STALL;
                        stop = true;
                    }
                    cycle_start = rel;
                    k = 0;
                    restarts++;
                }
            }
        }
        if (stop) break;
    }
    if (status != KM_OK) {
        km_solve_info sub;
        const int rc = km_bicgstab(ctx, NX, NY, coef, b, x, eps, max_iter > total ? max_iter - total : 1,
                                   KM_FLAG_X0 | KM_FLAG_TRUE, &sub);
        if (rc < 0) return rc;
        total += sub.iterations;
        status = sub.status;
        rel = sub.rel_resid_true;
        if (info) info->fallback = 1;
    }
    if (info) {
        info->iterations = total;
        info->status = status;
        info->refinements = restarts;
        info->rel_resid = rel;
        info->rel_resid_true = rel;
    }
    return KM_OK;
}

// n symmetric FP64 two-level solves from x = 0 (the p' systems of several
// cases): each system's coarse level set up and its smoothing tile chosen as
// its own solve would, then the systems of one tile shape, in groups of up to
// kmcoarse::kGroupSystems, run their first sweeps in shared launches
// (fused_inner_n), and each system's own continuation (further sweeps, the
// stall fallback) runs as in fbicg_solve.  mi: each system's iteration cap.
int fbicg_group(km_ctx* ctx, int n, const int* NX, const int* NY, const double* const* coef, const double* const* b,
                double* const* x, double eps, const int* mi, int flags, km_solve_info* info) {
    auto up = [](int64_t v) { return (size_t)((v * (int64_t)sizeof(double) + 255) & ~(int64_t)255); };
    // the first kGroupSystems systems' FP64 true residuals and 8n inner
    // workspaces, reused by each group in turn (a group takes at most that many)
    std::vector<int64_t> cells(n);
    for (int k = 0; k < n; ++k) cells[k] = (int64_t)NX[k] * NY[k];
    std::vector<int64_t> big(cells);
    std::sort(big.begin(), big.end(), std::greater<int64_t>());
    size_t bytes = 0;
    for (int k = 0; k < n && k < kmcoarse::kGroupSystems; ++k) bytes += up(big[k]) + up(8 * big[k]);
    for (int k = 0; k < n; ++k) bytes = std::max(bytes, fbicg_bytes<double>(cells[k]));
    if (int rc = km_ensure_workspace(ctx, bytes)) return rc;
    char* V = static_cast<char*>(km_ws_vectors(ctx));
    // every level set up (in system order, as separate solves would) and its tile chosen
    std::vector<km_coarse*> C(n);
    std::vector<int> tile(n);
    for (int k = 0; k < n; ++k) {
        if (int rc = kmcoarse::setup(ctx, NX[k], NY[k], coef[k], ctx->stream, true)) return rc;
        C[k] = ctx->coarse;
        double* scratch = reinterpret_cast<double*>(V + up(cells[k]));
        if (int rc = kmcoarse::group_tile(ctx, C[k], coef[k], b[k], scratch, &tile[k])) return rc;
    }
    const bool use_graph = !(flags & KM_FLAG_NO_GRAPH);
    std::vector<char> taken(n, 0);
    for (int a = 0; a < n; ++a) {
        if (taken[a]) continue;
        int g[kmcoarse::kGroupSystems];
        int m = 0;
        for (int c = a; c < n && m < kmcoarse::kGroupSystems; ++c)
            if (!taken[c] && tile[c] == tile[a]) {
                taken[c] = 1;
                g[m++] = c;
            }
        km_coarse* gC[kmcoarse::kGroupSystems];
        int gNX[kmcoarse::kGroupSystems], gNY[kmcoarse::kGroupSystems], gmi[kmcoarse::kGroupSystems];
        const double* gcoef[kmcoarse::kGroupSystems];
        const double* gb[kmcoarse::kGroupSystems];
        double* gx[kmcoarse::kGroupSystems];
        double* gr[kmcoarse::kGroupSystems];
        double* gw[kmcoarse::kGroupSystems];
        char* q = V;
        for (int j = 0; j < m; ++j) {
            const int k = g[j];
            gC[j] = C[k];
            gNX[j] = NX[k];
            gNY[j] = NY[k];
            gmi[j] = mi[k];
            gcoef[j] = coef[k];
            gb[j] = b[k];
            gx[j] = x[k];
            gr[j] = reinterpret_cast<double*>(q);
            q += up(cells[k]);
            gw[j] = reinterpret_cast<double*>(q);
            q += up(8 * cells[k]);
        }
        KState st[kmcoarse::kGroupSystems];
        // the true residuals are enqueued behind the solves and read with their states
        const std::function<int()> tail = [&]() {
            for (int j = 0; j < m; ++j)
                if (int rc = km_residual_async(ctx, gNX[j], gNY[j], gcoef[j], gb[j], gx[j], gr[j], j)) return rc;
            return (int)KM_OK;
        };
        if (int rc = kmcoarse::fused_inner_n(ctx, m, gC, gNX, gNY, gcoef, gb, gx, gw, kInnerMargin * eps, gmi, use_graph, st,
                                             &tail, tile[a]))
            return rc;
        double rr[kmcoarse::kGroupSystems];
        for (int j = 0; j < m; ++j) rr[j] = km_mapped_res(ctx, j);   // a continuation may reuse the slots
        for (int j = 0; j < m; ++j) {
            const int k = g[j];
            ctx->coarse = C[k];
            if (int rc = fbicg_solve<double>(ctx, NX[k], NY[k], coef[k], b[k], x[k], eps, mi[k], flags, nullptr,
                                             info ? &info[k] : nullptr, &st[j], rr[j]))
                return rc;
        }
    }
    return KM_OK;
}

}  // namespace

extern "C" {

int km_fbicgstab_n(km_ctx* ctx, int n, const int* NX, const int* NY, const double* const* coef,
                   const double* const* b, double* const* x, double eps, int max_iter, int flags,
                   km_solve_info* info) {
    if (!ctx) return KM_EINVAL;
    KM_REQUIRE(ctx, n >= 0 && n <= KM_MAX_GROUP_SOLVES, "%d systems (at most %d)", n, KM_MAX_GROUP_SOLVES);
    if (n == 0) return KM_OK;
    KM_REQUIRE(ctx, NX && NY && coef && b && x, "null pointer");
    KM_REQUIRE(ctx, eps > 0.0, "eps must be positive");
    std::vector<int> mi(n);
    for (int k = 0; k < n; ++k) {
        KM_REQUIRE(ctx, coef[k] && b[k] && x[k], "system %d: null pointer", k);
        KM_REQUIRE(ctx, NX[k] >= 4 && NY[k] >= 4 && (int64_t)NX[k] * NY[k] < ((int64_t)1 << 31),
                   "system %d shape %d x %d", k, NX[k], NY[k]);
        for (int j = 0; j < k; ++j)
            KM_REQUIRE(ctx, x[j] != x[k] && coef[j] != coef[k], "systems %d and %d share x or coefficients", j, k);
        const int64_t cells = (int64_t)NX[k] * NY[k];
        mi[k] = max_iter > 0 ? max_iter : (int)(cells < (1 << 30) ? cells : (1 << 30));
    }
    const bool group = (flags & KM_FLAG_COARSE) && (flags & KM_FLAG_SYMMETRIC) &&
                       !(flags & (KM_FLAG_X0 | KM_FLAG_MIXED | KM_FLAG_SMOOTH | KM_FLAG_GCR | KM_FLAG_JACOBI |
                                  KM_FLAG_MG));
    if (!group || n == 1) {
        for (int k = 0; k < n; ++k)
            if (int rc = km_fbicgstab(ctx, NX[k], NY[k], coef[k], b[k], x[k], eps, max_iter, flags, nullptr,
                                      info ? &info[k] : nullptr))
                return rc;
        return KM_OK;
    }
    return fbicg_group(ctx, n, NX, NY, coef, b, x, eps, mi.data(), flags, info);
}

int km_fbicgstab(km_ctx* ctx, int NX, int NY, const double* coef, const double* b, double* x, double eps,
                 int max_iter, int flags, const float* params, km_solve_info* info) {
    if (!ctx) return KM_EINVAL;
    KM_REQUIRE(ctx, coef && b && x, "null pointer");
    KM_REQUIRE(ctx, NX >= 4 && NY >= 4 && (int64_t)NX * NY < ((int64_t)1 << 31), "system shape %d x %d", NX, NY);
    KM_REQUIRE(ctx, eps > 0.0, "eps must be positive");
    const int64_t n = (int64_t)NX * NY;
    if (max_iter <= 0) max_iter = (int)(n < (1 << 30) ? n : (1 << 30));
    if (flags & KM_FLAG_GCR)
        return (flags & KM_FLAG_MIXED) ? fgcr_solve<float>(ctx, NX, NY, coef, b, x, eps, max_iter, flags, params, info)
                                       : fgcr_solve<double>(ctx, NX, NY, coef, b, x, eps, max_iter, flags, params, info);
    if (flags & KM_FLAG_MIXED) return fbicg_solve<float>(ctx, NX, NY, coef, b, x, eps, max_iter, flags, params, info);
    return fbicg_solve<double>(ctx, NX, NY, coef, b, x, eps, max_iter, flags, params, info);
}

int km_fbicgstab2(km_ctx* ctx, int NX0, int NY0, const double* coef0, const double* b0, double* x0, int NX1, int NY1,
                  const double* coef1, const double* b1, double* x1, double eps, int max_iter, int flags,
                  km_solve_info* info0, km_solve_info* info1) {
    if (!ctx) return KM_EINVAL;
    KM_REQUIRE(ctx, coef0 && b0 && x0 && coef1 && b1 && x1, "null pointer");
    KM_REQUIRE(ctx, x0 != x1, "the two systems share their solution vector");
    const int NX[2] = {NX0, NX1}, NY[2] = {NY0, NY1};
    int64_t n[2];
    int mi[2];
    for (int q = 0; q < 2; ++q) {
        KM_REQUIRE(ctx, NX[q] >= 4 && NY[q] >= 4 && (int64_t)NX[q] * NY[q] < ((int64_t)1 << 31),
                   "system %d shape %d x %d", q, NX[q], NY[q]);
        n[q] = (int64_t)NX[q] * NY[q];
        mi[q] = max_iter > 0 ? max_iter : (int)(n[q] < (1 << 30) ? n[q] : (1 << 30));
    }
    KM_REQUIRE(ctx, eps > 0.0, "eps must be positive");
    const bool pair = (flags & KM_FLAG_SMOOTH) && !(flags & (KM_FLAG_MIXED | KM_FLAG_X0 | KM_FLAG_COARSE));
    if (!pair) {
        if (int rc = km_fbicgstab(ctx, NX0, NY0, coef0, b0, x0, eps, max_iter, flags, nullptr, info0)) return rc;
        return km_fbicgstab(ctx, NX1, NY1, coef1, b1, x1, eps, max_iter, flags, nullptr, info1);
    }
    // FP64 true residuals r (n0, n1), then the two inner workspaces (8n each)
    auto up = [](int64_t v) { return (size_t)((v * (int64_t)sizeof(double) + 255) & ~(int64_t)255); };
    size_t bytes = up(n[0]) + up(n[1]) + up(8 * n[0]) + up(8 * n[1]);
    bytes = std::max(bytes, std::max(fbicg_bytes<double>(n[0]), fbicg_bytes<double>(n[1])));
    if (int rc = km_ensure_workspace(ctx, bytes)) return rc;
    char* V = static_cast<char*>(km_ws_vectors(ctx));
    double* r[2];
    double* w[2];
    r[0] = reinterpret_cast<double*>(V);
    r[1] = reinterpret_cast<double*>(V + up(n[0]));
    w[0] = reinterpret_cast<double*>(V + up(n[0]) + up(n[1]));
    w[1] = reinterpret_cast<double*>(V + up(n[0]) + up(n[1]) + up(8 * n[0]));
    const double* coef[2] = {coef0, coef1};
    const double* b[2] = {b0, b1};
    double* x[2] = {x0, x1};
    KState st[2];
    // both true residuals are enqueued behind the solve and read with its states
    const std::function<int()> tail = [&]() {
        for (int q = 0; q < 2; ++q)
            if (int rc = km_residual_async(ctx, NX[q], NY[q], coef[q], b[q], x[q], r[q], q)) return rc;
        return (int)KM_OK;
    };
    if (int rc = kmcoarse::smooth_inner2<double>(ctx, NX, NY, coef, b, x, w, kInnerMargin * eps, mi,
                                                 !(flags & KM_FLAG_NO_GRAPH), st, &tail))
        return rc;
    const double rr1 = km_mapped_res(ctx, 1);   // the first system's continuation may reuse the slots
    if (int rc = fbicg_solve<double>(ctx, NX0, NY0, coef0, b0, x0, eps, mi[0], flags, nullptr, info0, &st[0],
                                     km_mapped_res(ctx, 0)))
        return rc;
    return fbicg_solve<double>(ctx, NX1, NY1, coef1, b1, x1, eps, mi[1], flags, nullptr, info1, &st[1], rr1);
}

}  // extern "C"
