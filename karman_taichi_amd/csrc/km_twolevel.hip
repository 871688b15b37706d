// Fused flexible BiCGSTAB with the two-level Jacobi preconditioner
//   M(r) = z1, then kSweeps damped Jacobi sweeps on r - A z,  z1 = P A_c^-1 P^T r
// (the aggregation coarse level of km_coarse.hip), the preconditioner of the
// p^ / s^ hooks of bicgsolver.py:48-58.
//
// Six launches per inner iteration i (ring slot ph = i mod KM_RING):
//   K_P  S workgroups per aggregate (row slices).  omega = <t,s> / <t,t>,
//        rho = <s,r~> - omega <t,r~>, beta = (rho / rho') (alpha / omega);
//        x += alpha p^ + omega s^, r = s - omega t (iteration i-1),
//        p = r + beta (p - omega v); rc = P^T p; partial <r,r>
//   K_M  yc = A_c^-1 rc
//   K_V  2-D tiles with a (kSweeps + 1)-cell halo: p^ = M(p) (z1 from yc,
//        kSweeps sweeps on shrinking halos in LDS), v = A p^; partial <r~,v>
//   K_S  S workgroups per aggregate.  Stopping test on <r,r>;
//        alpha = rho / <r~,v>; s = r - alpha v; rc = P^T s
//   K_M  yc = A_c^-1 rc
//   K_T  tiles: s^ = M(s), t = A s^; partials <t,s>, <t,t>, <s,r~>, <t,r~>
// The unfused sequence (restrict, matvec, prolongation, the sweeps and the
// Krylov updates as separate kernels) takes fifteen launches; every vector of
// the fused form is written once per iteration.
#include "km_coarse.h"

#include <cmath>
#include <vector>

namespace {

constexpr int kMaxAggWg = 8192;            // aggregate-slice workgroups
constexpr int kMaxTileWg = 8192;           // smoothing workgroups (one tile each up to 8192 tiles)
constexpr int TL_PR = 0;                   // [kMaxAggWg] <r, r> (aggregate-slice workgroups)
constexpr int TL_PV = kMaxAggWg;           // [kMaxTileWg] <r~, v> (tile workgroups)
constexpr int TL_PT = kMaxAggWg + kMaxTileWg;   // [4][kMaxTileWg] <t, s>, <t, t>, <s, r~>, <t, r~>
constexpr int TL_TOT = kMaxAggWg + 5 * kMaxTileWg;   // totals: 0 <r~,v>, 1..4 the K_T planes, 5 <r,r>
// Consumers sum a producer's partials in every workgroup; when partials x
// consumer workgroups exceeds this, one workgroup sums them first (k_tl_total)
// and the consumers read the totals.
constexpr int64_t kTotalWork = 2000000;
constexpr int FIN_V = 1, FIN_T = 2, FIN_R = 4;
constexpr int kStallWindow = 300;        // as the unfused flexible solve
constexpr int kChunk = 8;
// Tile shapes of the smoothing launches (rows x columns, columns contiguous);
// the fastest for a system is chosen on its first solve.
constexpr int kNumTiles = 6;
constexpr int kTileI[kNumTiles] = {32, 16, 16, 8, 8, 16};
constexpr int kTileJ[kNumTiles] = {32, 64, 32, 64, 128, 32};
// minimum waves per SIMD the compiler must allow (a VGPR budget): 16 x 32 also
// fits 128 VGPRs without spills, so a fourth wave per SIMD can hide latency
constexpr int kTileW[kNumTiles] = {1, 1, 1, 1, 1, 4};

__device__ __forceinline__ bool fin(double v) { return v == v && v - v == 0.0; }

__device__ __forceinline__ void finish_state(KState* st, int status, int iter, double rr) {
    st->status = status;
    st->iter = iter;
    st->rTr = rr;
    st->done = 1;
}

// Geometry of the aggregate launches: aggregate (I, J) covers the interior
// block [1 + I B, 1 + (I + 1) B) x [1 + J B, 1 + (J + 1) B), clipped; edge
// aggregates also own the adjacent halo rows and columns, so every cell of the
// NX x NY system belongs to exactly one aggregate.  Each aggregate's rows are
// cut into S slices, one workgroup each (G = m S workgroups).
struct AggGeom {
    int NX, NY, B, mI, mJ, m;
    int A;      // aggregates per workgroup, adjacent along J (small aggregates)
    int ngrp;   // groups of A per aggregate row
    int S;      // row slices per group (large aggregates)
    int G;      // workgroups: mI ngrp S
};
constexpr int kMaxA = 8;

// Cells of one workgroup: rows [ra, ra + cnt / ej) x columns [je0, je0 + ej)
// of aggregates (I, J0 .. J0 + nA - 1); cell t is k(t), and loc(t) is the
// index (0 .. nA - 1) of its aggregate within the group.
struct Slice {
    int ra, je0, ej, cnt, NY, B, mJ, I, J0, nA, sl;
    __device__ __forceinline__ int64_t k(int t) const { return (int64_t)(ra + t / ej) * NY + (je0 + t % ej); }
    __device__ __forceinline__ int loc(int t) const {
        const int j = je0 + t % ej;
        int J = j < 1 ? 0 : (j - 1) / B;
        if (J > mJ - 1) J = mJ - 1;
        return J - J0;
    }
};

__device__ __forceinline__ Slice slice_of(const AggGeom& g, int wg) {
    const int sl = wg % g.S, rest = wg / g.S;
    const int grp = rest % g.ngrp, I = rest / g.ngrp;
    const int J0 = grp * g.A, J1 = (J0 + g.A < g.mJ ? J0 + g.A : g.mJ) - 1;
    const int nx = g.NX - 2, ny = g.NY - 2;
    const int i0 = 1 + I * g.B;
    const int ie0 = I == 0 ? 0 : i0, ie1 = I == g.mI - 1 ? g.NX - 1 : (i0 + g.B - 1 < nx ? i0 + g.B - 1 : nx);
    const int je0 = J0 == 0 ? 0 : 1 + J0 * g.B;
    const int je1 = J1 == g.mJ - 1 ? g.NY - 1 : (J1 * g.B + g.B < ny ? J1 * g.B + g.B : ny);
    const int er = ie1 - ie0 + 1;
    Slice q;
    q.ej = je1 - je0 + 1;
    q.je0 = je0;
    q.ra = ie0 + (er * sl) / g.S;
    q.cnt = (ie0 + (er * (sl + 1)) / g.S - q.ra) * q.ej;
    q.NY = g.NY;
    q.B = g.B;
    q.mJ = g.mJ;
    q.I = I;
    q.J0 = J0;
    q.nA = J1 - J0 + 1;
    q.sl = sl;
    return q;
}

// The group's restriction sums: each thread's per-aggregate sums, reduced over
// the workgroup in a fixed order (waves, then the four waves) and written as
// slice partials rc[sl m + agg].
__device__ __forceinline__ void write_restriction(const Slice& q, int m, const double (&ac)[kMaxA],
                                                  double* __restrict__ rc) {
    __shared__ double shr[4 * kMaxA];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (q.nA == 1) {   // one aggregate: one sum
        const double v = wave_sum(ac[0]);
        if (lane == 0) shr[w] = v;
    } else {
#pragma unroll
        for (int u = 0; u < kMaxA; ++u) {
            const double v = wave_sum(ac[u]);
            if (lane == 0) shr[u * 4 + w] = v;
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < q.nA) {
        const int u = threadIdx.x;
        rc[(int64_t)q.sl * m + (int64_t)q.I * q.mJ + q.J0 + u] =
            (shr[u * 4] + shr[u * 4 + 1]) + (shr[u * 4 + 2] + shr[u * 4 + 3]);
    }
}

__device__ __forceinline__ void add_to(double (&ac)[kMaxA], int a, double v) {
#pragma unroll
    for (int u = 0; u < kMaxA; ++u) ac[u] += u == a ? v : 0.0;
}

// Cells per thread whose loads are issued before the scalars are reduced
// (slices hold about kStage workgroups' worth; larger ones loop over the rest).
constexpr int kStage = 4;

template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_tl_init(Flat fl, const T* __restrict__ rhs, T* __restrict__ x, T* __restrict__ r, T* __restrict__ rt, double eps,
          int max_iter, KState* __restrict__ st) {
    const int64_t lo = (int64_t)blockIdx.x * fl.chunk;
    const int64_t hi = lo + fl.chunk < fl.n ? lo + fl.chunk : fl.n;
    for (int64_t k = lo + threadIdx.x; k < hi; k += KM_BLOCK) {
        const T b = rhs[k];
        x[k] = T(0);
        r[k] = b;
        rt[k] = b;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
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

// K_P: the x / r update of iteration i-1, the p update of iteration i, P^T p
template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_tl_p(AggGeom g, int Gt, int tot, int ph, const uint8_t* __restrict__ mask, T* __restrict__ x, T* __restrict__ r,
       T* __restrict__ p, const T* __restrict__ v, const T* __restrict__ s, const T* __restrict__ t,
       const T* __restrict__ ph_, const T* __restrict__ sh_, double* __restrict__ rc, KState* __restrict__ st,
       double* __restrict__ P) {
    __shared__ double sh[8];
    if (st->done) return;
    const KScal sp = st->ring[(ph + KM_RING - 1) % KM_RING];
    const int i = sp.iter + 1;
    // each XCD takes a contiguous run of aggregates: the cache lines that
    // neighbouring aggregates share are fetched into one L2
    const int wg = xcd_remap(blockIdx.x, g.G);
    const Slice sq = slice_of(g, wg);
    int64_t kk[kStage];
    int la[kStage];
    bool ok[kStage], fr[kStage];
    T xv[kStage], hv[kStage], s2v[kStage], sv[kStage], tv[kStage], pv[kStage], vv[kStage];
#pragma unroll
    for (int c = 0; c < kStage; ++c) {
        const int tt = threadIdx.x + c * KM_BLOCK;
        ok[c] = tt < sq.cnt;
        kk[c] = ok[c] ? sq.k(tt) : 0;
        la[c] = ok[c] ? sq.loc(tt) : 0;
        const int64_t k = kk[c];
        fr[c] = ok[c] && mask[k] != 0;
        const bool ld = ok[c] && i > 0;
        sv[c] = ok[c] ? (i == 0 ? r[k] : s[k]) : T(0);
        xv[c] = ld ? x[k] : T(0);
        hv[c] = ld ? ph_[k] : T(0);
        s2v[c] = ld ? sh_[k] : T(0);
        tv[c] = ld ? t[k] : T(0);
        pv[c] = ld ? p[k] : T(0);
        vv[c] = ld ? v[k] : T(0);
    }
    double omega = 0.0, beta = 0.0;
    if (i > 0) {
        double pt[4];
        if (tot & FIN_T) {
#pragma unroll
            for (int q = 0; q < 4; ++q) pt[q] = P[TL_TOT + 1 + q];
        } else {
            sum_partials_wide<4>(P + TL_PT, Gt, pt, kMaxTileWg);
        }
        omega = pt[1] != 0.0 ? pt[0] / pt[1] : 0.0;                  // bicgsolver.py:118
        const double rho = pt[2] - omega * pt[3];                    // <r_i, r~> (bicgsolver.py:95)
        beta = omega != 0.0 ? (rho / sp.rho) * (sp.alpha / omega) : 0.0;   // bicgsolver.py:102
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            st->ring[ph].rho = rho;
            st->ring[ph].omega = omega;
            st->ring[ph].beta = beta;
            st->ring[ph].iter = i;
        }
    } else if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->ring[ph].iter = 0;
        st->ring[ph].omega = 1.0;
    }
    const double alpha = sp.alpha;
    const T fo = (T)omega, fb = (T)beta;
    double rr = 0.0;
    double ac[kMaxA];
#pragma unroll
    for (int u = 0; u < kMaxA; ++u) ac[u] = 0.0;
    auto upd = [&](int64_t k, bool free, int a, T xk, T hk, T sk2, T sk, T tk, T pk0, T vk) {
        T rn, pk;
        if (i == 0) {
            rn = sk;                       // r_0 (staged in the s slot)
            pk = rn;
        } else {
            x[k] = xk + (T)(alpha * (double)hk + omega * (double)sk2);
            rn = sk - fo * tk;
            r[k] = rn;
            pk = rn + fb * (pk0 - fo * vk);
        }
        p[k] = pk;
        rr += (double)rn * (double)rn;
        if (free) add_to(ac, a, (double)pk);
    };
#pragma unroll
    for (int c = 0; c < kStage; ++c)
        if (ok[c]) upd(kk[c], fr[c], la[c], xv[c], hv[c], s2v[c], sv[c], tv[c], pv[c], vv[c]);
    for (int tt = threadIdx.x + kStage * KM_BLOCK; tt < sq.cnt; tt += KM_BLOCK) {
        const int64_t k = sq.k(tt);
        if (i == 0)
            upd(k, mask[k] != 0, sq.loc(tt), T(0), T(0), T(0), r[k], T(0), T(0), T(0));
        else
            upd(k, mask[k] != 0, sq.loc(tt), x[k], ph_[k], sh_[k], s[k], t[k], p[k], v[k]);
    }
    write_restriction(sq, g.m, ac, rc);
    rr = block_sum(rr, sh);
    if (threadIdx.x == 0) P[TL_PR + wg] = rr;
}

// K_S: stopping test, alpha, s = r - alpha v, P^T s
template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_tl_s(AggGeom g, int Gt, int Gp, int tot, int ph, const uint8_t* __restrict__ mask, const T* __restrict__ r,
       const T* __restrict__ v, T* __restrict__ s, double* __restrict__ rc, KState* __restrict__ st,
       double* __restrict__ P) {
    __shared__ double sh[8];
    if (st->done) return;
    const int wg = xcd_remap(blockIdx.x, g.G);
    const Slice sq = slice_of(g, wg);
    int64_t kk[kStage];
    int la[kStage];
    bool ok[kStage], fr[kStage];
    T rv_[kStage], vv[kStage];
#pragma unroll
    for (int c = 0; c < kStage; ++c) {
        const int tt = threadIdx.x + c * KM_BLOCK;
        ok[c] = tt < sq.cnt;
        kk[c] = ok[c] ? sq.k(tt) : 0;
        la[c] = ok[c] ? sq.loc(tt) : 0;
        fr[c] = ok[c] && mask[kk[c]] != 0;
        rv_[c] = ok[c] ? r[kk[c]] : T(0);
        vv[c] = ok[c] ? v[kk[c]] : T(0);
    }
    double tv[1], tr[1];
    if (tot & FIN_V) tv[0] = P[TL_TOT];
    else sum_partials_wide<1>(P + TL_PV, Gt, tv);
    if (tot & FIN_R) tr[0] = P[TL_TOT + 5];
    else sum_partials_wide<1>(P + TL_PR, Gp, tr);   // K_P's workgroups
    const double rv = tv[0], rr = tr[0];
    const KScal sp = st->ring[ph];
    const KScal spp = st->ring[(ph + KM_RING - 1) % KM_RING];
    const int i = sp.iter;
    const double rho = i == 0 ? rr : sp.rho;
    const double r0 = i == 0 ? rr : st->r0;
    const bool lead0 = blockIdx.x == 0 && threadIdx.x == 0;
    int status = -1;
    if (!fin(rr)) status = KM_NONFINITE;
    else if (i == 0 && rr == 0.0) status = KM_OK;
    else if (i > 0 && sqrt(rr / r0) < st->eps) status = KM_OK;   // bicgsolver.py:130
    else if (i > 0 && sp.omega == 0.0) status = KM_BREAKDOWN;
    else if (i >= st->max_iter) status = KM_MAXITER;
    else if (!fin(rho)) status = KM_NONFINITE;
    else if (rho == 0.0) status = KM_BREAKDOWN;
    else if (!fin(rv)) status = KM_NONFINITE;
    else if (rv == 0.0) status = KM_BREAKDOWN;
    const bool better = rr < 0.81 * spp.best;
    if (status < 0 && !better && i - spp.best_it >= kStallWindow) status = KM_STALL;
    if (lead0 && i == 0) st->r0 = r0;
    if (status >= 0) {
        if (lead0) finish_state(st, status, i, rr);
        return;
    }
    const double alpha = rho / rv;   // bicgsolver.py:108
    if (lead0) {
        st->ring[ph].alpha = alpha;
        if (i == 0) st->ring[ph].rho = rho;
        st->ring[ph].best = better ? rr : spp.best;
        st->ring[ph].best_it = better ? i : spp.best_it;
        st->rTr = rr;
        st->iter = i;
    }
    const T fa = (T)alpha;
    double ac[kMaxA];
#pragma unroll
    for (int u = 0; u < kMaxA; ++u) ac[u] = 0.0;
    auto upd = [&](int64_t k, bool free, int a, T rk, T vk) {
        const T sk = rk - fa * vk;
        s[k] = sk;
        if (free) add_to(ac, a, (double)sk);
    };
#pragma unroll
    for (int c = 0; c < kStage; ++c)
        if (ok[c]) upd(kk[c], fr[c], la[c], rv_[c], vv[c]);
    for (int tt = threadIdx.x + kStage * KM_BLOCK; tt < sq.cnt; tt += KM_BLOCK) {
        const int64_t k = sq.k(tt);
        upd(k, mask[k] != 0, sq.loc(tt), r[k], v[k]);
    }
    write_restriction(sq, g.m, ac, rc);
}

// K_V / K_T / the standalone smoother: z = M(src) on the tiles, then (MODE 1)
// v = A z with <r~, v> or (MODE 2) t = A z with <t,s>, <t,t>, <s,r~>, <t,r~>.
// The coefficients and src of the tile's kSweeps-cell halo frame stay in
// registers (Q cells per thread); z lives in LDS on the (kSweeps + 1)-cell
// halo frame, ping-ponged between two buffers over the sweeps.  The arithmetic is that of one
// damped Jacobi sweep per cell: z + (omega / aP) (src - A z), neighbours
// outside the system read as zero, rows with aP = 0 left unchanged.
template <typename T, int TI, int TJ, int MODE, int WPE>
__global__ void __launch_bounds__(KM_BLOCK) __attribute__((amdgpu_waves_per_eu(WPE, 8)))
k_tl_smooth(Tile tl, int B, int mJ, const T* __restrict__ c, const T* __restrict__ src,
            const uint8_t* __restrict__ mask, const double* __restrict__ yc, const T* __restrict__ rt,
            T* __restrict__ z, T* __restrict__ y, const KState* __restrict__ st, double* __restrict__ P) {
    constexpr int NS = kmcoarse::kSweeps;
    constexpr int WZ = TJ + 2 * (NS + 1), HZ = TI + 2 * (NS + 1);   // z frame: halo NS + 1
    constexpr int WR = TJ + 2 * NS, HR = TI + 2 * NS;               // register frame: halo NS
    constexpr int NR = WR * HR;
    constexpr int Q = (NR + KM_BLOCK - 1) / KM_BLOCK;
    __shared__ T za[HZ * WZ];
    __shared__ T zb[HZ * WZ];
    __shared__ double sh[8];
    if (st && st->done) return;
    const int L = xcd_remap(blockIdx.x, tl.G);
    const int NX = tl.NX, NY = tl.NY;
    const int64_t n = (int64_t)NX * NY;
    const T om = (T)kmcoarse::kOmega;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    const int t0 = L * tl.tpb, t1 = t0 + tl.tpb < tl.T ? t0 + tl.tpb : tl.T;
    for (int tile = t0; tile < t1; ++tile) {
        const int ti0 = (tile / tl.tiles_j) * TI, tj0 = (tile % tl.tiles_j) * TJ;
        T cr[Q][5], sr[Q], rq[Q], wd[Q];
        int64_t kq[Q];
        bool in[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int idx = threadIdx.x + q * KM_BLOCK;
            const int a = idx / WR, b = idx - (idx / WR) * WR;
            const int i = ti0 - NS + a, j = tj0 - NS + b;
            in[q] = idx < NR && i >= 0 && i < NX && j >= 0 && j < NY;
            kq[q] = in[q] ? (int64_t)i * NY + j : 0;
#pragma unroll
            for (int e = 0; e < 5; ++e) cr[q][e] = in[q] ? c[e * n + kq[q]] : T(0);
            sr[q] = in[q] ? src[kq[q]] : T(0);
            wd[q] = cr[q][0] != T(0) ? om * (T(1) / cr[q][0]) : T(0);   // omega D^-1
            const bool centre = in[q] && a >= NS && a < TI + NS && b >= NS && b < TJ + NS;
            rq[q] = MODE > 0 && centre ? rt[kq[q]] : T(0);
        }
        // z1 = P yc on the z frame (zero off the free cells and outside the
        // system); every mask and coarse value load is issued before the first
        // one is used
        constexpr int QZ = (HZ * WZ + KM_BLOCK - 1) / KM_BLOCK;
        const int nx = NX - 2, ny = NY - 2;
        double yq[QZ];
        uint8_t mq[QZ];
#pragma unroll
        for (int q = 0; q < QZ; ++q) {
            const int idx = threadIdx.x + q * KM_BLOCK;
            const int a = idx / WZ, b = idx - (idx / WZ) * WZ;
            const int i = ti0 - NS - 1 + a, j = tj0 - NS - 1 + b;
            const bool ing = idx < HZ * WZ && i >= 0 && i < NX && j >= 0 && j < NY;
            const int ic = i < 1 ? 0 : (i > nx ? nx - 1 : i - 1), jc = j < 1 ? 0 : (j > ny ? ny - 1 : j - 1);
            yq[q] = yc[(ic / B) * mJ + jc / B];
            mq[q] = ing ? mask[(int64_t)i * NY + j] : 0;
        }
#pragma unroll
        for (int q = 0; q < QZ; ++q) {
            const int idx = threadIdx.x + q * KM_BLOCK;
            if (idx < HZ * WZ) za[idx] = mq[q] ? (T)yq[q] : T(0);
        }
        __syncthreads();
        // sweep h (h = NS .. 1) on the halo-h frame, za and zb alternating
#pragma unroll
        for (int h = NS; h >= 1; --h) {
            const T* __restrict__ zi = ((NS - h) & 1) ? zb : za;
            T* __restrict__ zo = ((NS - h) & 1) ? za : zb;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int idx = threadIdx.x + q * KM_BLOCK;
                const int a = idx / WR, b = idx - (idx / WR) * WR;
                if (idx >= NR || a < NS - h || a >= TI + NS + h || b < NS - h || b >= TJ + NS + h) continue;
                const int pz = (a + 1) * WZ + (b + 1);
                T out = T(0);
                if (in[q]) {
                    const T zc = zi[pz];
                    T az = cr[q][0] * zc;
                    az = az + cr[q][1] * zi[pz - WZ];
                    az = az + cr[q][2] * zi[pz + WZ];
                    az = az + cr[q][3] * zi[pz + 1];
                    az = az + cr[q][4] * zi[pz - 1];
                    out = zc + wd[q] * (sr[q] - az);
                }
                zo[pz] = out;
            }
            __syncthreads();
        }
        const T* __restrict__ zf = (NS & 1) ? zb : za;   // the last sweep's output
        // the tile itself: z out, then y = A z
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int idx = threadIdx.x + q * KM_BLOCK;
            const int a = idx / WR, b = idx - (idx / WR) * WR;
            if (idx >= NR || a < NS || a >= TI + NS || b < NS || b >= TJ + NS || !in[q]) continue;
            const int pz = (a + 1) * WZ + (b + 1);
            const int64_t k = kq[q];
            const T zc = zf[pz];
            z[k] = zc;
            if (MODE > 0) {
                T yv = cr[q][0] * zc;
                yv = yv + cr[q][1] * zf[pz - WZ];
                yv = yv + cr[q][2] * zf[pz + WZ];
                yv = yv + cr[q][3] * zf[pz + 1];
                yv = yv + cr[q][4] * zf[pz - 1];
                y[k] = yv;
                const double rtk = (double)rq[q];
                if (MODE == 1) {
                    a0 += rtk * (double)yv;
                } else {
                    const double sk = (double)sr[q];
                    a0 += (double)yv * sk;
                    a1 += (double)yv * (double)yv;
                    a2 += sk * rtk;
                    a3 += (double)yv * rtk;
                }
            }
        }
        __syncthreads();
    }
    if (MODE == 1) {
        a0 = block_sum(a0, sh);
        if (threadIdx.x == 0) P[TL_PV + L] = a0;
    } else if (MODE == 2) {
        a0 = block_sum(a0, sh);
        a1 = block_sum(a1, sh);
        a2 = block_sum(a2, sh);
        a3 = block_sum(a3, sh);
        if (threadIdx.x == 0) {
            P[TL_PT + L] = a0;
            P[TL_PT + kMaxTileWg + L] = a1;
            P[TL_PT + 2 * kMaxTileWg + L] = a2;
            P[TL_PT + 3 * kMaxTileWg + L] = a3;
        }
    }
}

// make_tile with up to kMaxTileWg workgroups (one tile each when it fits)
Tile tl_tile(int NX, int NY, int TI, int TJ) {
    Tile t;
    t.NX = NX; t.NY = NY; t.TI = TI; t.TJ = TJ;
    t.tiles_j = (NY + TJ - 1) / TJ;
    t.T = ((NX + TI - 1) / TI) * t.tiles_j;
    t.tpb = (t.T + kMaxTileWg - 1) / kMaxTileWg;
    t.G = (t.T + t.tpb - 1) / t.tpb;
    return t;
}

// One smoothing launch with tile shape `which` (an index into kTileI / kTileJ).
template <typename T, int MODE>
void launch_smooth(int which, hipStream_t q, int NX, int NY, int B, int mJ, const T* c, const T* src,
                   const uint8_t* mask, const double* yc, const T* rt, T* z, T* y, const KState* st, double* P) {
#define KM_TL_CASE(IDX)                                                                                           \
    case IDX: {                                                                                                   \
        const Tile tl = tl_tile(NX, NY, kTileI[IDX], kTileJ[IDX]);                                                \
        hipLaunchKernelGGL((k_tl_smooth<T, kTileI[IDX], kTileJ[IDX], MODE, kTileW[IDX]>), dim3(tl.G),              \
                           dim3(KM_BLOCK), 0, q, tl,                                                                \
                           B, mJ, c, src, mask, yc, rt, z, y, st, P);                                             \
        break;                                                                                                    \
    }
    switch (which) {
        KM_TL_CASE(0)
        KM_TL_CASE(1)
        KM_TL_CASE(2)
        KM_TL_CASE(3)
        KM_TL_CASE(4)
        default:
        KM_TL_CASE(5)
    }
#undef KM_TL_CASE
}

int smooth_workgroups(int which, int NX, int NY) { return tl_tile(NX, NY, kTileI[which], kTileJ[which]).G; }

// The producers' partials summed by one workgroup (when kTotalWork says so)
template <int NV>
__global__ void __launch_bounds__(KM_BLOCK)
k_tl_total(const double* __restrict__ P, int G, int stride, double* __restrict__ out, const KState* __restrict__ st) {
    if (st->done) return;
    double t[NV];
    sum_partials_wide<NV>(P, G, t, stride);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < NV; ++q) out[q] = t[q];
}

int total_flags(const km_coarse* C, int Gt, int Gp, int Gs) {
    if (C->totals == 0) return 0;
    if (C->totals > 0) return FIN_V | FIN_T | FIN_R;
    return ((int64_t)Gt * Gs > kTotalWork ? FIN_V : 0) | ((int64_t)Gt * Gp > kTotalWork ? FIN_T : 0) |
           ((int64_t)Gp * Gs > kTotalWork ? FIN_R : 0);
}

// About `target` cells per workgroup: large aggregates are cut into at most
// kMaxSlices row slices, small ones grouped kMaxA at most along J; at most
// kMaxAggWg workgroups.  K_P, with seven loads per cell and four partial
// planes to reduce first, takes kStage x 256 (fewer workgroups, fewer
// redundant partial-sum reads); K_S, two loads per cell, takes 256.
constexpr int64_t kTargetP = (int64_t)kStage * KM_BLOCK, kTargetS = KM_BLOCK;
AggGeom agg_geom(const km_coarse* C, int64_t target) {
    AggGeom g;
    g.NX = C->NX; g.NY = C->NY; g.B = C->B; g.mI = C->mI; g.mJ = C->mJ; g.m = C->m;
    const int64_t cells = (int64_t)C->B * C->B;
    int S = 1, A = 1;
    if (cells >= target) {
        S = (int)(cells / target);
        if (S > kmcoarse::kMaxSlices) S = kmcoarse::kMaxSlices;
    } else {
        A = (int)(target / cells);
        if (A > kMaxA) A = kMaxA;
        if (A > C->mJ) A = C->mJ;
        if (A < 1) A = 1;
    }
    while (S > 1 && (int64_t)C->m * S > kMaxAggWg) --S;
    g.A = A;
    g.ngrp = (C->mJ + A - 1) / A;
    g.S = S;
    g.G = C->mI * g.ngrp * S;
    return g;
}

}  // namespace

namespace kmcoarse {

// Tile shape of the smoothing launches for the coarse level's system in T:
// on the first call for a (shape, precision), each candidate runs 3 + 16
// back-to-back MODE 1 launches on scratch vectors (ws: its first 4n values of
// T are overwritten) and the fastest is kept.
template <typename T>
int tuned_tile(km_ctx* ctx, const T* coefT, const T* src, T* ws, int* out) {
    km_coarse* C = ctx->coarse;
    const auto key = std::make_tuple(C->NX, C->NY, (int)sizeof(T));
    auto it = C->tiles.find(key);
    if (it != C->tiles.end()) {
        *out = it->second;
        return KM_OK;
    }
    const int64_t n = (int64_t)C->NX * C->NY;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    KM_HIP(ctx, hipEventCreate(&e0));
    KM_HIP(ctx, hipEventCreate(&e1));
    int best = 0;
    float best_ms = 0.f;
    for (int w = 0; w < kNumTiles; ++w) {
        auto once = [&]() {
            launch_smooth<T, 1>(w, ctx->stream, C->NX, C->NY, C->B, C->mJ, coefT, src, (const uint8_t*)C->mask,
                                (const double*)C->yc, (const T*)ws, ws + 2 * n, ws + 3 * n, (const KState*)nullptr,
                                ctx->partials + TL_PV);
        };
        for (int k = 0; k < 3; ++k) once();
        (void)hipEventRecord(e0, ctx->stream);
        for (int k = 0; k < 16; ++k) once();
        (void)hipEventRecord(e1, ctx->stream);
        hipError_t se = hipEventSynchronize(e1);
        float ms = 0.f;
        if (se == hipSuccess) se = hipEventElapsedTime(&ms, e0, e1);
        if (se != hipSuccess) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return km_fail(ctx, KM_EHIP, "smoothing tile timing: %s", hipGetErrorString(se));
        }
        if (w == 0 || ms < best_ms) {
            best = w;
            best_ms = ms;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    C->tiles[key] = best;
    *out = best;
    return KM_OK;
}

int tile_info(km_ctx* ctx, int* ti, int* tj) {
    const km_coarse* C = ctx->coarse;
    KM_REQUIRE(ctx, C && C->arena, "no coarse level");
    auto it = C->tiles.find(std::make_tuple(C->NX, C->NY, (int)sizeof(double)));
    const int w = it == C->tiles.end() ? 0 : it->second;
    *ti = kTileI[w];
    *tj = kTileJ[w];
    return KM_OK;
}

template <typename T>
int smooth(km_ctx* ctx, const T* coefT, const T* r, T* z, hipStream_t s) {
    const km_coarse* C = ctx->coarse;
    KM_REQUIRE(ctx, C && C->arena && C->m > 0, "coarse level not set up");
    launch_smooth<T, 0>(0, s, C->NX, C->NY, C->B, C->mJ, coefT, r, (const uint8_t*)C->mask, (const double*)C->yc,
                        (const T*)nullptr, z, (T*)nullptr, (const KState*)nullptr, (double*)nullptr);
    KM_CHECK_LAUNCH(ctx);
    return KM_OK;
}

template <typename T>
int fused_inner(km_ctx* ctx, int NX, int NY, const T* coefT, const T* rhs, T* xd, T* ws, double eps, int max_iter,
                bool use_graph, KState* out) {
    km_coarse* C = ctx->coarse;
    KM_REQUIRE(ctx, C && C->arena && C->NX == NX && C->NY == NY, "coarse level not set up for this system");
    KM_REQUIRE(ctx, C->m <= kMaxAggWg, "coarse level of %d unknowns", C->m);
    const int64_t n = (int64_t)NX * NY;
    T* r = ws;
    T* rt = ws + n;
    T* p = ws + 2 * n;
    T* v = ws + 3 * n;
    T* s = ws + 4 * n;
    T* t = ws + 5 * n;
    T* ph = ws + 6 * n;
    T* sh = ws + 7 * n;
    KState* st = &ctx->st[4];
    double* P = ctx->partials;
    const Flat fl = make_flat(n);
    const AggGeom g = agg_geom(C, kTargetP), gs = agg_geom(C, kTargetS);
    const uint8_t* mask = C->mask;
    double* rc = C->rc;
    const double* yc = C->yc;
    int tile = 0;
    if (int rc_ = tuned_tile<T>(ctx, coefT, rhs, ws, &tile)) return rc_;
    const int Gt = smooth_workgroups(tile, NX, NY);
    const int fin = total_flags(C, Gt, g.G, gs.G);
    hipLaunchKernelGGL(k_tl_init<T>, dim3(fl.G), dim3(KM_BLOCK), 0, ctx->stream, fl, rhs, xd, r, rt, eps, max_iter, st);
    KM_CHECK_LAUNCH(ctx);
    const int* skip = &st->done;
    auto enqueue = [=](hipStream_t q, int phs) {
        hipLaunchKernelGGL(k_tl_p<T>, dim3(g.G), dim3(KM_BLOCK), 0, q, g, Gt, fin, phs, mask, xd, r, p, (const T*)v,
                           (const T*)s, (const T*)t, (const T*)ph, (const T*)sh, rc, st, P);
        if (fin & FIN_R)
            hipLaunchKernelGGL(k_tl_total<1>, dim3(1), dim3(KM_BLOCK), 0, q, (const double*)(P + TL_PR), g.G,
                               kMaxAggWg, P + TL_TOT + 5, (const KState*)st);
        (void)coarse_solve(ctx, g.S, skip, q);
        launch_smooth<T, 1>(tile, q, NX, NY, g.B, g.mJ, coefT, (const T*)p, mask, yc, (const T*)rt, ph, v,
                            (const KState*)st, P);
        if (fin & FIN_V)
            hipLaunchKernelGGL(k_tl_total<1>, dim3(1), dim3(KM_BLOCK), 0, q, (const double*)(P + TL_PV), Gt,
                               kMaxTileWg, P + TL_TOT, (const KState*)st);
        hipLaunchKernelGGL(k_tl_s<T>, dim3(gs.G), dim3(KM_BLOCK), 0, q, gs, Gt, g.G, fin, phs, mask, (const T*)r,
                           (const T*)v, s, rc, st, P);
        (void)coarse_solve(ctx, gs.S, skip, q);
        launch_smooth<T, 2>(tile, q, NX, NY, g.B, g.mJ, coefT, (const T*)s, mask, yc, (const T*)rt, sh, t,
                            (const KState*)st, P);
        if (fin & FIN_T)
            hipLaunchKernelGGL(k_tl_total<4>, dim3(1), dim3(KM_BLOCK), 0, q, (const double*)(P + TL_PT), Gt,
                               kMaxTileWg, P + TL_TOT + 1, (const KState*)st);
    };
    GraphKey key;
    key.v = {8, sizeof(T), (uintptr_t)NX, (uintptr_t)NY, (uintptr_t)coefT, (uintptr_t)rhs, (uintptr_t)xd,
             (uintptr_t)ws, (uintptr_t)st, (uintptr_t)C, (uintptr_t)C->arena, (uintptr_t)C->m, (uintptr_t)C->B,
             (uintptr_t)tile, (uintptr_t)fin};
    return run_chunks(ctx, st, key, use_graph, kChunk, max_iter, enqueue, out);
}

// Average device time of one launch of a fused two-level kernel (which: 4 K_P,
// 5 K_V, 6 K_S, 7 K_T) inside FP64 iterations on (coef, b) with the stopping
// test disabled; events around every launch of that kernel, or (b2b) around
// reps identical launches after eight warm iterations.
int time_kernel(km_ctx* ctx, int which, int NX, int NY, const double* coef, const double* b, int reps, bool b2b,
                double* avg_us) {
    KM_REQUIRE(ctx, which >= 4 && which <= 7, "unknown kernel id %d", which);
    if (int rc = setup(ctx, NX, NY, coef, ctx->stream)) return rc;
    const int64_t n = (int64_t)NX * NY;
    if (int rc = km_ensure_workspace(ctx, (size_t)9 * n * sizeof(double))) return rc;
    double* ws = static_cast<double*>(km_ws_vectors(ctx));
    km_coarse* C = ctx->coarse;
    double *r = ws, *rt = ws + n, *p = ws + 2 * n, *v = ws + 3 * n, *sv = ws + 4 * n, *t = ws + 5 * n;
    double *ph = ws + 6 * n, *sh = ws + 7 * n, *x = ws + 8 * n;
    KState* st = &ctx->st[4];
    double* P = ctx->partials;
    const Flat fl = make_flat(n);
    const AggGeom g = agg_geom(C, kTargetP), gs = agg_geom(C, kTargetS);
    const uint8_t* mask = C->mask;
    int tile = 0;
    if (int rc = tuned_tile<double>(ctx, coef, b, ws, &tile)) return rc;
    const int Gt = smooth_workgroups(tile, NX, NY);
    const int fin = total_flags(C, Gt, g.G, gs.G);
    std::vector<hipEvent_t> ev(2 * (size_t)reps, nullptr);
    for (auto& e : ev) KM_HIP(ctx, hipEventCreate(&e));
    auto cleanup = [&]() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
    };
    hipLaunchKernelGGL(k_tl_init<double>, dim3(fl.G), dim3(KM_BLOCK), 0, ctx->stream, fl, b, x, r, rt, 0.0, 1 << 30,
                       st);
    hipStream_t q = ctx->stream;
    // the kernel alone
    auto launch = [&](int k, int phs) {
        switch (k) {
        case 4:
            hipLaunchKernelGGL(k_tl_p<double>, dim3(g.G), dim3(KM_BLOCK), 0, q, g, Gt, fin, phs, mask, x, r, p,
                               (const double*)v, (const double*)sv, (const double*)t, (const double*)ph,
                               (const double*)sh, C->rc, st, P);
            break;
        case 5:
            launch_smooth<double, 1>(tile, q, NX, NY, g.B, g.mJ, coef, (const double*)p, mask, (const double*)C->yc,
                                     (const double*)rt, ph, v, (const KState*)st, P);
            break;
        case 6:
            hipLaunchKernelGGL(k_tl_s<double>, dim3(gs.G), dim3(KM_BLOCK), 0, q, gs, Gt, g.G, fin, phs, mask,
                               (const double*)r, (const double*)v, sv, C->rc, st, P);
            break;
        default:
            launch_smooth<double, 2>(tile, q, NX, NY, g.B, g.mJ, coef, (const double*)sv, mask, (const double*)C->yc,
                                     (const double*)rt, sh, t, (const KState*)st, P);
        }
    };
    // one() adds what follows each kernel in the iteration: the totals and the coarse solves
    auto one = [&](int k, int phs, hipEvent_t e0, hipEvent_t e1) {
        if (e0) (void)hipEventRecord(e0, q);
        launch(k, phs);
        if (e1) (void)hipEventRecord(e1, q);
        if (k == 4 && (fin & FIN_R))
            hipLaunchKernelGGL(k_tl_total<1>, dim3(1), dim3(KM_BLOCK), 0, q, (const double*)(P + TL_PR), g.G,
                               kMaxAggWg, P + TL_TOT + 5, (const KState*)st);
        if (k == 5 && (fin & FIN_V))
            hipLaunchKernelGGL(k_tl_total<1>, dim3(1), dim3(KM_BLOCK), 0, q, (const double*)(P + TL_PV), Gt,
                               kMaxTileWg, P + TL_TOT, (const KState*)st);
        if (k == 7 && (fin & FIN_T))
            hipLaunchKernelGGL(k_tl_total<4>, dim3(1), dim3(KM_BLOCK), 0, q, (const double*)(P + TL_PT), Gt,
                               kMaxTileWg, P + TL_TOT + 1, (const KState*)st);
        if (k == 4) (void)coarse_solve(ctx, g.S, nullptr, q);
        if (k == 6) (void)coarse_solve(ctx, gs.S, nullptr, q);
    };
    const int warm = 8;
    for (int it = 0; it < warm; ++it)
        for (int k = 4; k <= 7; ++k) one(k, it % KM_RING, nullptr, nullptr);
    if (b2b) {
        // K_P re-runs the scalars and updates of the next phase, the others
        // read what the last warm iteration left
        const int phs = which == 4 ? warm % KM_RING : (warm - 1) % KM_RING;
        KM_HIP(ctx, hipEventRecord(ev[0], q));
        for (int k = 0; k < reps; ++k) launch(which, phs);
        KM_HIP(ctx, hipEventRecord(ev[1], q));
    } else {
        for (int it = 0; it < reps; ++it)
            for (int k = 4; k <= 7; ++k)
                one(k, (warm + it) % KM_RING, k == which ? ev[2 * it] : nullptr, k == which ? ev[2 * it + 1] : nullptr);
    }
    hipError_t le = hipGetLastError();
    hipError_t se = hipStreamSynchronize(q);
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
    KM_HIP(ctx, hipMemcpy(&ctx->h_st[3], st, sizeof(KState), hipMemcpyDeviceToHost));
    KM_REQUIRE(ctx, !ctx->h_st[3].done, "km_time_kernel: the iteration stopped during timing (status %d)",
               ctx->h_st[3].status);
    *avg_us = 1e3 * total_ms / reps;
    return KM_OK;
}

template int smooth<float>(km_ctx*, const float*, const float*, float*, hipStream_t);
template int smooth<double>(km_ctx*, const double*, const double*, double*, hipStream_t);
template int fused_inner<float>(km_ctx*, int, int, const float*, const float*, float*, float*, double, int, bool,
                                KState*);
template int fused_inner<double>(km_ctx*, int, int, const double*, const double*, double*, double*, double, int,
                                 bool, KState*);

}  // namespace kmcoarse

extern "C" {

int km_set_coarse_totals(km_ctx* ctx, int mode) {
    if (!ctx) return KM_EINVAL;
    KM_REQUIRE(ctx, ctx->coarse != nullptr, "no coarse level yet (solve once first)");
    ctx->coarse->totals = mode < 0 ? -1 : (mode > 0 ? 1 : 0);
    return KM_OK;
}

int km_twolevel_tile(km_ctx* ctx, int* ti, int* tj) {
    if (!ctx || !ti || !tj) return KM_EINVAL;
    return kmcoarse::tile_info(ctx, ti, tj);
}

int km_twolevel_precond(km_ctx* ctx, int NX, int NY, const double* coef, const double* r, double* z) {
    if (!ctx) return KM_EINVAL;
    KM_REQUIRE(ctx, coef && r && z, "null pointer");
    KM_REQUIRE(ctx, NX >= 3 && NY >= 3 && (int64_t)NX * NY < ((int64_t)1 << 31), "system shape %d x %d", NX, NY);
    if (int rc = kmcoarse::setup(ctx, NX, NY, coef, ctx->stream)) return rc;
    if (int rc = kmcoarse::restrict_solve<double>(ctx, r, nullptr, ctx->stream)) return rc;
    return kmcoarse::smooth<double>(ctx, coef, r, z, ctx->stream);
}

}  // extern "C"
