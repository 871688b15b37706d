// Fused flexible BiCGSTAB with the two-level preconditioner
//   M(r) = z1, then kSweeps red-black SOR sweeps on A z = r,  z1 = P A_c^-1 P^T r
// (the aggregation coarse level of km_coarse.hip), the preconditioner of the
// p^ / s^ hooks of bicgsolver.py:48-58.
//
// The restriction is linear, so the coarse level follows the Krylov vectors
// through their recurrences: with Y(w) = A_c^-1 P^T w,
//   Y(s) = Y(r) - alpha Y(v),  Y(r') = Y(s) - omega Y(t),
//   Y(p') = Y(r') + beta (Y(p) - omega Y(v)),
// and only v = A p^ and t = A s^ are restricted, by the tile launches that
// compute them.  Four launches per inner iteration i (ring slot ph = i mod
// KM_RING):
//   K_PV  2-D tiles with a (2 kSweeps + 1)-cell halo.  x += alpha p^ + omega s^,
//         r = s - omega t (iteration i-1) and p = r + beta (p - omega v) on the
//         tile and its halo; p^ = M(p) with z1 from Y(p), v = A p^;
//         partials <r~,v>, <r,r>; slot partials of P^T v
//   K_CV  stopping test on <r,r>, alpha = rho / <r~,v>;
//         Y(v) = A_c^-1 P^T v, Y(s) = Y(r) - alpha Y(v)
//   K_ST  tiles: s = r - alpha v on the tile and its halo; s^ = M(s) with z1
//         from Y(s), t = A s^; partials <t,s>, <t,t>, <s,r~>, <t,r~>; P^T t
//   K_CT  omega, rho, beta of the next iteration; Y(t) = A_c^-1 P^T t,
//         Y(r) = Y(s) - omega Y(t), Y(p) = Y(r) + beta (Y(p) - omega Y(v))
// p and v alternate between two buffers by iteration parity (K_PV reads the
// previous p and v on its halo while other tiles write the new ones).
//
// Tiles are aligned with the aggregates: tile (a, b) owns the interior rows
// [1 + a TI, 1 + (a + 1) TI) and columns [1 + b TJ, ...), clipped, and the
// boundary row or column next to it at the edges of the system (one cell
// into the halo, where M(.) is still exact).  An aggregate's restriction
// then comes from the few tiles that overlap it, each writing its partial sum
// into its own slot: slot (a - I B / TI) SJ + (b - J B / TJ) of aggregate
// (I, J).  The coarse launches sum the slots in slot order.
#include "km_coarse.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace {

constexpr int kMaxTileWg = 8192;                 // smoothing workgroups (one tile each up to 8192 tiles)
constexpr int TL_PA = 0;                         // [2][kMaxTileWg] <r~,v>, <r,r> (K_PV)
constexpr int TL_PT = 2 * kMaxTileWg;            // [4][kMaxTileWg] <t,s>, <t,t>, <s,r~>, <t,r~> (K_ST)
constexpr int TL_TOT = 6 * kMaxTileWg;           // totals: 0..1 the K_PV planes, 2..5 the K_ST planes
// Every coarse workgroup sums the tile partials; when partials x coarse
// workgroups exceeds this, one workgroup sums them first (k_tl_total) and the
// coarse launches read the totals.
constexpr int64_t kTotalWork = 2000000;
constexpr int FIN_A = 1, FIN_T = 2;
constexpr int kStallWindow = 300;        // as the unfused flexible solve
constexpr int kChunk = 8;
// Tile shapes of the smoothing launches (rows x columns, columns contiguous);
// the fastest for a system is chosen on its first solve.
constexpr int kNumTiles = 6;
constexpr int kTileI[kNumTiles] = {32, 16, 16, 8, 8, 16};
constexpr int kTileJ[kNumTiles] = {32, 64, 32, 64, 128, 32};
// minimum waves per SIMD the compiler must allow (a VGPR budget): 16 x 32 also
// fits 128 VGPRs, so a fourth wave per SIMD can hide latency
constexpr int kTileW[kNumTiles] = {1, 1, 3, 1, 1, 4};

__device__ __forceinline__ bool fin(double v) { return v == v && v - v == 0.0; }

__device__ __forceinline__ void finish_state(KState* st, int status, int iter, double rr) {
    st->status = status;
    st->iter = iter;
    st->rTr = rr;
    st->done = 1;
}

// Tiles aligned with the aggregates (see the file header).
struct ATile {
    int NX, NY, TI, TJ, tiles_i, tiles_j, T, tpb, G;
};

ATile a_tile(int NX, int NY, int TI, int TJ) {
    ATile t;
    t.NX = NX; t.NY = NY; t.TI = TI; t.TJ = TJ;
    t.tiles_i = (NX - 2 + TI - 1) / TI;
    t.tiles_j = (NY - 2 + TJ - 1) / TJ;
    t.T = t.tiles_i * t.tiles_j;
    t.tpb = (t.T + kMaxTileWg - 1) / kMaxTileWg;
    t.G = (t.T + t.tpb - 1) / t.tpb;
    return t;
}

// Slots per aggregate of a tile shape: the most tiles that overlap one
// aggregate along I (SI) and along J (SJ).
int slots_along(int n, int B, int Tt) {
    int most = 1;
    for (int I = 0; I * B < n; ++I) {
        const int e = ((I + 1) * B < n ? (I + 1) * B : n) - 1;
        const int c = e / Tt - (I * B) / Tt + 1;
        if (c > most) most = c;
    }
    return most;
}

// Coarse geometry as the tile launches see it.
struct CGeom {
    int B, mJ, m, SJ;
};

// Launch arguments of the tile kernels.  Frame inputs f0..f3 are read on the
// tile and its 2 kSweeps-cell halo:
//   MODE 0  z = M(f0)
//   MODE 1  f0 = s, f1 = t, f2 = p (previous), f3 = v (previous); r read at i = 0
//   MODE 2  f0 = r, f1 = v
template <typename T>
struct TlArgs {
    const T* c;              // 5 coefficient planes
    const uint8_t* mask;     // free cells
    const T* f0;
    const T* f1;
    const T* f2;
    const T* f3;
    const T* rt;             // r~ (MODE 1, 2)
    T* x;                    // MODE 1: x += alpha p^ + omega s^ (i > 0)
    T* r;                    // MODE 1: r out (i > 0), r_0 in (i = 0)
    const T* sh;             // MODE 1: the previous s^
    T* src;                  // MODE 1: p out; MODE 2: s out
    T* z;                    // M(src) out (MODE 1: also the previous p^ in)
    T* y;                    // A z out (MODE 1, 2)
    const double* yc;        // coarse vector prolongated as z1
    double* rcs;             // slot partials of P^T y
    const KState* st;        // null: no state (tuning, MODE 0)
    double* P;               // partial planes of this launch
    int ph;
};

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
        st->ring[0].iter = 0;
        st->ring[0].omega = 1.0;
        st->ring[0].beta = 0.0;
        st->ring[KM_RING - 1].iter = -1;
        st->ring[KM_RING - 1].alpha = 0.0;
        st->ring[KM_RING - 1].best = 1e300;
        st->ring[KM_RING - 1].best_it = 0;
    }
}

// Y(r) = Y(p) = yc (the coarse solve of r_0) and the slot partials cleared
// (slots no tile writes stay zero for the whole solve).
__global__ void __launch_bounds__(KM_BLOCK)
k_tl_cinit(int m, int S, const double* __restrict__ yc, double* __restrict__ Y, double* __restrict__ rcv,
           double* __restrict__ rct) {
    const int64_t k = (int64_t)blockIdx.x * KM_BLOCK + threadIdx.x;
    if (k < m) {
        Y[k] = yc[k];
        Y[m + k] = yc[k];
    }
    if (k < (int64_t)S * m) {
        rcv[k] = 0.0;
        rct[k] = 0.0;
    }
}

// K_PV (MODE 1) / K_ST (MODE 2) / the standalone smoother (MODE 0).  The
// coefficients and the frame input of the tile's NS-cell halo frame stay in
// registers (Q cells per thread, NS = 2 kSweeps half-sweeps); z lives in LDS
// on the (NS + 1)-cell halo frame and is updated in place: a red half-sweep
// changes only red cells ((i + j) even) and reads only black ones, and the
// other way round.  The arithmetic per cell is z + (omega / aP) (src - A z),
// neighbours outside the system read as zero, rows with aP = 0 left
// unchanged.  Half-sweep s covers the halo-(NS - s) frame, so after the last
// one z is exact one cell into the halo, which covers the boundary cells the
// edge tiles own.
template <typename T, int TI, int TJ, int MODE, int WPE>
__global__ void __launch_bounds__(KM_BLOCK) __attribute__((amdgpu_waves_per_eu(WPE, 8)))
k_tl_tile(ATile tl, CGeom cg, TlArgs<T> g) {
    constexpr int NS = 2 * kmcoarse::kSweeps;                       // half-sweeps = halo
    constexpr int WZ = TJ + 2 * (NS + 1), HZ = TI + 2 * (NS + 1);   // z frame: halo NS + 1
    constexpr int WR = TJ + 2 * NS, HR = TI + 2 * NS;               // register frame: halo NS
    constexpr int NR = WR * HR;
    constexpr int Q = (NR + KM_BLOCK - 1) / KM_BLOCK;
    __shared__ T za[HZ * WZ];
    __shared__ double yb[MODE > 0 ? TI * TJ : 1];   // y on the free interior cells (tile-local)
    __shared__ double sh[8];
    const KState* st = g.st;
    if (st && st->done) return;
    const int L = xcd_remap(blockIdx.x, tl.G);
    const int NX = tl.NX, NY = tl.NY, nx = NX - 2, ny = NY - 2;
    const int64_t n = (int64_t)NX * NY;
    const T om = (T)kmcoarse::kOmega;
    // scalars of the iteration
    int it = 0;
    double alpha = 0.0, omega = 0.0, beta = 0.0;
    if (MODE == 1 && st) {
        const KScal sp = st->ring[g.ph];
        it = sp.iter;
        omega = sp.omega;
        beta = sp.beta;
        alpha = st->ring[(g.ph + KM_RING - 1) % KM_RING].alpha;
    } else if (MODE == 2 && st) {
        alpha = st->ring[g.ph].alpha;
    }
    const T fo = (T)omega, fb = (T)beta, fa = (T)alpha;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    const int t0 = L * tl.tpb, t1 = t0 + tl.tpb < tl.T ? t0 + tl.tpb : tl.T;
    for (int tile = t0; tile < t1; ++tile) {
        const int ta = tile / tl.tiles_j, tb = tile - ta * tl.tiles_j;
        const int ti0 = 1 + ta * TI, tj0 = 1 + tb * TJ;
        // cells this tile writes: its interior block plus the boundary row or
        // column next to it at the edges of the system
        const int olo = ta == 0 ? 0 : ti0, ohi = ta == tl.tiles_i - 1 ? NX : ti0 + TI;
        const int jlo = tb == 0 ? 0 : tj0, jhi = tb == tl.tiles_j - 1 ? NY : tj0 + TJ;
        // every load of the tile is issued before the first store (the
        // argument pointers may alias as far as the compiler knows): the
        // register frame, the tile's own cells and the z frame
        constexpr int NF = MODE == 1 ? 4 : (MODE == 2 ? 2 : 1);   // frame inputs per cell
        constexpr int QZ = (HZ * WZ + KM_BLOCK - 1) / KM_BLOCK;
        T cr[Q][5], sr[Q], rq[Q], wd[Q];
        T fv[Q][NF];
        int64_t kq[Q];
        bool in[Q], ou[Q];
        unsigned ib = 0, fm = 0;   // bit q: interior cell of the tile; free cell
        unsigned blk = 0;          // bit q: black cell ((i + j) odd)
        double yq[QZ];
        uint8_t mq[QZ];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int idx = threadIdx.x + q * KM_BLOCK;
            const int a = idx / WR, b = idx - (idx / WR) * WR;
            const int i = ti0 - NS + a, j = tj0 - NS + b;
            in[q] = idx < NR && i >= 0 && i < NX && j >= 0 && j < NY;
            ou[q] = in[q] && i >= olo && i < ohi && j >= jlo && j < jhi;
            kq[q] = in[q] ? (int64_t)i * NY + j : 0;
            const int64_t k = kq[q];
            if ((i + j) & 1) blk |= 1u << q;
            if (MODE > 0 && ou[q] && i >= 1 && i <= nx && j >= 1 && j <= ny) {
                ib |= 1u << q;
                if (g.mask[k]) fm |= 1u << q;
            }
#pragma unroll
            for (int e = 0; e < 5; ++e) cr[q][e] = in[q] ? g.c[e * n + k] : T(0);
            if (MODE == 1 && it == 0) {
                fv[q][0] = in[q] ? g.r[k] : T(0);   // p_0 = r_0
#pragma unroll
                for (int e = 1; e < NF; ++e) fv[q][e] = T(0);
            } else {
                const T* fp[4] = {g.f0, g.f1, g.f2, g.f3};
#pragma unroll
                for (int e = 0; e < NF; ++e) fv[q][e] = in[q] ? fp[e][k] : T(0);
            }
            rq[q] = MODE > 0 && ou[q] ? g.rt[k] : T(0);
        }
        // z1 = P yc on the z frame (zero off the free cells and outside the system)
#pragma unroll
        for (int q = 0; q < QZ; ++q) {
            const int idx = threadIdx.x + q * KM_BLOCK;
            const int a = idx / WZ, b = idx - (idx / WZ) * WZ;
            const int i = ti0 - NS - 1 + a, j = tj0 - NS - 1 + b;
            const bool ing = idx < HZ * WZ && i >= 0 && i < NX && j >= 0 && j < NY;
            const int ic = i < 1 ? 0 : (i > nx ? nx - 1 : i - 1), jc = j < 1 ? 0 : (j > ny ? ny - 1 : j - 1);
            yq[q] = g.yc[(ic / cg.B) * cg.mJ + jc / cg.B];
            mq[q] = ing ? g.mask[(int64_t)i * NY + j] : 0;
        }
        // the source of M(.) on the frame; the Krylov updates of the tile's cells
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int64_t k = kq[q];
            T sv = T(0);
            if (MODE == 0) {
                sv = fv[q][0];
            } else if (MODE == 1) {
                if (it == 0) {
                    sv = fv[q][0];
                    if (ou[q]) a1 += (double)sv * (double)sv;
                } else if (in[q]) {
                    const T rn = fv[q][0] - fo * fv[q][1];             // bicgsolver.py:134
                    sv = rn + fb * (fv[q][2] - fo * fv[q][3]);         // bicgsolver.py:103
                    if (ou[q]) {
                        g.r[k] = rn;
                        a1 += (double)rn * (double)rn;
                    }
                }
                if (ou[q]) g.src[k] = sv;
            } else {
                sv = in[q] ? fv[q][0] - fa * fv[q][1] : T(0);        // bicgsolver.py:110
                if (ou[q]) g.src[k] = sv;
            }
            sr[q] = sv;
            wd[q] = cr[q][0] != T(0) ? om * (T(1) / cr[q][0]) : T(0);   // omega D^-1
        }
#pragma unroll
        for (int q = 0; q < QZ; ++q) {
            const int idx = threadIdx.x + q * KM_BLOCK;
            if (idx < HZ * WZ) za[idx] = mq[q] ? (T)yq[q] : T(0);
        }
        __syncthreads();
        // MODE 1: x, the previous p^ and s^ on the tile's cells, in flight
        // during the sweeps (x += alpha p^ + omega s^ before p^ is replaced)
        T ex[Q][3];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const bool xo = MODE == 1 && it > 0 && ou[q];
            ex[q][0] = xo ? g.x[kq[q]] : T(0);
            ex[q][1] = xo ? g.z[kq[q]] : T(0);
            ex[q][2] = xo ? g.sh[kq[q]] : T(0);
        }
        // half-sweep hs: red cells at even hs, black at odd, on the halo-h frame
#pragma unroll
        for (int hs = 0; hs < NS; ++hs) {
            const int h = NS - hs;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int idx = threadIdx.x + q * KM_BLOCK;
                const int a = idx / WR, b = idx - (idx / WR) * WR;
                if (idx >= NR || !in[q] || (int)((blk >> q) & 1u) != (hs & 1) || a < NS - h || a >= TI + NS + h ||
                    b < NS - h || b >= TJ + NS + h)
                    continue;
                const int pz = (a + 1) * WZ + (b + 1);
                const T zc = za[pz];
                T az = cr[q][0] * zc;
                az = az + cr[q][1] * za[pz - WZ];
                az = az + cr[q][2] * za[pz + WZ];
                az = az + cr[q][3] * za[pz + 1];
                az = az + cr[q][4] * za[pz - 1];
                za[pz] = zc + wd[q] * (sr[q] - az);
            }
            __syncthreads();
        }
        const T* __restrict__ zf = za;
        // the tile's cells: z out, then y = A z (and y on the free interior
        // cells into yb for the restriction)
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            if (!ou[q]) continue;
            const int idx = threadIdx.x + q * KM_BLOCK;
            const int a = idx / WR, b = idx - (idx / WR) * WR;
            const int pz = (a + 1) * WZ + (b + 1);
            const int64_t k = kq[q];
            const T zc = zf[pz];
            if (MODE == 1 && it > 0)   // bicgsolver.py:128
                g.x[k] = ex[q][0] + (T)(alpha * (double)ex[q][1] + omega * (double)ex[q][2]);
            if (MODE == 0) {
                g.z[k] = zc;
            } else {
                g.z[k] = zc;
                T yv = cr[q][0] * zc;
                yv = yv + cr[q][1] * zf[pz - WZ];
                yv = yv + cr[q][2] * zf[pz + WZ];
                yv = yv + cr[q][3] * zf[pz + 1];
                yv = yv + cr[q][4] * zf[pz - 1];
                g.y[k] = yv;
                if ((ib >> q) & 1u) yb[(a - NS) * TJ + (b - NS)] = ((fm >> q) & 1u) ? (double)yv : 0.0;
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
        if (MODE > 0) {
            // P^T y over the aggregates this tile overlaps: one wave per
            // aggregate piece of yb, summed in a fixed order
            __syncthreads();
            const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
            const int r0 = ti0 - 1, c0 = tj0 - 1;   // interior coordinates of the tile origin
            const int r1 = r0 + TI < nx ? r0 + TI : nx, c1 = c0 + TJ < ny ? c0 + TJ : ny;
            const int B = cg.B;
            const int I0 = r0 / B, J0 = c0 / B;
            const int nJ = (c1 - 1) / B - J0 + 1, np = ((r1 - 1) / B - I0 + 1) * nJ;
            for (int u = w; u < np; u += 4) {
                const int I = I0 + u / nJ, J = J0 + u % nJ;
                const int ra = I * B > r0 ? I * B : r0, rb = (I + 1) * B < r1 ? (I + 1) * B : r1;
                const int ca = J * B > c0 ? J * B : c0, cb = (J + 1) * B < c1 ? (J + 1) * B : c1;
                const int wdt = cb - ca, cnt = (rb - ra) * wdt;
                double acc = 0.0;
                for (int e = lane; e < cnt; e += 64) {
                    const int er = e / wdt;
                    acc += yb[(ra - r0 + er) * TJ + (ca - c0) + (e - er * wdt)];
                }
                acc = wave_sum(acc);
                if (lane == 0) {
                    const int slot = (ta - (I * B) / TI) * cg.SJ + (tb - (J * B) / TJ);
                    g.rcs[(int64_t)slot * cg.m + I * cg.mJ + J] = acc;
                }
            }
        }
        __syncthreads();
    }
    if (MODE == 1) {
        a0 = block_sum(a0, sh);
        a1 = block_sum(a1, sh);
        if (threadIdx.x == 0) {
            g.P[L] = a0;
            g.P[kMaxTileWg + L] = a1;
        }
    } else if (MODE == 2) {
        a0 = block_sum(a0, sh);
        a1 = block_sum(a1, sh);
        a2 = block_sum(a2, sh);
        a3 = block_sum(a3, sh);
        if (threadIdx.x == 0) {
            g.P[L] = a0;
            g.P[kMaxTileWg + L] = a1;
            g.P[2 * kMaxTileWg + L] = a2;
            g.P[3 * kMaxTileWg + L] = a3;
        }
    }
}

// One tile launch with tile shape `which` (an index into kTileI / kTileJ).
template <typename T, int MODE>
void launch_tile(int which, hipStream_t q, int NX, int NY, const CGeom& cg, const TlArgs<T>& a) {
#define KM_TL_CASE(IDX)                                                                                           \
    case IDX: {                                                                                                   \
        const ATile tl = a_tile(NX, NY, kTileI[IDX], kTileJ[IDX]);                                               \
        hipLaunchKernelGGL((k_tl_tile<T, kTileI[IDX], kTileJ[IDX], MODE, kTileW[IDX]>), dim3(tl.G), dim3(KM_BLOCK), \
                           0, q, tl, cg, a);                                                                      \
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

int tile_workgroups(int which, int NX, int NY) { return a_tile(NX, NY, kTileI[which], kTileJ[which]).G; }

CGeom c_geom(const km_coarse* C, int which) {
    CGeom g;
    g.B = C->B;
    g.mJ = C->mJ;
    g.m = C->m;
    g.SJ = slots_along(C->NY - 2, C->B, kTileJ[which]);
    return g;
}

int tile_slots(const km_coarse* C, int which) {
    return slots_along(C->NX - 2, C->B, kTileI[which]) * slots_along(C->NY - 2, C->B, kTileJ[which]);
}

// A state for timing K_PV outside a solve: iteration 1, not done.
__global__ void k_tl_tune_state(KState* st) {
    st->done = 0;
    st->ring[0].iter = 1;
    st->ring[0].omega = 0.5;
    st->ring[0].beta = 0.5;
    st->ring[KM_RING - 1].alpha = 0.5;
}

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

int total_flags(const km_coarse* C, int Gt, int Gc) {
    if (C->totals == 0) return 0;
    if (C->totals > 0) return FIN_A | FIN_T;
    return (int64_t)Gt * Gc > kTotalWork ? FIN_A | FIN_T : 0;
}

// K_CV (MODE 1) / K_CT (MODE 2): the scalars from the tile launch's partials
// (every workgroup sums them in the same order), then one wave per coarse row
// of y = A_c^-1 rc, rc the slot sum (staged in LDS).  A lane's first
// kCoarseCh row values, rc and the row's coarse values are loaded before the
// partials are summed, so one memory round trip serves them all (rows of up to
// 64 kCoarseCh unknowns).  Y holds Y(r), Y(p), Y(v), Y(s).
constexpr int kCoarseCh = 32;

template <int MODE>
__global__ void __launch_bounds__(KM_BLOCK)
k_tl_coarse(int m, int S, int Gt, int tot, int ph, const double* __restrict__ inv, const double* __restrict__ rcs,
            double* __restrict__ Y, KState* __restrict__ st, const double* __restrict__ P) {
    extern __shared__ double rs_[];
    if (st->done) return;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    const bool live = row < m;
    const int rw = live ? row : 0;
    const double* a = inv + (int64_t)rw * m;
    double y0 = 0.0, y1 = 0.0, y2 = 0.0, y3 = 0.0;
    if (lane == 0) {
        y0 = Y[rw];
        y1 = Y[m + rw];
        y2 = Y[2 * m + rw];
        y3 = Y[3 * m + rw];
    }
    double av[kCoarseCh];
#pragma unroll
    for (int u = 0; u < kCoarseCh; ++u) av[u] = lane + 64 * u < m ? a[lane + 64 * u] : 0.0;
    for (int k = threadIdx.x; k < m; k += KM_BLOCK) {
        double acc = rcs[k];
        for (int q = 1; q < S; ++q) acc += rcs[(int64_t)q * m + k];
        rs_[k] = acc;
    }
    const bool lead0 = blockIdx.x == 0 && threadIdx.x == 0;
    double c1 = 0.0, c2 = 0.0;   // MODE 1: alpha; MODE 2: omega, beta
    if (MODE == 1) {
        double pa[2];
        if (tot & FIN_A) {
            pa[0] = P[TL_TOT];
            pa[1] = P[TL_TOT + 1];
        } else {
            sum_partials_wide<2>(P + TL_PA, Gt, pa, kMaxTileWg);
        }
        const double rvv = pa[0], rr = pa[1];
        const KScal sp = st->ring[ph];
        const KScal spp = st->ring[(ph + KM_RING - 1) % KM_RING];
        const int i = sp.iter;
        const double rho = i == 0 ? rr : sp.rho;
        const double r0 = i == 0 ? rr : st->r0;
        int status = -1;
        if (!fin(rr)) status = KM_NONFINITE;
        else if (i == 0 && rr == 0.0) status = KM_OK;
        else if (i > 0 && sqrt(rr / r0) < st->eps) status = KM_OK;   // bicgsolver.py:130
        else if (i > 0 && sp.omega == 0.0) status = KM_BREAKDOWN;
        else if (i >= st->max_iter) status = KM_MAXITER;
        else if (!fin(rho)) status = KM_NONFINITE;
        else if (rho == 0.0) status = KM_BREAKDOWN;
        else if (!fin(rvv)) status = KM_NONFINITE;
        else if (rvv == 0.0) status = KM_BREAKDOWN;
        const bool better = rr < 0.81 * spp.best;
        if (status < 0 && !better && i - spp.best_it >= kStallWindow) status = KM_STALL;
        if (lead0 && i == 0) st->r0 = r0;
        if (status >= 0) {
            if (lead0) finish_state(st, status, i, rr);
            return;
        }
        c1 = rho / rvv;   // alpha, bicgsolver.py:108
        if (lead0) {
            st->ring[ph].alpha = c1;
            if (i == 0) st->ring[ph].rho = rho;
            st->ring[ph].best = better ? rr : spp.best;
            st->ring[ph].best_it = better ? i : spp.best_it;
            st->rTr = rr;
            st->iter = i;
        }
    } else {
        double pt[4];
        if (tot & FIN_T) {
#pragma unroll
            for (int q = 0; q < 4; ++q) pt[q] = P[TL_TOT + 2 + q];
        } else {
            sum_partials_wide<4>(P + TL_PT, Gt, pt, kMaxTileWg);
        }
        const KScal sp = st->ring[ph];
        const double omega = pt[1] != 0.0 ? pt[0] / pt[1] : 0.0;                  // bicgsolver.py:118
        const double rho = pt[2] - omega * pt[3];                                 // <r_i+1, r~> (bicgsolver.py:95)
        const double beta = omega != 0.0 ? (rho / sp.rho) * (sp.alpha / omega) : 0.0;   // bicgsolver.py:102
        c1 = omega;
        c2 = beta;
        if (lead0) {
            KScal& nxt = st->ring[(ph + 1) % KM_RING];
            nxt.rho = rho;
            nxt.omega = omega;
            nxt.beta = beta;
            nxt.iter = sp.iter + 1;
        }
    }
    __syncthreads();
    if (!live) return;
    // summed in k order per lane
    double s = 0.0;
#pragma unroll
    for (int u = 0; u < kCoarseCh; ++u)
        if (lane + 64 * u < m) s += av[u] * rs_[lane + 64 * u];
    for (int k0 = lane + 64 * kCoarseCh; k0 < m; k0 += 64 * kCoarseCh) {
#pragma unroll
        for (int u = 0; u < kCoarseCh; ++u) av[u] = k0 + 64 * u < m ? a[k0 + 64 * u] : 0.0;
#pragma unroll
        for (int u = 0; u < kCoarseCh; ++u)
            if (k0 + 64 * u < m) s += av[u] * rs_[k0 + 64 * u];
    }
    s = wave_sum(s);
    if (lane == 0) {
        if (MODE == 1) {
            Y[2 * m + row] = s;                  // Y(v)
            Y[3 * m + row] = y0 - c1 * s;        // Y(s)
        } else {
            const double yr = y3 - c1 * s;       // Y(r)
            Y[row] = yr;
            Y[m + row] = yr + c2 * (y1 - c1 * y2);   // Y(p)
        }
    }
}

}  // namespace

namespace kmcoarse {

// Tile shape of the smoothing launches for the coarse level's system in T:
// among the shapes with the fewest restriction slots, on the first call for a
// (shape, precision) each runs 3 + 16 back-to-back K_PV + K_ST pairs on
// scratch vectors (ws: 10n values of T, overwritten; src is read) with a
// scratch state, and the fastest is kept.
template <typename T>
int tuned_tile(km_ctx* ctx, const T* coefT, const T* src, T* ws, int* out) {
    km_coarse* C = ctx->coarse;
    const auto key = std::make_tuple(C->NX, C->NY, (int)sizeof(T));
    auto it = C->tiles.find(key);
    if (it != C->tiles.end()) {
        *out = it->second;
        return KM_OK;
    }
    int fewest = 1 << 30;
    for (int w = 0; w < kNumTiles; ++w) fewest = std::min(fewest, tile_slots(C, w));
    KM_REQUIRE(ctx, fewest <= kMaxVS, "aggregates of %d cells overlap %d tiles", C->B, fewest);
    const int64_t n = (int64_t)C->NX * C->NY;
    KState* tst = &ctx->st[5];
    hipLaunchKernelGGL(k_tl_tune_state, dim3(1), dim3(1), 0, ctx->stream, tst);
    KM_CHECK_LAUNCH(ctx);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    KM_HIP(ctx, hipEventCreate(&e0));
    KM_HIP(ctx, hipEventCreate(&e1));
    int best = -1;
    float best_ms = 0.f;
    for (int w = 0; w < kNumTiles; ++w) {
        if (tile_slots(C, w) != fewest) continue;
        TlArgs<T> a{};
        a.c = coefT;
        a.mask = C->mask;
        a.f0 = ws;
        a.f1 = ws + n;
        a.f2 = ws + 2 * n;
        a.f3 = ws + 3 * n;
        a.rt = src;
        a.x = ws + 4 * n;
        a.r = ws + 5 * n;
        a.sh = ws + 6 * n;
        a.src = ws + 7 * n;
        a.z = ws + 8 * n;
        a.y = ws + 9 * n;
        a.yc = C->yc;
        a.rcs = C->frc;
        a.st = tst;
        a.P = ctx->partials + TL_PA;
        a.ph = 0;
        TlArgs<T> b = a;   // K_ST: s = r - alpha v, s^, t
        b.f0 = ws + 5 * n;
        b.f1 = ws + 9 * n;
        b.src = ws;
        b.z = ws + 6 * n;
        b.y = ws + n;
        b.rcs = C->frc + (size_t)kMaxVS * C->m;
        b.P = ctx->partials + TL_PT;
        const CGeom cg = c_geom(C, w);
        auto once = [&]() {
            launch_tile<T, 1>(w, ctx->stream, C->NX, C->NY, cg, a);
            launch_tile<T, 2>(w, ctx->stream, C->NX, C->NY, cg, b);
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
        if (best < 0 || ms < best_ms) {
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
    TlArgs<T> a{};
    a.c = coefT;
    a.mask = C->mask;
    a.f0 = r;
    a.z = z;
    a.yc = C->yc;
    a.st = nullptr;
    launch_tile<T, 0>(0, s, C->NX, C->NY, c_geom(C, 0), a);
    KM_CHECK_LAUNCH(ctx);
    return KM_OK;
}

namespace {

// The vectors and launch arguments of one fused solve (ws: 10 n values of T).
template <typename T>
struct Fused {
    int NX, NY, tile, Gt, Gc, S, fin;
    CGeom cg;
    T *r, *rt, *p[2], *v[2], *s, *t, *ph, *sh;
    T* x;
    const T* coefT;
    double *Y, *rcv, *rct;
    km_coarse* C;
    KState* st;
    double* P;

    Fused(km_ctx* ctx, int NX_, int NY_, const T* coef, T* ws, T* xd, int tile_)
        : NX(NX_), NY(NY_), tile(tile_), x(xd), coefT(coef), C(ctx->coarse), st(&ctx->st[4]), P(ctx->partials) {
        const int64_t n = (int64_t)NX * NY;
        r = ws;
        rt = ws + n;
        p[0] = ws + 2 * n;
        p[1] = ws + 3 * n;
        v[0] = ws + 4 * n;
        v[1] = ws + 5 * n;
        s = ws + 6 * n;
        t = ws + 7 * n;
        ph = ws + 8 * n;
        sh = ws + 9 * n;
        Gt = tile_workgroups(tile, NX, NY);
        Gc = (C->m + 3) / 4;
        S = tile_slots(C, tile);
        cg = c_geom(C, tile);
        fin = total_flags(C, Gt, Gc);
        Y = C->fy;
        rcv = C->frc;
        rct = C->frc + (size_t)kMaxVS * C->m;
    }

    void pv(hipStream_t q, int phs) const {
        const int par = phs & 1;
        TlArgs<T> a{};
        a.c = coefT;
        a.mask = C->mask;
        a.f0 = s;
        a.f1 = t;
        a.f2 = p[par ^ 1];
        a.f3 = v[par ^ 1];
        a.rt = rt;
        a.x = x;
        a.r = r;
        a.sh = sh;
        a.src = p[par];
        a.z = ph;
        a.y = v[par];
        a.yc = Y + C->m;   // Y(p)
        a.rcs = rcv;
        a.st = st;
        a.P = P + TL_PA;
        a.ph = phs;
        launch_tile<T, 1>(tile, q, NX, NY, cg, a);
    }
    void st_(hipStream_t q, int phs) const {
        TlArgs<T> a{};
        a.c = coefT;
        a.mask = C->mask;
        a.f0 = r;
        a.f1 = v[phs & 1];
        a.rt = rt;
        a.src = s;
        a.z = sh;
        a.y = t;
        a.yc = Y + 3 * C->m;   // Y(s)
        a.rcs = rct;
        a.st = st;
        a.P = P + TL_PT;
        a.ph = phs;
        launch_tile<T, 2>(tile, q, NX, NY, cg, a);
    }
    void total_a(hipStream_t q) const {
        if (fin & FIN_A)
            hipLaunchKernelGGL(k_tl_total<2>, dim3(1), dim3(KM_BLOCK), 0, q, (const double*)(P + TL_PA), Gt,
                               kMaxTileWg, P + TL_TOT, (const KState*)st);
    }
    void total_t(hipStream_t q) const {
        if (fin & FIN_T)
            hipLaunchKernelGGL(k_tl_total<4>, dim3(1), dim3(KM_BLOCK), 0, q, (const double*)(P + TL_PT), Gt,
                               kMaxTileWg, P + TL_TOT + 2, (const KState*)st);
    }
    template <int MODE>
    void coarse(hipStream_t q, int phs) const {
        hipLaunchKernelGGL(k_tl_coarse<MODE>, dim3(Gc), dim3(KM_BLOCK), (size_t)C->m * sizeof(double), q,
                           C->m, S, Gt, fin, phs, (const double*)C->inv, (const double*)(MODE == 1 ? rcv : rct), Y,
                           st, (const double*)P);
    }
    // one inner iteration
    void iteration(hipStream_t q, int phs) const {
        pv(q, phs);
        total_a(q);
        coarse<1>(q, phs);
        st_(q, phs);
        total_t(q);
        coarse<2>(q, phs);
    }
};

// x = 0, r = r~ = rhs, the state, Y(r) = Y(p) = A_c^-1 P^T rhs, the slots cleared
template <typename T>
int fused_start(km_ctx* ctx, const Fused<T>& f, const T* rhs, double eps, int max_iter) {
    km_coarse* C = f.C;
    const int64_t n = (int64_t)f.NX * f.NY;
    const Flat fl = make_flat(n);
    hipLaunchKernelGGL(k_tl_init<T>, dim3(fl.G), dim3(KM_BLOCK), 0, ctx->stream, fl, rhs, f.x, f.r, f.rt, eps,
                       max_iter, f.st);
    KM_CHECK_LAUNCH(ctx);
    if (int rc = restrict_solve<T>(ctx, rhs, nullptr, ctx->stream)) return rc;
    const int64_t cells = std::max<int64_t>(C->m, (int64_t)f.S * C->m);
    hipLaunchKernelGGL(k_tl_cinit, dim3((unsigned)((cells + KM_BLOCK - 1) / KM_BLOCK)), dim3(KM_BLOCK), 0,
                       ctx->stream, C->m, f.S, (const double*)C->yc, f.Y, f.rcv, f.rct);
    KM_CHECK_LAUNCH(ctx);
    return KM_OK;
}

}  // namespace

template <typename T>
int fused_inner(km_ctx* ctx, int NX, int NY, const T* coefT, const T* rhs, T* xd, T* ws, double eps, int max_iter,
                bool use_graph, KState* out) {
    km_coarse* C = ctx->coarse;
    KM_REQUIRE(ctx, C && C->arena && C->NX == NX && C->NY == NY, "coarse level not set up for this system");
    int tile = 0;
    if (int rc = tuned_tile<T>(ctx, coefT, rhs, ws, &tile)) return rc;
    const Fused<T> f(ctx, NX, NY, coefT, ws, xd, tile);
    if (int rc = fused_start<T>(ctx, f, rhs, eps, max_iter)) return rc;
    auto enqueue = [&f](hipStream_t q, int phs) { f.iteration(q, phs); };
    GraphKey key;
    key.v = {9, sizeof(T), (uintptr_t)NX, (uintptr_t)NY, (uintptr_t)coefT, (uintptr_t)rhs, (uintptr_t)xd,
             (uintptr_t)ws, (uintptr_t)f.st, (uintptr_t)C, (uintptr_t)C->arena, (uintptr_t)C->m, (uintptr_t)C->B,
             (uintptr_t)tile, (uintptr_t)f.fin};
    return run_chunks(ctx, f.st, key, use_graph, kChunk, max_iter, enqueue, out);
}

// Average device time of one launch of a fused two-level kernel (which: 4 K_PV,
// 5 K_CV, 6 K_ST, 7 K_CT) inside FP64 iterations on (coef, b) with the stopping
// test disabled; events around every launch of that kernel, or (b2b) around
// reps identical launches after eight warm iterations.
int time_kernel(km_ctx* ctx, int which, int NX, int NY, const double* coef, const double* b, int reps, bool b2b,
                double* avg_us) {
    KM_REQUIRE(ctx, which >= 4 && which <= 7, "unknown kernel id %d", which);
    if (int rc = setup(ctx, NX, NY, coef, ctx->stream)) return rc;
    const int64_t n = (int64_t)NX * NY;
    if (int rc = km_ensure_workspace(ctx, (size_t)11 * n * sizeof(double))) return rc;
    double* ws = static_cast<double*>(km_ws_vectors(ctx));
    double* x = ws + 10 * n;
    int tile = 0;
    if (int rc = tuned_tile<double>(ctx, coef, b, ws, &tile)) return rc;
    const Fused<double> f(ctx, NX, NY, coef, ws, x, tile);
    if (int rc = fused_start<double>(ctx, f, b, 0.0, 1 << 30)) return rc;
    std::vector<hipEvent_t> ev(2 * (size_t)reps, nullptr);
    for (auto& e : ev) KM_HIP(ctx, hipEventCreate(&e));
    auto cleanup = [&]() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
    };
    hipStream_t q = ctx->stream;
    auto launch = [&](int k, int phs) {
        switch (k) {
        case 4: f.pv(q, phs); break;
        case 5: f.coarse<1>(q, phs); break;
        case 6: f.st_(q, phs); break;
        default: f.coarse<2>(q, phs);
        }
    };
    // one() adds the totals that follow the tile launches
    auto one = [&](int k, int phs, hipEvent_t e0, hipEvent_t e1) {
        if (e0) (void)hipEventRecord(e0, q);
        launch(k, phs);
        if (e1) (void)hipEventRecord(e1, q);
        if (k == 4) f.total_a(q);
        if (k == 6) f.total_t(q);
    };
    const int warm = 8;
    for (int it = 0; it < warm; ++it)
        for (int k = 4; k <= 7; ++k) one(k, it % KM_RING, nullptr, nullptr);
    if (b2b) {
        // K_PV re-runs the scalars and updates of the next phase, the others
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
    KM_HIP(ctx, hipMemcpy(&ctx->h_st[3], f.st, sizeof(KState), hipMemcpyDeviceToHost));
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
