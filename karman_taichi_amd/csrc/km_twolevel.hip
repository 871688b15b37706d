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
//   K_PV  2-D tiles with a (2 kSweeps + 1)-cell halo: p^ = M(p) with z1 from
//         Y(p), v = A p^; partials <r~,v>; slot partials of P^T v
//   K_CV  stopping test on <r,r>, alpha = rho / <r~,v>;
//         Y(v) = A_c^-1 P^T v, Y(s) = Y(r) - alpha Y(v)
//   K_ST  tiles: s = r - alpha v on the tile and its halo; s^ = M(s) with z1
//         from Y(s), t = A s^; partials <t,s>, <t,t>, <s,r~>, <t,r~>; P^T t
//   K_CT  omega, rho, beta of the next iteration.  Coarse workgroups:
//         Y(t) = A_c^-1 P^T t, Y(r) = Y(s) - omega Y(t),
//         Y(p) = Y(r) + beta (Y(p) - omega Y(v)).  Flat workgroups (no halo):
//         x += alpha p^ + omega s^, r = s - omega t, p = r + beta (p - omega v);
//         partials <r,r>
// The flat updates run beside the latency-bound coarse rows, and K_PV reads
// only p on its halo.

// Tiles are aligned with the aggregates: tile (a, b) owns the interior rows
// [1 + a TI, 1 + (a + 1) TI) and columns [1 + b TJ, ...), clipped, and the
// boundary row or column next to it at the edges of the system (one cell
// into the halo, where M(.) is still exact).  An aggregate's restriction
// then comes from the few tiles that overlap it, each writing its partial sum
// into its own slot: slot (a - I B / TI) SJ + (b - J B / TJ) of aggregate
// (I, J).  The coarse launches sum the slots in slot order.
#include "km_coarse.h"

#include <algorithm>
#include <type_traits>
#include <cmath>
#include <vector>

namespace {

constexpr int kMaxTileWg = 8192;                 // smoothing workgroups (one tile each up to 8192 tiles)
constexpr int kMaxFlatWg = 8192;                 // flat workgroups of K_CT and the start
constexpr int TL_PV = 0;                         // [kMaxTileWg] <r~,v> (K_PV)
constexpr int TL_PR = kMaxTileWg;                // [kMaxFlatWg] <r,r> (K_CT's flat workgroups, the start)
constexpr int TL_PT = kMaxTileWg + kMaxFlatWg;   // [4][kMaxTileWg] <t,s>, <t,t>, <s,r~>, <t,r~> (K_ST)
constexpr int TL_TOT = TL_PT + 4 * kMaxTileWg;   // totals: 0 <r~,v>, 1..4 the K_ST planes, 5 <r,r>
constexpr int kPairPartials = 1 << 16;           // the second system of a pair solve: its partials' offset
static_assert(kPairPartials >= TL_TOT + 8 && 2 * kPairPartials <= KM_MAX_PARTIALS, "pair partials");
// The coarse (and flat) workgroups sum the tile partials; when partials x
// consumer workgroups exceeds this, one workgroup sums them first
// (k_tl_total) and the consumers read the totals.
constexpr int64_t kTotalWork = 2000000;
constexpr int FIN_A = 1, FIN_T = 2, FIN_R = 4;
constexpr int kStallWindow = 300;        // as the unfused flexible solve
constexpr int64_t kTileCellsMax = (int64_t)1 << 29;   // the tile kernels' 32-bit byte offsets
constexpr int kFlatPer = 8;              // cells per thread of a flat workgroup, loads in flight
// Tile shapes of the smoothing launches (rows x columns, columns contiguous);
// the fastest for a system is chosen on its first solve.
constexpr int kNumTiles = 9;
constexpr int kTileI[kNumTiles] = {32, 16, 16, 8, 8, 16, 16, 32, 16};
constexpr int kTileJ[kNumTiles] = {32, 64, 32, 64, 128, 32, 32, 32, 64};
// minimum waves per SIMD the compiler must allow (a VGPR budget): 16 x 32 also
// fits 128 VGPRs, so a fourth wave per SIMD can hide latency
constexpr int kTileW[kNumTiles] = {1, 1, 3, 1, 1, 4, 4, 2, 2};
// threads per workgroup: with 512 a thread holds half as many frame cells, so
// each one's chain of loads and half-sweep updates is shorter
constexpr int kTileT[kNumTiles] = {256, 256, 256, 256, 256, 256, 512, 512, 512};

__device__ __forceinline__ bool fin(double v) { return v == v && v - v == 0.0; }

// 1 / d for finite non-zero d: the hardware reciprocal estimate refined by two
// Newton steps x' = x + x (1 - d x) (each squares the relative error; the
// result is within an ulp of the rounded quotient)
__device__ __forceinline__ double recip(double d) {
    double x = __builtin_amdgcn_rcp(d);
    x = __builtin_fma(x, __builtin_fma(-d, x, 1.0), x);
    x = __builtin_fma(x, __builtin_fma(-d, x, 1.0), x);
    return x;
}
__device__ __forceinline__ float recip(float d) { return 1.0f / d; }

// base[byte offset]: a uniform base and a 32-bit per-lane offset (the
// scalar-base addressing form of the vector loads)
template <typename T>
__device__ __forceinline__ T ld_at(const T* base, uint32_t off) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off);
}

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

// Flat workgroups: contiguous chunks of kFlatPer KM_BLOCK cells (more when
// that would need over kMaxFlatWg workgroups).
struct FGeom {
    int64_t n;
    int chunk, G;
};

FGeom f_geom(int64_t n) {
    FGeom f;
    f.n = n;
    int64_t chunk = (int64_t)kFlatPer * KM_BLOCK;
    if ((n + chunk - 1) / chunk > kMaxFlatWg) chunk = ((n + kMaxFlatWg - 1) / kMaxFlatWg + KM_BLOCK - 1) / KM_BLOCK * KM_BLOCK;
    f.chunk = (int)chunk;
    f.G = (int)((n + chunk - 1) / chunk);
    return f;
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

// Launch arguments of the tile kernels.  Frame inputs f0, f1 are read on the
// tile and its 2 kSweeps-cell halo:
//   MODE 0  z = M(f0)
//   MODE 1  f0 = p: z = p^, y = v
//   MODE 2  f0 = r, f1 = v: src = s, z = s^, y = t
template <typename T>
struct TlArgs {
    const T* c;              // 5 coefficient planes
    const double* cd;        // the same in FP64 (the C64 kernels of FP32 vectors)
    const uint8_t* mask;     // free cells
    const T* f0;
    const T* f1;
    const T* rt;             // r~ (MODE 1, 2)
    T* src;                  // MODE 2: s out
    T* z;                    // M(.) out
    T* y;                    // A z out (MODE 1, 2)
    const double* yc;        // coarse vector prolongated as z1
    double* rcs;             // slot partials of P^T y
    const KState* st;        // null: no state (tuning, MODE 0)
    double* P;               // partial planes of this launch
    int ph;
    T om;                    // relaxation of the red-black sweeps
    // K_ST with its own scalars (a solve without a coarse level): the
    // partials of <r~,v> (Gin[0] workgroups) and <r,r> (Gin[1]) at Pin, or
    // their totals (tot: FIN_A and FIN_R), and the state it updates
    const double* Pin;
    int Gin[2];
    int tot;
    KState* stw;
    int* flag;               // the host's progress words (cv_scalars)
};

// x = 0, r = r~ = p = rhs, partials of <r,r> on the flat workgroups, the state
template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_tl_init(FGeom fl, const T* __restrict__ rhs, T* __restrict__ x, T* __restrict__ r, T* __restrict__ rt,
          T* __restrict__ p, double eps, int max_iter, KState* __restrict__ st, double* __restrict__ P) {
    __shared__ double sh[8];
    const int64_t lo = (int64_t)blockIdx.x * fl.chunk;
    const int64_t hi = lo + fl.chunk < fl.n ? lo + fl.chunk : fl.n;
    double rr = 0.0;
    for (int64_t k = lo + threadIdx.x; k < hi; k += KM_BLOCK) {
        const T b = rhs[k];
        x[k] = T(0);
        r[k] = b;
        rt[k] = b;
        p[k] = b;
        rr += (double)b * (double)b;
    }
    rr = block_sum(rr, sh);
    if (threadIdx.x == 0) P[TL_PR + blockIdx.x] = rr;
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

// The stopping test of iteration ph on <r,r> and alpha = rho / <r~,v>
// (bicgsolver.py:108,130) with the stall test of the flexible solve; every
// workgroup derives the same, lead0 (one thread of the system's first
// workgroup) writes the state and, to the host's words `flag` (may be null),
// the iteration reached and ``done``.
// false: the solve has ended (``done`` set).
__device__ __forceinline__ bool cv_scalars(KState* __restrict__ st, int ph, double rvv, double rr, double* alpha,
                                           int* flag, bool lead0) {
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
        if (lead0) {
            finish_state(st, status, i, rr);
            if (flag) {
                // the host's copy of the final state, then done (release: after the rest)
                __hip_atomic_store(flag + 1, i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                __hip_atomic_store(flag + 2, status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                __hip_atomic_store(reinterpret_cast<double*>(flag + 4), r0, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_SYSTEM);
                __hip_atomic_store(reinterpret_cast<double*>(flag + 6), rr, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_SYSTEM);
                __hip_atomic_store(flag, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
            }
        }
        return false;
    }
    if (lead0 && flag) __hip_atomic_store(flag + 1, i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    *alpha = rho / rvv;   // bicgsolver.py:108
    if (lead0) {
        st->ring[ph].alpha = *alpha;
        if (i == 0) st->ring[ph].rho = rho;
        st->ring[ph].best = better ? rr : spp.best;
        st->ring[ph].best_it = better ? i : spp.best_it;
        st->rTr = rr;
        st->iter = i;
    }
    return true;
}

// K_PV (MODE 1) / K_ST (MODE 2) / the standalone smoother (MODE 0).  The
// coefficients and the frame input of the tile's NS-cell halo frame stay in
// registers (NS = 2 SW half-sweeps); z lives in LDS on the (NS + 1)-cell halo
// frame and is updated in place: a red half-sweep changes only red cells
// ((i + j) even) and reads only black ones, and the other way round.  The
// arithmetic per cell is z + (omega / aP) (src - A z), neighbours outside the
// system read as zero, rows with aP = 0 left unchanged.  Half-sweep s covers
// the halo-(NS - s) frame, so after the last one z is exact one cell into the
// halo, which covers the boundary cells the edge tiles own.  Without a coarse
// vector (yc null) z starts from zero; without slots (rcs null) nothing is
// restricted.  SELF (MODE 2): the launch derives alpha and the stopping test
// itself (cv_scalars).
//
// Each thread holds pairs of neighbouring cells along j, (a, b0) and
// (a, b0 + 1) with b0 even: one red and one black cell, kept in registers by
// colour, so every lane has a cell to update in every half-sweep.  Pair slot
// q holds pair t + q NT of the frame, the tile's own block first (NPI
// pairs), then the ring of halo cells (the top and bottom bands, then the
// side bands row by row); the output pass (z, A z, partial sums) covers the
// slots holding the block, and the others only in the tiles at the system's
// edges (whose boundary row or column lies in the ring).
//
// SYM (a symmetric system, aW(i, j) = aE(i - 1, j) and aS(i, j) = aN(i, j - 1)
// bit for bit, as the assembled p' systems are): only the aP, aE and aN
// planes are read; each cell's aE and aN go to LDS (zero outside the system,
// where the halo's links are zero) with one more row above the frame and
// one more column left of it, and aW, aS are read back from there, so the
// arithmetic and its result are those of the five-plane kernel.
//
// C64 (FP32 vectors, T = float): the coefficients come from the FP64 planes
// (g.cd) and the smoother and y = A z run in FP64, z rounded to T in LDS.
// A p' row's five terms cancel to about lambda_min / aP of aP z on smooth z,
// so FP32 coefficients or FP32 sums err there by ~6e-8 times the condition
// number (order nx^2): on the largest channels that stalls the FP64
// refinement around the FP32 solve.  With FP64 coefficients and sums, y is A
// applied to the stored (rounded) z to FP64 accuracy, and the FP32 iteration
// solves the FP64 system.
template <typename T, int TI, int TJ, int MODE, int SW, bool SELF, int NT, bool SYM = false, bool C64 = false>
__device__ __forceinline__ void tile_body(const ATile& tl, const CGeom& cg, const TlArgs<T>& g, int L) {
    // The smoother and the matvec are issue-bound: fused multiply-adds (one
    // rounding instead of two per term) and the reciprocal of aP from the
    // hardware estimate with two Newton steps instead of a full division.
#pragma clang fp contract(fast)
    using CT = typename std::conditional<C64, double, T>::type;   // coefficient and arithmetic type
    const CT* __restrict__ cbase;
    if constexpr (C64) cbase = g.cd;
    else cbase = g.c;
    static_assert((NT == 256 || NT == 512) && (!SELF || NT == KM_BLOCK), "4 or 8 waves; SELF sums as KM_BLOCK");
    constexpr int NW = NT / 64;                                     // waves per workgroup
    constexpr int NS = 2 * SW;                                      // half-sweeps = halo
    constexpr int WZ = TJ + 2 * (NS + 1), HZ = TI + 2 * (NS + 1);   // z frame: halo NS + 1
    constexpr int WR = TJ + 2 * NS, HR = TI + 2 * NS;               // register frame: halo NS
    static_assert(TJ % 2 == 0 && NS % 2 == 0, "cell pairs along j");
    constexpr int NPI = TI * TJ / 2;          // pairs of the tile's own block
    constexpr int NP = HR * WR / 2;           // pairs of the register frame
    constexpr int QI = (NPI + NT - 1) / NT;   // slots holding the block's pairs
    constexpr int Q = (NP + NT - 1) / NT;     // pair slots
    constexpr int BAND = NS * (WR / 2);       // pairs of the top (and of the bottom) band
    __shared__ T za[HZ * WZ];
    __shared__ double yb[MODE > 0 ? TI * TJ : 1];   // y on the free interior cells (tile-local)
    // SYM: aE of frame rows -1 .. HR - 1 (row a at a + 1), aN of frame columns
    // -1 .. WR - 1 (column b at b + 1)
    __shared__ CT aes[SYM ? (HR + 1) * WR : 1];
    __shared__ CT ans[SYM ? HR * (WR + 1) : 1];
    static_assert(!SYM || WR + HR <= NT, "one extra aE / aN cell per thread");
    __shared__ double sh[4 * NW];
    const KState* st = g.st;
    if (st && st->done) return;
    const int NX = tl.NX, NY = tl.NY, nx = NX - 2, ny = NY - 2;
    const int64_t n = (int64_t)NX * NY;
    const T om = g.om;
    double alpha = MODE == 2 && st && !SELF ? st->ring[g.ph].alpha : 0.0;
    if (SELF) {
        double rvv, rr;
        if ((g.tot & (FIN_A | FIN_R)) == (FIN_A | FIN_R)) {
            rvv = g.Pin[TL_TOT];
            rr = g.Pin[TL_TOT + 5];
        } else {
            const double* planes[2] = {g.Pin + TL_PV, g.Pin + TL_PR};
            double t2[2];
            sum_partials_each<2>(planes, g.Gin, t2);
            rvv = t2[0];
            rr = t2[1];
        }
        if (!cv_scalars(g.stw, g.ph, rvv, rr, &alpha, g.flag, L == 0 && threadIdx.x == 0)) return;
    }
    const T fa = (T)alpha;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    const int t0 = L * tl.tpb, t1 = t0 + tl.tpb < tl.T ? t0 + tl.tpb : tl.T;
    for (int tile = t0; tile < t1; ++tile) {
        const int ta = tile / tl.tiles_j, tb = tile - ta * tl.tiles_j;
        const int ti0 = 1 + ta * TI, tj0 = 1 + tb * TJ;
        // cells this tile writes: its interior block plus the boundary row or
        // column next to it at the edges of the system
        const bool edge = ta == 0 || ta == tl.tiles_i - 1 || tb == 0 || tb == tl.tiles_j - 1;
        const int olo = ta == 0 ? 0 : ti0, ohi = ta == tl.tiles_i - 1 ? NX : ti0 + TI;
        const int jlo = tb == 0 ? 0 : tj0, jhi = tb == tl.tiles_j - 1 ? NY : tj0 + TJ;
        const int par0 = (ti0 + tj0) & 1;   // (i + j) parity of frame cell (0, 0)

        // every load of the tile is issued before the first store (the
        // argument pointers may alias as far as the compiler knows): the
        // register frame, the tile's own cells and the z frame.  Index c:
        // the pair's red (0) and black (1) cell.
        constexpr int NF = MODE == 2 ? 2 : 1;   // frame inputs per cell
        constexpr int QZ = (HZ * WZ + NT - 1) / NT;
        // Every load reads a valid address (cells outside the system read
        // cell 0) and sits in one basic block with the others; what it
        // contributes is selected only after all of them are issued, so the
        // tile's loads are in flight together instead of one round trip per
        // conditional load.
        CT cr[Q][2][5], wd[Q][2];
        T sr[Q][2], rq[Q][2];
        T fv[Q][2][NF];
        uint8_t mk[Q][2];
        int64_t kq[Q][2];
        int pa[Q], pb[Q][2];       // frame row of the pair, frame column of each colour's cell
        uint32_t inb = 0, oub = 0, ib = 0, fm = 0;   // bit 2 q + c: in the system; written; interior; free
        double yq[QZ];
        uint8_t mq[QZ];
        // the free-cell mask where there are slots to restrict into (ib is
        // set only then), else the coefficient array (in bounds, unused)
        const uint8_t* mkp = g.rcs ? g.mask : reinterpret_cast<const uint8_t*>(g.c);
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int P = threadIdx.x + q * NT;
            int a, b0;
            if (P < NPI) {
                a = NS + P / (TJ / 2);
                b0 = NS + 2 * (P - (P / (TJ / 2)) * (TJ / 2));
            } else {
                const int r = P - NPI;
                if (r < BAND) {
                    a = r / (WR / 2);
                    b0 = 2 * (r - a * (WR / 2));
                } else if (r < 2 * BAND) {
                    const int u = r - BAND;
                    a = NS + TI + u / (WR / 2);
                    b0 = 2 * (u - (u / (WR / 2)) * (WR / 2));
                } else {
                    const int u = r - 2 * BAND, e = u - (u / NS) * NS;
                    a = NS + u / NS;
                    b0 = e < NS / 2 ? 2 * e : TJ + NS + 2 * (e - NS / 2);
                }
            }
            const bool valid = P < NP;
            const int par = (par0 + a) & 1;   // parity of (a, b0); colour c sits at b0 + (par ^ c)
            pa[q] = a;
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const int b = b0 + (par ^ c);
                pb[q][c] = b;
                const int i = ti0 - NS + a, j = tj0 - NS + b;
                const bool inc = valid && i >= 0 && i < NX && j >= 0 && j < NY;
                const bool ouc = inc && i >= olo && i < ohi && j >= jlo && j < jhi;
                const int64_t k = inc ? (int64_t)i * NY + j : 0;
                kq[q][c] = k;
                const uint32_t bit = 1u << (2 * q + c);
                if (inc) inb |= bit;
                if (ouc) oub |= bit;
                if (MODE > 0 && g.rcs && ouc && i >= 1 && i <= nx && j >= 1 && j <= ny) ib |= bit;
                // 32-bit byte offsets from uniform bases (systems below 2^29
                // cells, kTileCellsMax: checked by the coarse set-up and the
                // smoothed solves)
                const uint32_t ko = (uint32_t)k * (uint32_t)sizeof(T);
                const uint32_t kc = (uint32_t)k * (uint32_t)sizeof(CT);
#pragma unroll
                for (int e = 0; e < 5; ++e)
                    if (!SYM || e == 0 || e == 2 || e == 3) cr[q][c][e] = ld_at(cbase + e * n, kc);
                const T* fp[2] = {g.f0, g.f1};
#pragma unroll
                for (int e = 0; e < NF; ++e) fv[q][c][e] = ld_at(fp[e], ko);
                if (MODE > 0) rq[q][c] = ld_at(g.rt, ko);
                if (MODE > 0) mk[q][c] = mkp[(uint32_t)k];
            }
        }
        // SYM: aE of the row above the frame (threads 0 .. WR - 1) and aN of
        // the column left of it (threads WR .. WR + HR - 1)
        CT xtra = CT(0);
        bool xin = false;
        if (SYM) {
            const int t = threadIdx.x;
            int i = -1, j = -1, e = 2;
            if (t < WR) {
                i = ti0 - NS - 1;
                j = tj0 - NS + t;
            } else if (t < WR + HR) {
                i = ti0 - NS + (t - WR);
                j = tj0 - NS - 1;
                e = 3;
            }
            xin = i >= 0 && i < NX && j >= 0 && j < NY;
            const uint32_t kc = xin ? ((uint32_t)i * (uint32_t)NY + (uint32_t)j) * (uint32_t)sizeof(CT) : 0u;
            xtra = ld_at(cbase + e * n, kc);
        }
        // z1 = P yc on the z frame (zero off the free cells and outside the
        // system).  Without a coarse vector the loads read the coefficient
        // array instead (in bounds: index < NY doubles, < NX NY bytes) and
        // nothing they read is used.
        uint32_t zin = 0;   // bit q: z-frame slot q lies in the system
        {
            const double* ycp = g.yc ? g.yc : reinterpret_cast<const double*>(g.c);
            const uint8_t* mzp = g.yc ? g.mask : reinterpret_cast<const uint8_t*>(g.c);
            const int B = g.yc ? cg.B : 1, mJ = g.yc ? cg.mJ : 0;
#pragma unroll
            for (int q = 0; q < QZ; ++q) {
                const int idx = threadIdx.x + q * NT;
                const int a = idx / WZ, b = idx - (idx / WZ) * WZ;
                const int i = ti0 - NS - 1 + a, j = tj0 - NS - 1 + b;
                const bool ing = idx < HZ * WZ && i >= 0 && i < NX && j >= 0 && j < NY;
                const int ic = i < 1 ? 0 : (i > nx ? nx - 1 : i - 1), jc = j < 1 ? 0 : (j > ny ? ny - 1 : j - 1);
                if (ing) zin |= 1u << q;
                yq[q] = ycp[(ic / B) * mJ + jc / B];
                mq[q] = mzp[ing ? (int64_t)i * NY + j : 0];
            }
        }
        // what the loads contribute, now that all of them are issued.  A cell
        // outside the system read cell 0: its values are never used (the
        // half-sweeps skip cells without their inb bit, the output pass
        // covers oub cells, a subset), and its z stays the zero of za.
#pragma unroll
        for (int q = 0; q < Q; ++q)
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const uint32_t bit = 1u << (2 * q + c);
                if (MODE > 0 && (ib & bit) && mk[q][c]) fm |= bit;
                if (SYM && threadIdx.x + q * NT < NP) {
                    const bool inc = inb & bit;
                    aes[(pa[q] + 1) * WR + pb[q][c]] = inc ? cr[q][c][2] : CT(0);
                    ans[pa[q] * (WR + 1) + pb[q][c] + 1] = inc ? cr[q][c][3] : CT(0);
                }
            }
        if (SYM) {
            const int t = threadIdx.x;
            if (t < WR) aes[t] = xin ? xtra : CT(0);
            else if (t < WR + HR) ans[(t - WR) * (WR + 1)] = xin ? xtra : CT(0);
        }
        // the source of M(.) on the frame (MODE 2: s, written on the tile's cells)
#pragma unroll
        for (int q = 0; q < Q; ++q)
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                T sv = fv[q][c][0];
                if (MODE == 2) {
                    sv = fv[q][c][0] - fa * fv[q][c][1];        // bicgsolver.py:110
                    if ((oub >> (2 * q + c)) & 1u) g.src[kq[q][c]] = sv;
                }
                sr[q][c] = sv;
                wd[q][c] = cr[q][c][0] != CT(0) ? (CT)om * recip(cr[q][c][0]) : CT(0);   // omega D^-1
            }
#pragma unroll
        for (int q = 0; q < QZ; ++q) {
            const int idx = threadIdx.x + q * NT;
            if (idx < HZ * WZ) za[idx] = g.yc && ((zin >> q) & 1u) && mq[q] ? (T)yq[q] : T(0);
        }
        __syncthreads();
        if (SYM) {
            // aW = aE of the cell above, aS = aN of the cell to the left
#pragma unroll
            for (int q = 0; q < Q; ++q)
                if (threadIdx.x + q * NT < NP)
#pragma unroll
                    for (int c = 0; c < 2; ++c) {
                        cr[q][c][1] = aes[pa[q] * WR + pb[q][c]];
                        cr[q][c][4] = ans[pa[q] * (WR + 1) + pb[q][c]];
                    }
        }
        // half-sweep hs: red cells at even hs, black at odd, on the halo-h frame
#pragma unroll
        for (int hs = 0; hs < NS; ++hs) {
            const int h = NS - hs, c = hs & 1;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int a = pa[q], b = pb[q][c];
                if (!((inb >> (2 * q + c)) & 1u) || a < NS - h || a >= TI + NS + h || b < NS - h ||
                    b >= TJ + NS + h)
                    continue;
                const int pz = (a + 1) * WZ + (b + 1);
                const CT zc = za[pz];
                CT az = cr[q][c][0] * zc;
                az = az + cr[q][c][1] * (CT)za[pz - WZ];
                az = az + cr[q][c][2] * (CT)za[pz + WZ];
                az = az + cr[q][c][3] * (CT)za[pz + 1];
                az = az + cr[q][c][4] * (CT)za[pz - 1];
                za[pz] = (T)(zc + wd[q][c] * ((CT)sr[q][c] - az));
            }
            __syncthreads();
        }
        const T* __restrict__ zf = za;
        // the tile's cells: z out, then y = A z (and y on the free interior
        // cells into yb for the restriction)
        auto emit = [&](int q, int c) {
            const int a = pa[q], b = pb[q][c];
            const int pz = (a + 1) * WZ + (b + 1);
            const int64_t k = kq[q][c];
            const T zc = zf[pz];
            g.z[k] = zc;
            if (MODE > 0) {
                CT yv = cr[q][c][0] * (CT)zc;
                yv = yv + cr[q][c][1] * (CT)zf[pz - WZ];
                yv = yv + cr[q][c][2] * (CT)zf[pz + WZ];
                yv = yv + cr[q][c][3] * (CT)zf[pz + 1];
                yv = yv + cr[q][c][4] * (CT)zf[pz - 1];
                g.y[k] = (T)yv;
                const uint32_t bit = 1u << (2 * q + c);
                if (ib & bit) yb[(a - NS) * TJ + (b - NS)] = (fm & bit) ? (double)yv : 0.0;
                const double rtk = (double)rq[q][c];
                if (MODE == 1) {
                    a0 += rtk * (double)yv;
                } else {
                    const double sk = (double)sr[q][c];
                    a0 += (double)yv * sk;
                    a1 += (double)yv * (double)yv;
                    a2 += sk * rtk;
                    a3 += (double)yv * rtk;
                }
            }
        };
#pragma unroll
        for (int q = 0; q < QI; ++q)
#pragma unroll
            for (int c = 0; c < 2; ++c)
                if ((oub >> (2 * q + c)) & 1u) emit(q, c);
        if (edge) {
            // the boundary row or column of the system next to an edge tile
#pragma unroll
            for (int q = QI; q < Q; ++q)
#pragma unroll
                for (int c = 0; c < 2; ++c)
                    if ((oub >> (2 * q + c)) & 1u) emit(q, c);
        }
        if (MODE > 0 && g.rcs) {
            // P^T y over the aggregates this tile overlaps: one wave per
            // aggregate piece of yb, summed in a fixed order
            __syncthreads();
            const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
            const int r0 = ti0 - 1, c0 = tj0 - 1;   // interior coordinates of the tile origin
            const int r1 = r0 + TI < nx ? r0 + TI : nx, c1 = c0 + TJ < ny ? c0 + TJ : ny;
            const int B = cg.B;
            const int I0 = r0 / B, J0 = c0 / B;
            const int nJ = (c1 - 1) / B - J0 + 1, np = ((r1 - 1) / B - I0 + 1) * nJ;
            for (int u = w; u < np; u += NW) {
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
        if (tile + 1 < t1) __syncthreads();   // the next tile reuses za
    }
    if (MODE > 0) {
        // the partial sums, each reduced in the fixed order of block_sum
        constexpr int NV = MODE == 1 ? 1 : 4;
        double v[NV];
        v[0] = a0;
        if (NV > 1) {
            v[NV > 1 ? 1 : 0] = a1;
            v[NV > 2 ? 2 : 0] = a2;
            v[NV > 3 ? 3 : 0] = a3;
        }
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
        for (int q = 0; q < NV; ++q) v[q] = wave_sum(v[q]);
        if (lane == 0)
#pragma unroll
            for (int q = 0; q < NV; ++q) sh[q * NW + w] = v[q];
        __syncthreads();
        if (threadIdx.x == 0)
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                const double* hq = sh + q * NW;
                double t = (hq[0] + hq[1]) + (hq[2] + hq[3]);
                if (NW == 8) t = t + ((hq[4 % NW] + hq[5 % NW]) + (hq[6 % NW] + hq[7 % NW]));
                g.P[q * kMaxTileWg + L] = t;
            }
    }
}

template <typename T, int TI, int TJ, int MODE, int WPE, int SW, bool SELF, int NT = KM_BLOCK, bool SYM = false,
          bool C64 = false>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE, 8)))
k_tl_tile(ATile tl, CGeom cg, TlArgs<T> g) {
    tile_body<T, TI, TJ, MODE, SW, SELF, NT, SYM, C64>(tl, cg, g, xcd_remap(blockIdx.x, tl.G));
}

// Two independent systems in one launch: workgroups [0, tl0.G) take the
// first, the rest the second (each with its own arguments and state).
template <typename T, int TI, int TJ, int MODE, int WPE, int SW, bool SELF>
__global__ void __launch_bounds__(KM_BLOCK) __attribute__((amdgpu_waves_per_eu(WPE, 8)))
k_tl_tile2(ATile tl0, ATile tl1, CGeom cg, TlArgs<T> g0, TlArgs<T> g1) {
    if ((int)blockIdx.x < tl0.G)
        tile_body<T, TI, TJ, MODE, SW, SELF, KM_BLOCK>(tl0, cg, g0, xcd_remap(blockIdx.x, tl0.G));
    else
        tile_body<T, TI, TJ, MODE, SW, SELF, KM_BLOCK>(tl1, cg, g1, xcd_remap(blockIdx.x - tl0.G, tl1.G));
}

// One tile launch of the two-level smoother (kmcoarse::kSweeps red-black
// sweeps) with tile shape `which` (an index into kTileI / kTileJ); SYM: the
// three-plane kernel of a symmetric system; C64: FP64 coefficients and
// arithmetic for FP32 vectors (a.cd).
template <typename T, int MODE, bool SYM = false, bool C64 = false>
void launch_tile(int which, hipStream_t q, int NX, int NY, const CGeom& cg, const TlArgs<T>& a) {
#define KM_TL_CASE(IDX)                                                                                           \
    case IDX: {                                                                                                   \
        const ATile tl = a_tile(NX, NY, kTileI[IDX], kTileJ[IDX]);                                               \
        hipLaunchKernelGGL((k_tl_tile<T, kTileI[IDX], kTileJ[IDX], MODE, kTileW[IDX], kmcoarse::kSweeps, false,    \
                                      kTileT[IDX], SYM, C64>),                                                    \
                           dim3(tl.G), dim3(kTileT[IDX]), 0, q, tl, cg, a);                                      \
        break;                                                                                                    \
    }
    switch (which) {
        KM_TL_CASE(0)
        KM_TL_CASE(1)
        KM_TL_CASE(2)
        KM_TL_CASE(3)
        KM_TL_CASE(4)
        KM_TL_CASE(5)
        KM_TL_CASE(6)
        KM_TL_CASE(7)
        default:
        KM_TL_CASE(8)
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

// FIN_A: K_PV's partials (read by K_CV's Gc workgroups); FIN_T: K_ST's
// partials (read by K_CT's Gc + Gf workgroups)
int total_flags(const km_coarse* C, int Gt, int Gc, int Gf) {
    if (C->totals == 0) return 0;
    if (C->totals > 0) return FIN_A | FIN_T;
    return ((int64_t)Gt * Gc > kTotalWork ? FIN_A : 0) | ((int64_t)Gt * (Gc + Gf) > kTotalWork ? FIN_T : 0);
}

// The coarse rows of K_CV / K_CT: one wave per row of y = A_c^-1 rc, rc the
// slot sum staged in LDS by load_rc.  A lane's first kCoarseCh row values are
// loaded (first_row) before the launch's partials are summed, so one memory
// round trip serves them all (rows of up to 64 kCoarseCh unknowns); the rest
// of longer rows follow.  Summed in k order per lane.
constexpr int kCoarseCh = 24;

__device__ __forceinline__ void first_row(const double* __restrict__ a, int m, int lane, double (&av)[kCoarseCh]) {
#pragma unroll
    for (int u = 0; u < kCoarseCh; ++u) av[u] = lane + 64 * u < m ? a[lane + 64 * u] : 0.0;
}

__device__ __forceinline__ void load_rc(const double* __restrict__ rcs, int m, int S, double* rs) {
    for (int k = threadIdx.x; k < m; k += KM_BLOCK) {
        double acc = rcs[k];
        for (int q = 1; q < S; ++q) acc += rcs[(int64_t)q * m + k];
        rs[k] = acc;
    }
}

__device__ __forceinline__ double row_dot(const double* __restrict__ a, int m, int lane, const double (&av)[kCoarseCh],
                                          const double* rs) {
    double s = 0.0;
#pragma unroll
    for (int u = 0; u < kCoarseCh; ++u)
        if (lane + 64 * u < m) s += av[u] * rs[lane + 64 * u];
    for (int k0 = lane + 64 * kCoarseCh; k0 < m; k0 += 64 * kCoarseCh) {
        double bv[kCoarseCh];
#pragma unroll
        for (int u = 0; u < kCoarseCh; ++u) bv[u] = k0 + 64 * u < m ? a[k0 + 64 * u] : 0.0;
#pragma unroll
        for (int u = 0; u < kCoarseCh; ++u)
            if (k0 + 64 * u < m) s += bv[u] * rs[k0 + 64 * u];
    }
    return wave_sum(s);
}

// K_CV: the stopping test on <r,r> and alpha (every workgroup sums the
// partials in the same order), then Y(v) = A_c^-1 P^T v and Y(s) = Y(r) -
// alpha Y(v) on coarse row bx * 4 + wave (bx: the workgroup's index within
// its system).  Y holds Y(r), Y(p), Y(v), Y(s).
__device__ __forceinline__ void cv_body(int bx, int m, int S, int Gt, int Gf, int tot, int ph,
                                        const double* __restrict__ inv, const double* __restrict__ rcs,
                                        double* __restrict__ Y, KState* __restrict__ st,
                                        const double* __restrict__ P, int* flag) {
    extern __shared__ double rs_[];
    const int done = st->done;   // tested once the row loads are in flight
    const int row = bx * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    const bool live = row < m;
    const int rw = live ? row : 0;
    const double* a = inv + (int64_t)rw * m;
    const double yr = lane == 0 ? Y[rw] : 0.0;
    double av[kCoarseCh];
    first_row(a, m, lane, av);
    load_rc(rcs, m, S, rs_);
    if (done) return;
    double rvv, rr;
    if (tot & FIN_A) {
        double t1[1];
        sum_partials_wide<1>(P + TL_PR, Gf, t1);
        rvv = P[TL_TOT];
        rr = t1[0];
    } else {
        const double* planes[2] = {P + TL_PV, P + TL_PR};
        const int counts[2] = {Gt, Gf};
        double t2[2];
        sum_partials_each<2>(planes, counts, t2);
        rvv = t2[0];
        rr = t2[1];
    }
    double alpha;
    if (!cv_scalars(st, ph, rvv, rr, &alpha, flag, bx == 0 && threadIdx.x == 0)) return;
    __syncthreads();   // rs_
    if (!live) return;
    const double s = row_dot(a, m, lane, av, rs_);
    if (lane == 0) {
        Y[2 * m + row] = s;                  // Y(v)
        Y[3 * m + row] = yr - alpha * s;     // Y(s)
    }
}

__global__ void __launch_bounds__(KM_BLOCK)
k_tl_cv(int m, int S, int Gt, int Gf, int tot, int ph, const double* __restrict__ inv, const double* __restrict__ rcs,
        double* __restrict__ Y, KState* __restrict__ st, const double* __restrict__ P, int* flag) {
    cv_body(blockIdx.x, m, S, Gt, Gf, tot, ph, inv, rcs, Y, st, P, flag);
}

// omega, rho, beta of the next iteration from K_ST's partials (bicgsolver.py:
// 95,102,118; every workgroup derives the same, `lead` writes the next ring
// slot) and this iteration's alpha.
__device__ __forceinline__ void ct_scalars(const double* __restrict__ P, int Gt, int tot, KState* __restrict__ st,
                                           int ph, bool lead, double* omega_out, double* beta_out, double* alpha_out) {
    double pt[4];
    if (tot & FIN_T) {
#pragma unroll
        for (int q = 0; q < 4; ++q) pt[q] = P[TL_TOT + 1 + q];
    } else {
        sum_partials_wide<4>(P + TL_PT, Gt, pt, kMaxTileWg);
    }
    const KScal sp = st->ring[ph];
    const double omega = pt[1] != 0.0 ? pt[0] / pt[1] : 0.0;                  // bicgsolver.py:118
    const double rho = pt[2] - omega * pt[3];                                 // <r_i+1, r~> (bicgsolver.py:95)
    const double beta = omega != 0.0 ? (rho / sp.rho) * (sp.alpha / omega) : 0.0;   // bicgsolver.py:102
    if (lead) {
        KScal& nxt = st->ring[(ph + 1) % KM_RING];
        nxt.rho = rho;
        nxt.omega = omega;
        nxt.beta = beta;
        nxt.iter = sp.iter + 1;
    }
    *omega_out = omega;
    *beta_out = beta;
    *alpha_out = sp.alpha;
}

// x += alpha p^ + omega s^ (bicgsolver.py:128), r = s - omega t (:134),
// p = r + beta (p - omega v) (:103) on flat chunk wf, and its <r,r> partial.
// scalars(alpha, omega, beta) runs first (the whole workgroup calls it);
// false: the solve has ended, nothing is written.
template <typename T, typename F>
__device__ __forceinline__ void flat_update(const FGeom& fl, int wf, F scalars, T* __restrict__ x,
                                            T* __restrict__ r, T* __restrict__ p, const T* __restrict__ v,
                                            const T* __restrict__ s, const T* __restrict__ t,
                                            const T* __restrict__ ph_, const T* __restrict__ sh_,
                                            double* __restrict__ P, double* shr) {
    double alpha = 0.0, omega = 0.0, beta = 0.0;
    if (!scalars(alpha, omega, beta)) return;
    const T fo = (T)omega, fb = (T)beta;
    const int64_t lo = (int64_t)wf * fl.chunk;
    const int64_t hi = lo + fl.chunk < fl.n ? lo + fl.chunk : fl.n;
    double rr = 0.0;
    for (int64_t base = lo; base < hi; base += (int64_t)kFlatPer * KM_BLOCK) {
        T sv[kFlatPer], tv[kFlatPer], pv[kFlatPer], vv[kFlatPer], xv[kFlatPer], hv[kFlatPer], s2v[kFlatPer];
#pragma unroll
        for (int c = 0; c < kFlatPer; ++c) {
            const int64_t k = base + threadIdx.x + c * KM_BLOCK;
            const bool ok = k < hi;
            sv[c] = ok ? s[k] : T(0);
            tv[c] = ok ? t[k] : T(0);
            pv[c] = ok ? p[k] : T(0);
            vv[c] = ok ? v[k] : T(0);
            xv[c] = ok ? x[k] : T(0);
            hv[c] = ok ? ph_[k] : T(0);
            s2v[c] = ok ? sh_[k] : T(0);
        }
#pragma unroll
        for (int c = 0; c < kFlatPer; ++c) {
            const int64_t k = base + threadIdx.x + c * KM_BLOCK;
            if (k >= hi) continue;
            x[k] = xv[c] + (T)(alpha * (double)hv[c] + omega * (double)s2v[c]);
            const T rn = sv[c] - fo * tv[c];
            r[k] = rn;
            p[k] = rn + fb * (pv[c] - fo * vv[c]);
            rr += (double)rn * (double)rn;
        }
    }
    rr = block_sum(rr, shr);
    if (threadIdx.x == 0) P[TL_PR + wf] = rr;
}

// K_CT: omega, rho, beta of the next iteration from K_ST's partials (every
// workgroup), then workgroups [0, Gc) take coarse rows (Y(t) = A_c^-1 P^T t,
// Y(r) = Y(s) - omega Y(t), Y(p) = Y(r) + beta (Y(p) - omega Y(v))) and the
// rest flat chunks of the fine vectors: x += alpha p^ + omega s^ (bicgsolver.py:128),
// r = s - omega t (:134), p = r + beta (p - omega v) (:103) of the next
// iteration, partials of <r,r>.
// bx: the workgroup's index within its system.
template <typename T>
__device__ __forceinline__ void ct_body(int bx, int m, int S, int Gt, int Gc, const FGeom& fl, int tot, int ph,
                                        const double* __restrict__ inv, const double* __restrict__ rcs,
                                        double* __restrict__ Y, T* __restrict__ x, T* __restrict__ r,
                                        T* __restrict__ p, const T* __restrict__ v, const T* __restrict__ s,
                                        const T* __restrict__ t, const T* __restrict__ ph_,
                                        const T* __restrict__ sh_, KState* __restrict__ st, double* __restrict__ P) {
    extern __shared__ double rs_[];
    __shared__ double shr[8];
    const int done = st->done;   // tested once the first loads are in flight
    const bool coarse = bx < Gc;
    const int row = bx * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    const bool live = coarse && row < m;
    const int rw = live ? row : 0;
    const double* a = inv + (int64_t)rw * m;
    double y1 = 0.0, y2 = 0.0, y3 = 0.0;
    double av[kCoarseCh];
    if (coarse) {
        if (lane == 0) {
            y1 = Y[m + rw];
            y2 = Y[2 * m + rw];
            y3 = Y[3 * m + rw];
        }
        first_row(a, m, lane, av);
        load_rc(rcs, m, S, rs_);
    }
    const bool lead = bx == 0 && threadIdx.x == 0;
    if (coarse) {
        double omega, beta, alpha;
        ct_scalars(P, Gt, tot, st, ph, lead && !done, &omega, &beta, &alpha);
        if (done) return;
        __syncthreads();   // rs_
        if (!live) return;
        const double yt = row_dot(a, m, lane, av, rs_);
        if (lane == 0) {
            const double yr = y3 - omega * yt;   // Y(r)
            Y[row] = yr;
            Y[m + row] = yr + beta * (y1 - omega * y2);   // Y(p)
        }
        return;
    }
    // the partial sums load beside the state; a finished solve writes nothing
    auto scalars = [&](double& al, double& om, double& be) {
        ct_scalars(P, Gt, tot, st, ph, lead && !done, &om, &be, &al);
        return !done;
    };
    flat_update<T>(fl, bx - Gc, scalars, x, r, p, v, s, t, ph_, sh_, P, shr);
}

template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_tl_ct(int m, int S, int Gt, int Gc, FGeom fl, int tot, int ph, const double* __restrict__ inv,
        const double* __restrict__ rcs, double* __restrict__ Y, T* __restrict__ x, T* __restrict__ r,
        T* __restrict__ p, const T* __restrict__ v, const T* __restrict__ s, const T* __restrict__ t,
        const T* __restrict__ ph_, const T* __restrict__ sh_, KState* __restrict__ st, double* __restrict__ P) {
    ct_body<T>(blockIdx.x, m, S, Gt, Gc, fl, tot, ph, inv, rcs, Y, x, r, p, v, s, t, ph_, sh_, st, P);
}

// The flat workgroups of K_CT for two independent systems in one launch
// (workgroups [0, fl0.G) the first): each derives its system's scalars and
// updates its chunk.  S holds both systems' arguments.
template <typename T>
struct FlatSys {
    FGeom fl;
    int Gt, tot;
    T *x, *r, *p;
    const T *v, *s, *t, *ph, *sh;
    KState* st;
    double* P;
};

template <typename T>
__global__ void __launch_bounds__(KM_BLOCK) k_tl_flat2(int ph, FlatSys<T> a, FlatSys<T> b) {
    __shared__ double shr[8];
    const bool first = (int)blockIdx.x < a.fl.G;
    const FlatSys<T>& S = first ? a : b;
    const int wf = first ? (int)blockIdx.x : (int)blockIdx.x - a.fl.G;
    const int done = S.st->done;
    auto scalars = [&](double& al, double& om, double& be) {
        ct_scalars(S.P, S.Gt, S.tot, S.st, ph, wf == 0 && threadIdx.x == 0 && !done, &om, &be, &al);
        return !done;
    };
    flat_update<T>(S.fl, wf, scalars, S.x, S.r, S.p, S.v, S.s, S.t, S.ph, S.sh, S.P, shr);
}

// ------------------------------------------------------------ system groups
// Up to kGroupMax independent systems (the p' systems of several cases) in
// the same launches: workgroups [start[q], start[q + 1]) run system q's
// single-system code on its own geometry, arguments, state and partials,
// with bx = blockIdx.x - start[q] as its workgroup index (the tile kernels'
// XCD-aware order within each system), so each system's arithmetic is that
// of its own solve.
constexpr int kGroupMax = kmcoarse::kGroupSystems;
static_assert(kGroupMax <= KM_HW_SYSTEMS && (int64_t)kGroupMax * kPairPartials <= KM_MAX_PARTIALS,
              "each system of a group needs its host words and partial region");

__device__ __forceinline__ int group_of(const int (&start)[kGroupMax + 1], int n, int bx) {
    int q = 0;
#pragma unroll
    for (int k = 1; k < kGroupMax; ++k)
        if (k < n && bx >= start[k]) q = k;
    return q;
}

template <typename T>
struct TileGroup {
    int n;
    int start[kGroupMax + 1];
    ATile tl[kGroupMax];
    CGeom cg[kGroupMax];
    TlArgs<T> a[kGroupMax];
};

template <typename T, int TI, int TJ, int MODE, int WPE, int SW, int NT, bool SYM>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE, 8)))
k_tl_tile_n(TileGroup<T> g) {
    const int q = group_of(g.start, g.n, blockIdx.x);
    tile_body<T, TI, TJ, MODE, SW, false, NT, SYM>(g.tl[q], g.cg[q], g.a[q],
                                                   xcd_remap(blockIdx.x - g.start[q], g.tl[q].G));
}

// One grouped tile launch of tile shape `which` (g.start filled by the caller).
template <typename T, int MODE, bool SYM>
void launch_tile_n(int which, hipStream_t q, const TileGroup<T>& g) {
#define KM_TL_CASE(IDX)                                                                                           \
    case IDX:                                                                                                     \
        hipLaunchKernelGGL((k_tl_tile_n<T, kTileI[IDX], kTileJ[IDX], MODE, kTileW[IDX], kmcoarse::kSweeps,        \
                                        kTileT[IDX], SYM>),                                                       \
                           dim3(g.start[g.n]), dim3(kTileT[IDX]), 0, q, g);                                       \
        break;
    switch (which) {
        KM_TL_CASE(0)
        KM_TL_CASE(1)
        KM_TL_CASE(2)
        KM_TL_CASE(3)
        KM_TL_CASE(4)
        KM_TL_CASE(5)
        KM_TL_CASE(6)
        KM_TL_CASE(7)
        default:
        KM_TL_CASE(8)
    }
#undef KM_TL_CASE
}

// K_CV / K_CT arguments of one system of a group
template <typename T>
struct CoarseSys {
    int m, S, Gt, Gc, Gf, tot;
    FGeom fl;
    const double* inv;
    const double* rcv;   // K_CV's slot partials (P^T v)
    const double* rct;   // K_CT's (P^T t)
    double* Y;
    T *x, *r, *p;
    const T *v, *s, *t, *ph, *sh;
    KState* st;
    double* P;
    int* flag;
};

template <typename T>
struct CoarseGroup {
    int n;
    int start[kGroupMax + 1];
    CoarseSys<T> a[kGroupMax];
};

template <typename T>
__global__ void __launch_bounds__(KM_BLOCK) k_tl_cv_n(CoarseGroup<T> g, int ph) {
    const int q = group_of(g.start, g.n, blockIdx.x);
    const CoarseSys<T>& a = g.a[q];
    cv_body(blockIdx.x - g.start[q], a.m, a.S, a.Gt, a.Gf, a.tot, ph, a.inv, a.rcv, a.Y, a.st, a.P, a.flag);
}

template <typename T>
__global__ void __launch_bounds__(KM_BLOCK) k_tl_ct_n(CoarseGroup<T> g, int ph) {
    const int q = group_of(g.start, g.n, blockIdx.x);
    const CoarseSys<T>& a = g.a[q];
    ct_body<T>(blockIdx.x - g.start[q], a.m, a.S, a.Gt, a.Gc, a.fl, a.tot, ph, a.inv, a.rct, a.Y, a.x, a.r, a.p,
               a.v, a.s, a.t, a.ph, a.sh, a.st, a.P);
}

}  // namespace

namespace kmcoarse {

// Tile shape of the smoothing launches for the coarse level's system in T:
// on the first call for a (shape, precision) each shape whose restriction
// slots fit kMaxVS runs 3 + 16 back-to-back K_PV + K_ST pairs on scratch
// vectors (ws: its first 5n values of T are overwritten; src is read) and
// the fastest is kept (more slots only lengthen the coarse launches' slot
// sums, m S values).
template <typename T>
int tuned_tile(km_ctx* ctx, const T* coefT, const T* src, T* ws, int* out, bool sym = false,
               const double* cd = nullptr) {
    constexpr bool c64 = sizeof(T) == sizeof(float);   // FP32 vectors: FP64 coefficients (cd)
    KM_REQUIRE(ctx, !c64 || cd, "FP32 tiles need the FP64 coefficients");
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
    hipEvent_t e0 = nullptr, e1 = nullptr;
    KM_HIP_SYNC(ctx, hipEventCreate(&e0));
    KM_HIP_SYNC(ctx, hipEventCreate(&e1));
    int best = -1;
    float best_ms = 0.f;
    for (int w = 0; w < kNumTiles; ++w) {
        if (tile_slots(C, w) > kMaxVS) continue;
        TlArgs<T> a{};   // K_PV: p^ = M(src), v
        a.c = coefT;
        a.cd = cd;
        a.mask = C->mask;
        a.f0 = src;
        a.rt = src;
        a.z = ws;
        a.y = ws + n;
        a.yc = C->yc;
        a.rcs = C->frc;
        a.P = ctx->partials + TL_PV;
        a.om = (T)kOmega;
        TlArgs<T> b = a;   // K_ST: s = r - alpha v (alpha 0 without a state), s^, t
        b.f0 = src;
        b.f1 = ws + n;
        b.src = ws + 2 * n;
        b.z = ws + 3 * n;
        b.y = ws + 4 * n;
        b.rcs = C->frc + (size_t)kMaxVS * C->m;
        b.P = ctx->partials + TL_PT;
        const CGeom cg = c_geom(C, w);
        auto once = [&]() {
            if (sym) {
                launch_tile<T, 1, true, c64>(w, ctx->stream, C->NX, C->NY, cg, a);
                launch_tile<T, 2, true, c64>(w, ctx->stream, C->NX, C->NY, cg, b);
            } else {
                launch_tile<T, 1, false, c64>(w, ctx->stream, C->NX, C->NY, cg, a);
                launch_tile<T, 2, false, c64>(w, ctx->stream, C->NX, C->NY, cg, b);
            }
        };
        for (int k = 0; k < 3; ++k) once();
        (void)hipEventRecord(e0, ctx->stream);
        for (int k = 0; k < 16; ++k) once();
        (void)hipEventRecord(e1, ctx->stream);
        hipError_t se;
        {
            KmHostLock lk(km_host_mutex());
            se = hipEventSynchronize(e1);
        }
        float ms = 0.f;
        if (se == hipSuccess) se = hipEventElapsedTime(&ms, e0, e1);
        if (se != hipSuccess) {
            { KmHostLock lk_(km_host_mutex()); (void)hipEventDestroy(e0); }
            { KmHostLock lk_(km_host_mutex()); (void)hipEventDestroy(e1); }
            return km_fail(ctx, KM_EHIP, "smoothing tile timing: %s", hipGetErrorString(se));
        }
        if (best < 0 || ms < best_ms) {
            best = w;
            best_ms = ms;
        }
    }
    { KmHostLock lk_(km_host_mutex()); (void)hipEventDestroy(e0); }
    { KmHostLock lk_(km_host_mutex()); (void)hipEventDestroy(e1); }
    C->tiles[key] = best;
    *out = best;
    return KM_OK;
}

int tile_info(km_ctx* ctx, int* ti, int* tj, int* threads) {
    const km_coarse* C = ctx->coarse;
    KM_REQUIRE(ctx, C && C->arena, "no coarse level");
    auto it = C->tiles.find(std::make_tuple(C->NX, C->NY, (int)sizeof(double)));
    const int w = it == C->tiles.end() ? 0 : it->second;
    *ti = kTileI[w];
    *tj = kTileJ[w];
    if (threads) *threads = kTileT[w];
    return KM_OK;
}

// Pin the smoothing tile of the coarse level's system (both precisions);
// which < 0 clears the choice, so the next solve times the shapes again.
int set_tile(km_ctx* ctx, int which) {
    km_coarse* C = ctx->coarse;
    KM_REQUIRE(ctx, C && C->arena, "no coarse level yet (solve once first)");
    KM_REQUIRE(ctx, which < kNumTiles, "tile %d of %d", which, kNumTiles);
    if (which >= 0)
        KM_REQUIRE(ctx, tile_slots(C, which) <= kMaxVS, "tile %d: %d restriction slots", which, tile_slots(C, which));
    const int NX = C->NX, NY = C->NY;
    for_each_level(ctx, [&](km_coarse* L) {
        for (int sz : {(int)sizeof(float), (int)sizeof(double)}) {
            const auto key = std::make_tuple(NX, NY, sz);
            if (which < 0) L->tiles.erase(key);
            else L->tiles[key] = which;
        }
    });
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
    a.om = (T)kOmega;
    launch_tile<T, 0>(0, s, C->NX, C->NY, c_geom(C, 0), a);
    KM_CHECK_LAUNCH(ctx);
    return KM_OK;
}

namespace {

// The vectors and launch arguments of one fused solve (ws: 8 n values of T).
template <typename T>
struct Fused {
    int NX, NY, tile, Gt, Gc, S, fin;
    FGeom fl;
    CGeom cg;
    T *r, *rt, *p, *v, *s, *t, *ph, *sh;
    T* x;
    const T* coefT;
    double *Y, *rcv, *rct;
    km_coarse* C;
    KState* st;
    double* P;
    int* flag;

    bool sym;            // symmetric system: the three-plane tile kernels
    const double* cd;    // FP64 coefficients (the FP32-vector tiles compute with them)

    // slot: the system's place in a group solve (its state st[4 + slot], its
    // partials and host words beyond the earlier systems'); C: ctx->coarse
    Fused(km_ctx* ctx, int NX_, int NY_, const T* coef, T* ws, T* xd, int tile_, bool sym_ = false,
          const double* cd_ = nullptr, int slot = 0)
        : NX(NX_), NY(NY_), tile(tile_), x(xd), coefT(coef), C(ctx->coarse), st(&ctx->st[4 + slot]),
          P(ctx->partials + (size_t)slot * kPairPartials), flag(ctx->d_flag + KM_HW * slot), sym(sym_), cd(cd_) {
        const int64_t n = (int64_t)NX * NY;
        r = ws;
        rt = ws + n;
        p = ws + 2 * n;
        v = ws + 3 * n;
        s = ws + 4 * n;
        t = ws + 5 * n;
        ph = ws + 6 * n;
        sh = ws + 7 * n;
        fl = f_geom(n);
        Gt = tile_workgroups(tile, NX, NY);
        Gc = (C->m + 3) / 4;
        S = tile_slots(C, tile);
        cg = c_geom(C, tile);
        fin = total_flags(C, Gt, Gc, fl.G);
        Y = C->fy;
        rcv = C->frc;
        rct = C->frc + (size_t)kMaxVS * C->m;
    }

    TlArgs<T> pv_args(int phs) const {
        TlArgs<T> a{};
        a.c = coefT;
        a.cd = cd;
        a.mask = C->mask;
        a.f0 = p;
        a.rt = rt;
        a.z = ph;
        a.y = v;
        a.yc = Y + C->m;   // Y(p)
        a.rcs = rcv;
        a.st = st;
        a.P = P + TL_PV;
        a.ph = phs;
        a.om = (T)kOmega;
        return a;
    }
    void pv(hipStream_t q, int phs) const { tiles<1>(q, pv_args(phs)); }
    // the tile launch of this solve's kernel variant
    template <int MODE>
    void tiles(hipStream_t q, TlArgs<T> a) const {
        constexpr bool c64 = sizeof(T) == sizeof(float);
        a.cd = cd;
        if (sym) launch_tile<T, MODE, true, c64>(tile, q, NX, NY, cg, a);
        else launch_tile<T, MODE, false, c64>(tile, q, NX, NY, cg, a);
    }
    TlArgs<T> st_args(int phs) const {
        TlArgs<T> a{};
        a.c = coefT;
        a.cd = cd;
        a.mask = C->mask;
        a.f0 = r;
        a.f1 = v;
        a.rt = rt;
        a.src = s;
        a.z = sh;
        a.y = t;
        a.yc = Y + 3 * C->m;   // Y(s)
        a.rcs = rct;
        a.st = st;
        a.P = P + TL_PT;
        a.ph = phs;
        a.om = (T)kOmega;
        return a;
    }
    void st_(hipStream_t q, int phs) const { tiles<2>(q, st_args(phs)); }
    CoarseSys<T> coarse_sys() const {
        CoarseSys<T> a;
        a.m = C->m;
        a.S = S;
        a.Gt = Gt;
        a.Gc = Gc;
        a.Gf = fl.G;
        a.tot = fin;
        a.fl = fl;
        a.inv = C->inv;
        a.rcv = rcv;
        a.rct = rct;
        a.Y = Y;
        a.x = x;
        a.r = r;
        a.p = p;
        a.v = v;
        a.s = s;
        a.t = t;
        a.ph = ph;
        a.sh = sh;
        a.st = st;
        a.P = P;
        a.flag = flag;
        return a;
    }
    void total_a(hipStream_t q) const {
        if (fin & FIN_A)
            hipLaunchKernelGGL(k_tl_total<1>, dim3(1), dim3(KM_BLOCK), 0, q, (const double*)(P + TL_PV), Gt,
                               kMaxTileWg, P + TL_TOT, (const KState*)st);
    }
    void total_t(hipStream_t q) const {
        if (fin & FIN_T)
            hipLaunchKernelGGL(k_tl_total<4>, dim3(1), dim3(KM_BLOCK), 0, q, (const double*)(P + TL_PT), Gt,
                               kMaxTileWg, P + TL_TOT + 1, (const KState*)st);
    }
    void cv(hipStream_t q, int phs) const {
        hipLaunchKernelGGL(k_tl_cv, dim3(Gc), dim3(KM_BLOCK), (size_t)C->m * sizeof(double), q, C->m, S, Gt, fl.G,
                           fin, phs, (const double*)C->inv, (const double*)rcv, Y, st, (const double*)P, flag);
    }
    void ct(hipStream_t q, int phs) const {
        hipLaunchKernelGGL(k_tl_ct<T>, dim3(Gc + fl.G), dim3(KM_BLOCK), (size_t)C->m * sizeof(double), q, C->m, S, Gt,
                           Gc, fl, fin, phs, (const double*)C->inv, (const double*)rct, Y, x, r, p, (const T*)v,
                           (const T*)s, (const T*)t, (const T*)ph, (const T*)sh, st, P);
    }
    // one inner iteration
    void iteration(hipStream_t q, int phs) const {
        pv(q, phs);
        total_a(q);
        cv(q, phs);
        st_(q, phs);
        total_t(q);
        ct(q, phs);
    }
};

// x = 0, r = r~ = p = rhs, the state, Y(r) = Y(p) = A_c^-1 P^T rhs, the slots cleared
template <typename T>
int fused_start(km_ctx* ctx, const Fused<T>& f, const T* rhs, double eps, int max_iter) {
    km_coarse* C = f.C;
    hipLaunchKernelGGL(k_tl_init<T>, dim3(f.fl.G), dim3(KM_BLOCK), 0, ctx->stream, f.fl, rhs, f.x, f.r, f.rt, f.p, eps,
                       max_iter, f.st, f.P);
    KM_CHECK_LAUNCH(ctx);
    if (int rc = restrict_solve<T>(ctx, rhs, nullptr, ctx->stream)) return rc;
    const int64_t cells = std::max<int64_t>(C->m, (int64_t)f.S * C->m);
    hipLaunchKernelGGL(k_tl_cinit, dim3((unsigned)((cells + KM_BLOCK - 1) / KM_BLOCK)), dim3(KM_BLOCK), 0,
                       ctx->stream, C->m, f.S, (const double*)C->yc, f.Y, f.rcv, f.rct);
    KM_CHECK_LAUNCH(ctx);
    return KM_OK;
}

// One iteration of a group of fused solves (symmetric FP64 systems, one tile
// shape): K_PV, K_CV, K_ST and K_CT each one launch over every system, and
// each system's totals where its sizes ask for them.
struct FusedGroup {
    int n = 0, tile = 0;
    const Fused<double>* f[kGroupMax] = {};

    void iteration(hipStream_t q, int phs) const {
        TileGroup<double> g{};
        g.n = n;
        CoarseGroup<double> c{};
        c.n = n;
        int mmax = 1;
        for (int k = 0; k < n; ++k) {
            g.tl[k] = a_tile(f[k]->NX, f[k]->NY, kTileI[tile], kTileJ[tile]);
            g.cg[k] = f[k]->cg;
            g.a[k] = f[k]->pv_args(phs);
            g.start[k + 1] = g.start[k] + g.tl[k].G;
            c.a[k] = f[k]->coarse_sys();
            c.start[k + 1] = c.start[k] + f[k]->Gc;
            mmax = std::max(mmax, f[k]->C->m);
        }
        const size_t lds = (size_t)mmax * sizeof(double);
        launch_tile_n<double, 1, true>(tile, q, g);
        for (int k = 0; k < n; ++k) f[k]->total_a(q);
        hipLaunchKernelGGL(k_tl_cv_n<double>, dim3(c.start[n]), dim3(KM_BLOCK), lds, q, c, phs);
        for (int k = 0; k < n; ++k) g.a[k] = f[k]->st_args(phs);
        launch_tile_n<double, 2, true>(tile, q, g);
        for (int k = 0; k < n; ++k) f[k]->total_t(q);
        for (int k = 0; k < n; ++k) c.start[k + 1] = c.start[k] + f[k]->Gc + f[k]->fl.G;
        hipLaunchKernelGGL(k_tl_ct_n<double>, dim3(c.start[n]), dim3(KM_BLOCK), lds, q, c, phs);
    }
};

// The first chunk length's hint of a level: where its short chunks begin
// (the shortest of its last solves, less the lead), 0 without history.
int expected_iterations(const km_coarse* C) {
    int expect = 0;
    for (int c : C->recent)
        if (c > 0 && (expect == 0 || c < expect)) expect = c;
    if (expect > 0) expect = std::max(1, expect - C->chunk_lead + 2);
    return expect;
}

// The level's record of a finished solve: its iteration count, and a solve
// over twice as long as the shortest recent one asks for a refactorisation
// (the lagged coarse inverse no longer serves).
void record_solve(km_coarse* C, int iters) {
    for (int k = 7; k > 0; --k) C->recent[k] = C->recent[k - 1];
    C->recent[0] = iters;
    int shortest = 0;
    for (int k = 1; k < 8; ++k)
        if (C->recent[k] > 0 && (shortest == 0 || C->recent[k] < shortest)) shortest = C->recent[k];
    if (shortest > 0 && iters > 2 * shortest + 4) C->refresh = true;
}

}  // namespace

template <typename T>
int fused_inner(km_ctx* ctx, int NX, int NY, const T* coefT, const T* rhs, T* xd, T* ws, double eps, int max_iter,
                bool use_graph, KState* out, const std::function<int()>* tail, bool sym, const double* coef64) {
    km_coarse* C = ctx->coarse;
    KM_REQUIRE(ctx, C && C->arena && C->NX == NX && C->NY == NY, "coarse level not set up for this system");
    if (sizeof(T) == sizeof(double)) coef64 = reinterpret_cast<const double*>(coefT);
    KM_REQUIRE(ctx, coef64, "the FP32-vector two-level solve needs the FP64 coefficients");
    int tile = 0;
    if (int rc = tuned_tile<T>(ctx, coefT, rhs, ws, &tile, sym, coef64)) return rc;
    const Fused<T> f(ctx, NX, NY, coefT, ws, xd, tile, sym, coef64);
    if (int rc = fused_start<T>(ctx, f, rhs, eps, max_iter)) return rc;
    auto enqueue = [&f](hipStream_t q, int phs) { f.iteration(q, phs); };
    GraphKey key;
    key.v = {10, sizeof(T), (uintptr_t)NX, (uintptr_t)NY, (uintptr_t)coefT, (uintptr_t)rhs, (uintptr_t)xd,
             (uintptr_t)ws, (uintptr_t)f.st, (uintptr_t)C, (uintptr_t)C->arena, (uintptr_t)C->m, (uintptr_t)C->B,
             (uintptr_t)tile, (uintptr_t)f.fin, (uintptr_t)sym, (uintptr_t)coef64};
    // the shortest of the last solves of this system shape (the outer steps
    // of a time step differ) sets where the short chunks begin
    const int expect = expected_iterations(C);
    const int rc = run_chunks_flag(ctx, key, use_graph, C->chunk_big, C->chunk_small, expect, max_iter,
                                   enqueue, out, tail);
    if (rc == KM_OK) record_solve(C, out->iter);
    return rc;
}

int group_tile(km_ctx* ctx, km_coarse* C, const double* coef, const double* rhs, double* ws, int* tile) {
    ctx->coarse = C;
    return tuned_tile<double>(ctx, coef, rhs, ws, tile, true, coef);
}

int fused_inner_n(km_ctx* ctx, int n, km_coarse* const* C, const int* NX, const int* NY, const double* const* coef,
                  const double* const* rhs, double* const* xd, double* const* ws, double eps, const int* max_iter,
                  bool use_graph, KState* out, const std::function<int()>* tail, int tile) {
    KM_REQUIRE(ctx, n >= 1 && n <= kGroupMax, "%d systems in one group", n);
    KM_REQUIRE(ctx, tile >= 0 && tile < kNumTiles, "tile %d", tile);
    std::vector<Fused<double>> f;
    f.reserve(n);
    int mi = 0, expect = 0;
    bool history = true;
    for (int k = 0; k < n; ++k) {
        KM_REQUIRE(ctx, C[k] && C[k]->arena && C[k]->NX == NX[k] && C[k]->NY == NY[k],
                   "coarse level of system %d not set up", k);
        KM_REQUIRE(ctx, (int64_t)NX[k] * NY[k] < kTileCellsMax, "system %d too large for the tile kernels", k);
        KM_REQUIRE(ctx, tile_slots(C[k], tile) <= kMaxVS, "tile %d: %d restriction slots on system %d", tile,
                   tile_slots(C[k], tile), k);
        for (int j = 0; j < k; ++j) KM_REQUIRE(ctx, C[j] != C[k] && xd[j] != xd[k], "systems %d and %d overlap", j, k);
        ctx->coarse = C[k];
        f.emplace_back(ctx, NX[k], NY[k], coef[k], ws[k], xd[k], tile, true, coef[k], k);
        if (int rc = fused_start<double>(ctx, f[k], rhs[k], eps, max_iter[k])) return rc;
        mi = std::max(mi, max_iter[k]);
        const int e = expected_iterations(C[k]);
        history = history && e > 0;
        expect = std::max(expect, e);
    }
    FusedGroup g;
    g.n = n;
    g.tile = tile;
    GraphKey key;
    key.v = {13, (uintptr_t)n, (uintptr_t)tile};
    for (int k = 0; k < n; ++k) {
        g.f[k] = &f[k];
        for (uintptr_t v : {(uintptr_t)NX[k], (uintptr_t)NY[k], (uintptr_t)coef[k], (uintptr_t)rhs[k], (uintptr_t)xd[k],
                            (uintptr_t)ws[k], (uintptr_t)C[k], (uintptr_t)C[k]->arena, (uintptr_t)C[k]->m,
                            (uintptr_t)C[k]->B, (uintptr_t)f[k].fin})
            key.v.push_back(v);
    }
    auto enqueue = [&g](hipStream_t q, int phs) { g.iteration(q, phs); };
    // the slowest system sets the pace: short chunks from the latest expected end
    const int rc = run_chunks_flag(ctx, key, use_graph, C[0]->chunk_big, C[0]->chunk_small, history ? expect : 0, mi,
                                   enqueue, out, tail, n);
    if (rc == KM_OK)
        for (int k = 0; k < n; ++k) record_solve(C[k], out[k].iter);
    return rc;
}

// Average device time of one launch of a fused two-level kernel (which: 4 K_PV,
// 5 K_CV, 6 K_ST, 7 K_CT) inside FP64 iterations on (coef, b) with the stopping
// test disabled; events around every launch of that kernel, or (b2b) around
// reps identical launches after eight warm iterations.
int time_kernel(km_ctx* ctx, int which, int NX, int NY, const double* coef, const double* b, int reps, bool b2b,
                double* avg_us, bool sym) {
    KM_REQUIRE(ctx, which >= 4 && which <= 7, "unknown kernel id %d", which);
    if (int rc = setup(ctx, NX, NY, coef, ctx->stream)) return rc;
    const int64_t n = (int64_t)NX * NY;
    if (int rc = km_ensure_workspace(ctx, (size_t)9 * n * sizeof(double))) return rc;
    double* ws = static_cast<double*>(km_ws_vectors(ctx));
    double* x = ws + 8 * n;
    int tile = 0;
    if (int rc = tuned_tile<double>(ctx, coef, b, ws, &tile, sym)) return rc;
    const Fused<double> f(ctx, NX, NY, coef, ws, x, tile, sym);
    if (int rc = fused_start<double>(ctx, f, b, 0.0, 1 << 30)) return rc;
    std::vector<hipEvent_t> ev(2 * (size_t)reps, nullptr);
    for (auto& e : ev) KM_HIP_SYNC(ctx, hipEventCreate(&e));
    auto cleanup = [&]() {
        for (auto& e : ev)
            if (e) { KmHostLock lk_(km_host_mutex()); (void)hipEventDestroy(e); }
    };
    hipStream_t q = ctx->stream;
    auto launch = [&](int k, int phs) {
        switch (k) {
        case 4: f.pv(q, phs); break;
        case 5: f.cv(q, phs); break;
        case 6: f.st_(q, phs); break;
        default: f.ct(q, phs);
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
        // each kernel re-run with the phase of the last warm iteration
        const int phs = (warm - 1) % KM_RING;
        KM_HIP(ctx, hipEventRecord(ev[0], q));
        for (int k = 0; k < reps; ++k) launch(which, phs);
        KM_HIP(ctx, hipEventRecord(ev[1], q));
    } else {
        for (int it = 0; it < reps; ++it)
            for (int k = 4; k <= 7; ++k)
                one(k, (warm + it) % KM_RING, k == which ? ev[2 * it] : nullptr, k == which ? ev[2 * it + 1] : nullptr);
    }
    hipError_t le = hipGetLastError();
    hipError_t se;
    {
        KmHostLock lk(km_host_mutex());
        se = hipStreamSynchronize(q);
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
    KM_HIP_SYNC(ctx, hipMemcpy(&ctx->h_st[3], f.st, sizeof(KState), hipMemcpyDeviceToHost));
    KM_REQUIRE(ctx, !ctx->h_st[3].done, "km_time_kernel: the iteration stopped during timing (status %d)",
               ctx->h_st[3].status);
    *avg_us = 1e3 * total_ms / reps;
    return KM_OK;
}

namespace {

// The smoothed flexible solve without a coarse level (the momentum systems):
// M(r) = red-black Gauss-Seidel sweeps on A z = r from z = 0, on 16 x 32
// tiles; three launches per iteration (K_PV, K_ST deriving alpha and the
// stopping test itself, K_CT with flat workgroups only).  One sweep, or two
// for FP64 systems of at least kMomTwoSweepCells cells (km_set_momentum_sweeps
// overrides): on the 1280 x 256 channel two sweeps take u / v from 6 / 7 to
// 5 / 6 iterations (SIMPLE steps 1.8% faster); where several streams share
// the GPU (the many-case steps) the extra sweep's work is not hidden and one
// sweep is faster, as it is on small systems, whose launches are
// latency-bound.
constexpr int kMomTile = 2;
constexpr int64_t kMomTwoSweepCells = 200000;
constexpr double kMomOmega = 1.0;   // over-relaxed sweeps diverge on the convective momentum systems

template <typename T>
int mom_sweeps(const km_ctx* ctx, int64_t cells) {
    if (ctx->mom_sweeps > 0) return ctx->mom_sweeps;
    return sizeof(T) == sizeof(double) && cells >= kMomTwoSweepCells ? 2 : 1;
}

template <typename T, int MODE, bool SELF>
void launch_mom(hipStream_t q, int NX, int NY, const TlArgs<T>& a, int sw) {
    const ATile tl = a_tile(NX, NY, kTileI[kMomTile], kTileJ[kMomTile]);
    CGeom cg;
    cg.B = 1;
    cg.mJ = 1;
    cg.m = 0;
    cg.SJ = 1;
    if (sw == 2)
        hipLaunchKernelGGL((k_tl_tile<T, kTileI[kMomTile], kTileJ[kMomTile], MODE, kTileW[kMomTile], 2, SELF>),
                           dim3(tl.G), dim3(KM_BLOCK), 0, q, tl, cg, a);
    else
        hipLaunchKernelGGL((k_tl_tile<T, kTileI[kMomTile], kTileJ[kMomTile], MODE, kTileW[kMomTile], 1, SELF>),
                           dim3(tl.G), dim3(KM_BLOCK), 0, q, tl, cg, a);
}

template <typename T, int MODE, bool SELF>
void launch_mom2(hipStream_t q, int NX0, int NY0, int NX1, int NY1, const TlArgs<T>& a0, const TlArgs<T>& a1,
                 int sw) {
    const ATile t0 = a_tile(NX0, NY0, kTileI[kMomTile], kTileJ[kMomTile]);
    const ATile t1 = a_tile(NX1, NY1, kTileI[kMomTile], kTileJ[kMomTile]);
    CGeom cg;
    cg.B = 1;
    cg.mJ = 1;
    cg.m = 0;
    cg.SJ = 1;
    if (sw == 2)
        hipLaunchKernelGGL((k_tl_tile2<T, kTileI[kMomTile], kTileJ[kMomTile], MODE, kTileW[kMomTile], 2, SELF>),
                           dim3(t0.G + t1.G), dim3(KM_BLOCK), 0, q, t0, t1, cg, a0, a1);
    else
        hipLaunchKernelGGL((k_tl_tile2<T, kTileI[kMomTile], kTileJ[kMomTile], MODE, kTileW[kMomTile], 1, SELF>),
                           dim3(t0.G + t1.G), dim3(KM_BLOCK), 0, q, t0, t1, cg, a0, a1);
}

// One system's vectors (ws: 8n values of T), state, partials and host words.
template <typename T>
struct SmoothFused {
    int NX, NY, Gt, fin;
    int sw;   // red-black sweeps of M (mom_sweeps)
    FGeom fl;
    T *r, *rt, *p, *v, *s, *t, *ph, *sh, *x;
    const T* coefT;
    KState* st;
    double* P;
    int* flag;

    SmoothFused(int NX_, int NY_, const T* coef, T* ws, T* xd, KState* st_, double* P_, int* flag_, int sw_)
        : NX(NX_), NY(NY_), sw(sw_), x(xd), coefT(coef), st(st_), P(P_), flag(flag_) {
        const int64_t n = (int64_t)NX * NY;
        r = ws;
        rt = ws + n;
        p = ws + 2 * n;
        v = ws + 3 * n;
        s = ws + 4 * n;
        t = ws + 5 * n;
        ph = ws + 6 * n;
        sh = ws + 7 * n;
        fl = f_geom(n);
        Gt = a_tile(NX, NY, kTileI[kMomTile], kTileJ[kMomTile]).G;
        // K_ST's Gt workgroups read <r~,v> and <r,r>, K_CT's fl.G read K_ST's planes
        fin = ((int64_t)Gt * (Gt + fl.G) > kTotalWork ? FIN_A | FIN_R : 0) |
              ((int64_t)Gt * fl.G > kTotalWork ? FIN_T : 0);
    }
    void init(hipStream_t q, const T* rhs, double eps, int max_iter) const {
        hipLaunchKernelGGL(k_tl_init<T>, dim3(fl.G), dim3(KM_BLOCK), 0, q, fl, rhs, x, r, rt, p, eps, max_iter, st, P);
        total_r(q);
    }
    // K_PV: p^ = M(p), v = A p^, <r~,v> partials
    TlArgs<T> pv_args(int phs) const {
        TlArgs<T> a{};
        a.c = coefT;
        a.f0 = p;
        a.rt = rt;
        a.z = ph;
        a.y = v;
        a.st = st;
        a.P = P + TL_PV;
        a.ph = phs;
        a.om = (T)kMomOmega;
        return a;
    }
    void pv_total(hipStream_t q) const {
        if (fin & FIN_A)
            hipLaunchKernelGGL(k_tl_total<1>, dim3(1), dim3(KM_BLOCK), 0, q, (const double*)(P + TL_PV), Gt,
                               kMaxTileWg, P + TL_TOT, (const KState*)st);
    }
    // K_ST: alpha and the stopping test, s = r - alpha v, s^ = M(s), t = A s^
    TlArgs<T> st_args(int phs) const {
        TlArgs<T> a{};
        a.c = coefT;
        a.f0 = r;
        a.f1 = v;
        a.rt = rt;
        a.src = s;
        a.z = sh;
        a.y = t;
        a.st = st;
        a.P = P + TL_PT;
        a.ph = phs;
        a.om = (T)kMomOmega;
        a.Pin = P;
        a.Gin[0] = Gt;
        a.Gin[1] = fl.G;
        a.tot = fin;
        a.stw = st;
        a.flag = flag;
        return a;
    }
    void st_total(hipStream_t q) const {
        if (fin & FIN_T)
            hipLaunchKernelGGL(k_tl_total<4>, dim3(1), dim3(KM_BLOCK), 0, q, (const double*)(P + TL_PT), Gt,
                               kMaxTileWg, P + TL_TOT + 1, (const KState*)st);
    }
    void total_r(hipStream_t q) const {
        if (fin & FIN_R)
            hipLaunchKernelGGL(k_tl_total<1>, dim3(1), dim3(KM_BLOCK), 0, q, (const double*)(P + TL_PR), fl.G,
                               kMaxFlatWg, P + TL_TOT + 5, (const KState*)st);
    }
    FlatSys<T> flat() const {
        FlatSys<T> f;
        f.fl = fl;
        f.Gt = Gt;
        f.tot = fin;
        f.x = x;
        f.r = r;
        f.p = p;
        f.v = v;
        f.s = s;
        f.t = t;
        f.ph = ph;
        f.sh = sh;
        f.st = st;
        f.P = P;
        return f;
    }
    void ct(hipStream_t q, int phs) const {
        hipLaunchKernelGGL(k_tl_ct<T>, dim3(fl.G), dim3(KM_BLOCK), 0, q, 0, 1, Gt, 0, fl, fin, phs,
                           (const double*)nullptr, (const double*)nullptr, (double*)nullptr, x, r, p, (const T*)v,
                           (const T*)s, (const T*)t, (const T*)ph, (const T*)sh, st, P);
        total_r(q);
    }
    void iteration(hipStream_t q, int phs) const {
        launch_mom<T, 1, false>(q, NX, NY, pv_args(phs), sw);
        pv_total(q);
        launch_mom<T, 2, true>(q, NX, NY, st_args(phs), sw);
        st_total(q);
        ct(q, phs);
    }
};

// One iteration of two independent systems in three launches: each launch
// covers both systems' workgroups, so the pair costs one system's latency.
template <typename T>
void pair_iteration(const SmoothFused<T>& a, const SmoothFused<T>& b, hipStream_t q, int phs) {
    launch_mom2<T, 1, false>(q, a.NX, a.NY, b.NX, b.NY, a.pv_args(phs), b.pv_args(phs), a.sw);
    a.pv_total(q);
    b.pv_total(q);
    launch_mom2<T, 2, true>(q, a.NX, a.NY, b.NX, b.NY, a.st_args(phs), b.st_args(phs), a.sw);
    a.st_total(q);
    b.st_total(q);
    hipLaunchKernelGGL(k_tl_flat2<T>, dim3(a.fl.G + b.fl.G), dim3(KM_BLOCK), 0, q, phs, a.flat(), b.flat());
    a.total_r(q);
    b.total_r(q);
}

}  // namespace

template <typename T>
int smooth_inner(km_ctx* ctx, int NX, int NY, const T* coefT, const T* rhs, T* xd, T* ws, double eps, int max_iter,
                 bool use_graph, KState* out, const std::function<int()>* tail) {
    KM_REQUIRE(ctx, (int64_t)NX * NY < kTileCellsMax, "system of %lld cells is too large for the tile kernels",
               (long long)NX * NY);
    const SmoothFused<T> f(NX, NY, coefT, ws, xd, &ctx->st[4], ctx->partials, ctx->d_flag,
                           mom_sweeps<T>(ctx, (int64_t)NX * NY));
    f.init(ctx->stream, rhs, eps, max_iter);
    KM_CHECK_LAUNCH(ctx);
    auto enqueue = [&f](hipStream_t q, int phs) { f.iteration(q, phs); };
    GraphKey key;
    key.v = {11, sizeof(T), (uintptr_t)NX, (uintptr_t)NY, (uintptr_t)coefT, (uintptr_t)rhs, (uintptr_t)xd,
             (uintptr_t)ws, (uintptr_t)f.st, (uintptr_t)f.fin, (uintptr_t)f.sw};
    // these solves take a handful of iterations: short graphs throughout
    return run_chunks_flag(ctx, key, use_graph, KM_RING, 2, 1, max_iter, enqueue, out, tail);
}

template <typename T>
int smooth_inner2(km_ctx* ctx, const int* NX, const int* NY, const T* const* coefT, const T* const* rhs, T* const* xd,
                  T* const* ws, double eps, const int* max_iter, bool use_graph, KState* out,
                  const std::function<int()>* tail) {
    KM_REQUIRE(ctx, (int64_t)NX[0] * NY[0] < kTileCellsMax && (int64_t)NX[1] * NY[1] < kTileCellsMax,
               "system too large for the tile kernels");
    // states st[4], st[5] (contiguous: one copy reads both), partials and
    // host words of the second system beyond the first's; both systems take
    // the larger one's sweep count (one kernel serves both)
    const int sw = mom_sweeps<T>(ctx, std::max((int64_t)NX[0] * NY[0], (int64_t)NX[1] * NY[1]));
    const SmoothFused<T> a(NX[0], NY[0], coefT[0], ws[0], xd[0], &ctx->st[4], ctx->partials, ctx->d_flag, sw);
    const SmoothFused<T> b(NX[1], NY[1], coefT[1], ws[1], xd[1], &ctx->st[5], ctx->partials + kPairPartials,
                           ctx->d_flag + KM_HW, sw);
    a.init(ctx->stream, rhs[0], eps, max_iter[0]);
    b.init(ctx->stream, rhs[1], eps, max_iter[1]);
    KM_CHECK_LAUNCH(ctx);
    auto enqueue = [&a, &b](hipStream_t q, int phs) { pair_iteration(a, b, q, phs); };
    GraphKey key;
    key.v = {12, sizeof(T), (uintptr_t)NX[0], (uintptr_t)NY[0], (uintptr_t)coefT[0], (uintptr_t)rhs[0],
             (uintptr_t)xd[0], (uintptr_t)ws[0], (uintptr_t)a.fin, (uintptr_t)NX[1], (uintptr_t)NY[1],
             (uintptr_t)coefT[1], (uintptr_t)rhs[1], (uintptr_t)xd[1], (uintptr_t)ws[1], (uintptr_t)b.fin,
             (uintptr_t)sw};
    return run_chunks_flag(ctx, key, use_graph, KM_RING, 2, 1, std::max(max_iter[0], max_iter[1]), enqueue, out,
                           tail, 2);
}

template int smooth_inner<float>(km_ctx*, int, int, const float*, const float*, float*, float*, double, int, bool,
                                 KState*, const std::function<int()>*);
template int smooth_inner<double>(km_ctx*, int, int, const double*, const double*, double*, double*, double, int,
                                  bool, KState*, const std::function<int()>*);
template int smooth_inner2<double>(km_ctx*, const int*, const int*, const double* const*, const double* const*,
                                   double* const*, double* const*, double, const int*, bool, KState*,
                                   const std::function<int()>*);

template int smooth<float>(km_ctx*, const float*, const float*, float*, hipStream_t);
template int smooth<double>(km_ctx*, const double*, const double*, double*, hipStream_t);
template int fused_inner<float>(km_ctx*, int, int, const float*, const float*, float*, float*, double, int, bool,
                                KState*, const std::function<int()>*, bool, const double*);
template int fused_inner<double>(km_ctx*, int, int, const double*, const double*, double*, double*, double, int,
                                 bool, KState*, const std::function<int()>*, bool, const double*);

}  // namespace kmcoarse

extern "C" {

int km_set_coarse_totals(km_ctx* ctx, int mode) {
    if (!ctx) return KM_EINVAL;
    KM_REQUIRE(ctx, ctx->coarse != nullptr, "no coarse level yet (solve once first)");
    kmcoarse::for_each_level(ctx, [&](km_coarse* C) { C->totals = mode < 0 ? -1 : (mode > 0 ? 1 : 0); });
    return KM_OK;
}

int km_set_coarse_chunks(km_ctx* ctx, int big, int small, int lead) {
    if (!ctx) return KM_EINVAL;
    KM_REQUIRE(ctx, ctx->coarse != nullptr, "no coarse level yet (solve once first)");
    KM_REQUIRE(ctx, big >= 0 && big % KM_RING == 0 && big <= 64, "chunk of %d iterations", big);
    KM_REQUIRE(ctx, small >= 0 && (small == 0 || KM_RING % small == 0), "short chunk of %d iterations", small);
    KM_REQUIRE(ctx, lead >= 0 && lead <= 1000, "lead of %d iterations", lead);
    kmcoarse::for_each_level(ctx, [&](km_coarse* C) {
        C->chunk_big = big ? big : 8;
        C->chunk_small = small ? small : 2;
        C->chunk_lead = lead;
    });
    return KM_OK;
}

int km_set_momentum_sweeps(km_ctx* ctx, int sweeps) {
    if (!ctx) return KM_EINVAL;
    KM_REQUIRE(ctx, sweeps >= 0 && sweeps <= 2, "%d momentum sweeps (0 automatic, 1 or 2)", sweeps);
    ctx->mom_sweeps = sweeps;
    return KM_OK;
}

int km_twolevel_tile(km_ctx* ctx, int* ti, int* tj) {
    if (!ctx || !ti || !tj) return KM_EINVAL;
    return kmcoarse::tile_info(ctx, ti, tj, nullptr);
}

int km_twolevel_tile_threads(km_ctx* ctx, int* threads) {
    if (!ctx || !threads) return KM_EINVAL;
    int ti = 0, tj = 0;
    return kmcoarse::tile_info(ctx, &ti, &tj, threads);
}

int km_twolevel_set_tile(km_ctx* ctx, int which) {
    if (!ctx) return KM_EINVAL;
    return kmcoarse::set_tile(ctx, which);
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
