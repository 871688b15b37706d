// Shared definitions of the gfx950 SIMPLE kernels: context, error handling,
// wave64 reductions, deterministic partial-sum reductions and the XCD-aware
// tile mapping.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <atomic>
#include <chrono>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <functional>
#include <map>
#include <mutex>
#include <set>
#include <shared_mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/karman.h"

#define KM_BLOCK 256
#define KM_NXCD 8           // MI355X: 8 XCDs, workgroups dispatched round-robin
#define KM_MAX_PARTIALS (1 << 18)

// Device-resident Krylov state.  The per-iteration scalars live in a ring of
// KM_RING slots indexed by the iteration's phase (iteration mod KM_RING, a
// kernel argument baked into each captured chunk); a kernel reads the slots
// of earlier kernels and only workgroup 0 writes the slot of the current
// iteration, so no slot is read and written by the same launch.
#define KM_RING 4
struct KScal {
    double rho, alpha, omega, beta;
    double best;          // flexible solve: smallest r.r so far (stall detection)
    int32_t iter;         // iteration index of this phase
    int32_t best_it;      // iteration of ``best``
};
struct KState {
    double r0;        // stopping-test denominator (r0.r0, or ||b||^2 for a warm start)
    double eps;       // relative tolerance on sqrt(r.r / r0)
    double rTr;       // latest r.r
    double ref;       // > 0: stopping denominator given by the caller
    int32_t max_iter, done, status, iter;
    KScal ring[KM_RING];
    double aux[8];
};

// Reductions: a reducing launch has at most KM_MAX_RED workgroups, each
// writing its partial sums with plain stores.  The following launch sums the
// partials itself, in every workgroup and in one fixed order, so all
// workgroups derive bitwise-identical scalars and only kernel boundaries
// order memory.  (Detecting the last workgroup with a device-scope atomic
// costs ~14 us per launch on gfx950, more than the whole stencil sweep of a
// 1280x256 system.)
#define KM_MAX_RED 2048
#define KM_PA 0                          // [2][KM_MAX_RED]
#define KM_PB (2 * KM_MAX_RED)           // [4][KM_MAX_RED]
#define KM_PC(par) (6 * KM_MAX_RED + (par) * 2 * KM_MAX_RED)   // [2][KM_MAX_RED], double-buffered
#define KM_PX (10 * KM_MAX_RED)          // helpers (dot, residual, mdiv): [2][KM_MAX_RED]
// The plain iteration's totals when a one-workgroup kernel sums a launch's
// partials once for all consumer workgroups (KM_PA's second plane):
// <r~,v> | <t,s>, <t,t> | <r,r>, <r,r~>
#define KM_PT (KM_PA + KM_MAX_RED)

// Geometry of a 2-D tiled stencil launch: TI rows x TJ columns per tile,
// tpb consecutive tiles per workgroup, G workgroups (<= KM_MAX_RED).
struct Tile {
    int NX, NY, TI, TJ, tiles_j, T, tpb, G;
};

static inline Tile make_tile(int NX, int NY, int TI, int TJ) {
    Tile t;
    t.NX = NX; t.NY = NY; t.TI = TI; t.TJ = TJ;
    t.tiles_j = (NY + TJ - 1) / TJ;
    const int tiles_i = (NX + TI - 1) / TI;
    t.T = tiles_i * t.tiles_j;
    t.tpb = (t.T + KM_MAX_RED - 1) / KM_MAX_RED;
    t.G = (t.T + t.tpb - 1) / t.tpb;
    return t;
}

// Tiles the register-staged iteration kernels accept: at most 4 cells per
// thread and a halo'd tile of at most (cells-per-thread + 2) positions per
// thread (cells per thread rounded up to 1, 2 or 4).
static inline bool tile_ok(int TI, int TJ) {
    if (TI < 1 || TJ < 1 || TI * TJ > 4 * KM_BLOCK) return false;
    const int c = (TI * TJ + KM_BLOCK - 1) / KM_BLOCK;
    const int cpt = c <= 1 ? 1 : (c <= 2 ? 2 : 4);
    return (TI + 2) * (TJ + 2) <= (cpt + 2) * KM_BLOCK;
}

// A flat (pointwise) launch over n elements: G workgroups, each owning a
// contiguous chunk (a multiple of the workgroup size) of the index range.
struct Flat {
    int64_t n, chunk;
    int G;
};

__host__ __device__ static inline Flat make_flat(int64_t n) {
    Flat f;
    f.n = n;
    int64_t g = (n + 255) / 256;
    if (g > KM_MAX_RED) g = KM_MAX_RED;
    if (g < 1) g = 1;
    int64_t chunk = (n + g - 1) / g;
    chunk = (chunk + 255) / 256 * 256;
    if (chunk < 256) chunk = 256;
    f.chunk = chunk;
    f.G = (int)((n + chunk - 1) / chunk);
    if (f.G < 1) f.G = 1;
    return f;
}

struct GraphKey {
    std::vector<uintptr_t> v;
    bool operator<(const GraphKey& o) const { return v < o.v; }
};

struct km_net;   // learned preconditioner buffers (km_unet.hip)
struct km_batch; // batched cases (km_batch.hip)
struct km_coarse; // coarse level of the two-level preconditioner (km_coarse.hip)
struct km_mg;     // multigrid hierarchy of the V-cycle preconditioner (km_mg.hip)

struct km_ctx {
    int device = 0;
    hipStream_t stream = nullptr;       // caller's stream: all work is enqueued here
    hipStream_t cap_stream = nullptr;   // private stream used only to capture graphs
    hipEvent_t ev[2] = {nullptr, nullptr};
    std::string err;
    // workspace arena (grow only)
    void* ws = nullptr;
    size_t ws_bytes = 0;
    uint64_t ws_gen = 0;
    KState* st = nullptr;          // device scalars (several slots)
    double* partials = nullptr;    // per-workgroup partial sums
    double* dres = nullptr;        // small device results
    KState* h_st = nullptr;        // pinned host mirrors of st[0] (two polling slots)
    volatile int* h_flag = nullptr;   // coherent pinned host words the fused solves write (KM_HW per system)
    int* d_flag = nullptr;         // the same words as the device addresses them
    double* h_res = nullptr;       // pinned host results
    std::map<std::tuple<int, int, int>, std::pair<int, int>> tiles;   // (NX, NY, prec) -> (TI, TJ)
    std::map<std::tuple<int, int, int>, std::pair<int, int>> tile_override;
    std::map<GraphKey, hipGraphExec_t> graphs;
    int autotune = 1;              // time tile candidates on the first solve of a shape
    int plain_totals = -1;         // km_set_plain_totals: -1 by size, 0 never, 1 always
    int mom_sweeps = 0;            // km_set_momentum_sweeps: 0 by size and precision, else 1 or 2
    std::set<std::tuple<int, int, int>> tuned;
    km_net* net = nullptr;
    km_batch* batch = nullptr;
    km_coarse* coarse = nullptr;   // the coarse level in use (one of coarse_pool once a fused solve ran)
    std::vector<km_coarse*> coarse_pool;   // the fused solves' levels, one per coefficient array
    uint64_t coarse_tick = 0;
    void* blas = nullptr;          // rocBLAS handle (the coarse levels' Newton-Schulz GEMM)
    void* scratch = nullptr;       // device address of this device's no-allocation scratch array
    km_mg* mg = nullptr;           // V-cycle hierarchy (KM_FLAG_MG)
};

void km_batch_free(km_ctx* ctx);

int km_fail(km_ctx* ctx, int code, const char* fmt, ...);

#define KM_HIP(ctx, expr)                                                               \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return km_fail((ctx), KM_EHIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

#define KM_CHECK_LAUNCH(ctx) KM_HIP(ctx, hipGetLastError())

// Graph captures and the HIP calls that may synchronise the device or
// allocate (hipMalloc, hipFree, synchronous copies and fills, stream and event
// waits, graph destruction, rocBLAS set-up and GEMMs) are ordered by one
// process-wide lock.  A capture holds it exclusively from
// hipStreamBeginCapture to hipGraphInstantiate; the other calls hold it
// shared while they run.  hipFree and hipMalloc synchronise the whole device,
// and a device-wide synchronisation while another host thread's stream is
// capturing invalidates that capture: hipStreamEndCapture then fails with
// "operation failed due to a previous error during capture".  Contexts driven
// from several host threads at once (one per worker, each on its own stream)
// therefore never overlap a capture with such a call.
std::shared_mutex& km_host_mutex();
using KmCaptureLock = std::unique_lock<std::shared_mutex>;
using KmHostLock = std::shared_lock<std::shared_mutex>;

#define KM_HIP_SYNC(ctx, expr)                     \
    do {                                           \
        KmHostLock lk_(km_host_mutex());           \
        KM_HIP(ctx, expr);                         \
    } while (0)

// Capture the work `body` enqueues on the stream it is given (the context's
// private capture stream) and instantiate it, under the exclusive lock.
int km_capture(km_ctx* ctx, const std::function<int(hipStream_t)>& body, hipGraphExec_t* out);

#define KM_REQUIRE(ctx, cond, ...)                     \
    do {                                               \
        if (!(cond)) return km_fail((ctx), KM_EINVAL, __VA_ARGS__); \
    } while (0)

// Device address of the no-allocation scratch array on ctx's device, into
// ctx->scratch (km_krylov.hip; called once per context at creation).
int km_resolve_scratch(km_ctx* ctx);

// Workspace management (km_ctx.hip)
int km_ensure_workspace(km_ctx* ctx, size_t vec_bytes);
void* km_ws_vectors(km_ctx* ctx);
int km_choose_tile(km_ctx* ctx, int NX, int NY, int prec, int* TI, int* TJ);
void km_drop_graphs(km_ctx* ctx);
// The coherent host words (km_ctx::h_flag / d_flag): KM_HW ints per system
// of a fused solve, which its stopping test writes: [0] done, [1] the
// iteration reached (the last one once done), [2] status, [4..5] r0 and
// [6..7] <r,r> at the stop (doubles); then, from KM_HW_RES, the doubles of
// km_residual_async.
#define KM_HW 16
#define KM_HW_SYSTEMS 4
#define KM_HW_RES (KM_HW * KM_HW_SYSTEMS)
#define KM_HW_WORDS (KM_HW_RES + 2 * 16)

// r = b - A x, with ||r||^2 written by the device into the host words'
// result slot `slot` (read it with km_mapped_res after synchronising; km_krylov.hip)
int km_residual_async(km_ctx* ctx, int NX, int NY, const double* coef, const double* b, const double* x, double* r,
                      int slot = 0);
inline double km_mapped_res(const km_ctx* ctx, int slot) {
    return reinterpret_cast<const volatile double*>(ctx->h_flag + KM_HW_RES)[slot];
}

// Poll loop: enqueue chunks of iterations (a captured graph, or direct
// launches), two chunks in flight, until the device sets ``done``.  The chunk
// length is a multiple of KM_RING so that every chunk starts at phase 0.
template <typename F>
inline int run_chunks(km_ctx* ctx, KState* st, GraphKey key, bool use_graph, int chunk, int max_iter, F enqueue,
               KState* final_state) {
    hipGraphExec_t exec = nullptr;
    if (use_graph) {
        key.v.push_back((uintptr_t)chunk);
        key.v.push_back((uintptr_t)ctx->ws_gen);
        auto it = ctx->graphs.find(key);
        if (it != ctx->graphs.end()) {
            exec = it->second;
        } else {
            const int rc = km_capture(
                ctx,
                [&](hipStream_t q) {
                    for (int c = 0; c < chunk; ++c) enqueue(q, c % KM_RING);
                    return KM_OK;
                },
                &exec);
            if (rc) return rc;
            ctx->graphs[key] = exec;
        }
    }
    auto launch_chunk = [&]() -> int {
        if (exec) {
            KM_HIP(ctx, hipGraphLaunch(exec, ctx->stream));
        } else {
            for (int c = 0; c < chunk; ++c) enqueue(ctx->stream, c % KM_RING);
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
        KM_HIP_SYNC(ctx, hipEventSynchronize(ctx->ev[(c - 1) & 1]));
        if (ctx->h_st[(c - 1) & 1].done) break;
    }
    KM_HIP(ctx, hipMemcpyAsync(&ctx->h_st[2], st, sizeof(KState), hipMemcpyDeviceToHost, ctx->stream));
    KM_HIP_SYNC(ctx, hipStreamSynchronize(ctx->stream));
    *final_state = ctx->h_st[2];
    return KM_OK;
}


// run_chunks with a chunk schedule: chunks of `big` iterations while the
// iterations enqueued stay `expect` - 2 or more short of `expect` (a guess of
// the solve's length, 0: none), then chunks of `small` (a divisor of KM_RING),
// so that when the device sets ``done`` few no-op launches remain in flight.
// A chunk of `big` starts at phase 0; a chunk of `small` at any multiple of
// `small`.  Graphs are cached per (key, length, first phase).
template <typename F>
inline int run_chunks_sched(km_ctx* ctx, KState* st, GraphKey key, bool use_graph, int big, int small, int expect,
                            int max_iter, F enqueue, KState* final_state) {
    auto graph_for = [&](int len, int ph0, hipGraphExec_t* out) -> int {
        GraphKey k = key;
        k.v.push_back((uintptr_t)len);
        k.v.push_back((uintptr_t)ph0);
        k.v.push_back((uintptr_t)ctx->ws_gen);
        auto it = ctx->graphs.find(k);
        if (it != ctx->graphs.end()) {
            *out = it->second;
            return KM_OK;
        }
        hipGraphExec_t exec = nullptr;
        const int rc = km_capture(
            ctx,
            [&](hipStream_t q) {
                for (int c = 0; c < len; ++c) enqueue(q, (ph0 + c) % KM_RING);
                return KM_OK;
            },
            &exec);
        if (rc) return rc;
        ctx->graphs[k] = exec;
        *out = exec;
        return KM_OK;
    };
    int64_t issued = 0;
    auto launch_chunk = [&]() -> int {
        const int ph0 = (int)(issued % KM_RING);
        const bool near = expect > 0 && issued + big > (int64_t)expect - 2;
        const int len = (near || ph0 != 0) ? small : big;
        if (use_graph) {
            hipGraphExec_t exec = nullptr;
            if (int rc = graph_for(len, ph0, &exec)) return rc;
            KM_HIP(ctx, hipGraphLaunch(exec, ctx->stream));
        } else {
            for (int c = 0; c < len; ++c) enqueue(ctx->stream, (ph0 + c) % KM_RING);
            KM_CHECK_LAUNCH(ctx);
        }
        issued += len;
        return KM_OK;
    };
    // a device-side cap ends every solve; the host cap only guards against a
    // state that never reports done because a launch failed
    const int64_t host_cap = (int64_t)max_iter / small + 4;
    if (int rc = launch_chunk()) return rc;
    KM_HIP(ctx, hipMemcpyAsync(&ctx->h_st[0], st, sizeof(KState), hipMemcpyDeviceToHost, ctx->stream));
    KM_HIP(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    for (int64_t c = 1;; ++c) {
        if (c > host_cap) return km_fail(ctx, KM_EHIP, "solver did not report completion after %lld chunks", (long long)c);
        if (int rc = launch_chunk()) return rc;
        KM_HIP(ctx, hipMemcpyAsync(&ctx->h_st[c & 1], st, sizeof(KState), hipMemcpyDeviceToHost, ctx->stream));
        KM_HIP(ctx, hipEventRecord(ctx->ev[c & 1], ctx->stream));
        KM_HIP_SYNC(ctx, hipEventSynchronize(ctx->ev[(c - 1) & 1]));
        if (ctx->h_st[(c - 1) & 1].done) break;
    }
    KM_HIP(ctx, hipMemcpyAsync(&ctx->h_st[2], st, sizeof(KState), hipMemcpyDeviceToHost, ctx->stream));
    KM_HIP_SYNC(ctx, hipStreamSynchronize(ctx->stream));
    *final_state = ctx->h_st[2];
    return KM_OK;
}


// run_chunks_sched for the fused solves, whose stopping-test launch writes
// the iteration it has reached and ``done`` into ctx's coherent host words
// (h_flag): the host keeps at most one chunk queued beyond the iterations the
// device has reached and stops launching as soon as it reports done, with no
// copy or event between chunks.  Chunk lengths as run_chunks_sched.  `tail`
// (may be null) enqueues more work after the last chunk, before the one
// synchronisation at the end.  nsys independent systems iterate in the same
// launches (system s: host words KM_HW s onwards): the solve ends when all
// are done, and the slowest one still running sets the pace; final_state[s]
// is filled from the words the stopping test wrote.
template <typename F>
inline int run_chunks_flag(km_ctx* ctx, GraphKey key, bool use_graph, int big, int small, int expect,
                           int max_iter, F enqueue, KState* final_state, const std::function<int()>* tail = nullptr,
                           int nsys = 1) {
    auto graph_for = [&](int len, int ph0, hipGraphExec_t* out) -> int {
        GraphKey k = key;
        k.v.push_back((uintptr_t)len);
        k.v.push_back((uintptr_t)ph0);
        k.v.push_back((uintptr_t)ctx->ws_gen);
        auto it = ctx->graphs.find(k);
        if (it != ctx->graphs.end()) {
            *out = it->second;
            return KM_OK;
        }
        hipGraphExec_t exec = nullptr;
        const int rc = km_capture(
            ctx,
            [&](hipStream_t q) {
                for (int c = 0; c < len; ++c) enqueue(q, (ph0 + c) % KM_RING);
                return KM_OK;
            },
            &exec);
        if (rc) return rc;
        ctx->graphs[k] = exec;
        *out = exec;
        return KM_OK;
    };
    volatile int* hf = ctx->h_flag;
    if (nsys < 1 || nsys > KM_HW_SYSTEMS) return km_fail(ctx, KM_EINVAL, "%d systems in one solve", nsys);
    for (int q = 0; q < nsys; ++q) {
        hf[KM_HW * q] = 0;
        hf[KM_HW * q + 1] = -1;
    }
    std::atomic_thread_fence(std::memory_order_seq_cst);
    auto all_done = [&]() {
        for (int q = 0; q < nsys; ++q)
            if (!hf[KM_HW * q]) return false;
        return true;
    };
    auto slowest = [&]() {
        int m = 1 << 30;
        for (int q = 0; q < nsys; ++q)
            if (!hf[KM_HW * q] && hf[KM_HW * q + 1] < m) m = hf[KM_HW * q + 1];
        return m == (1 << 30) ? -1 : m;
    };
    int64_t issued = 0;
    const int64_t cap = (int64_t)max_iter + 2 * big;   // the device's own cap ends the solve before this
    int seen = -1;
    auto t_seen = std::chrono::steady_clock::now();
    int spins = 0;
    while (!all_done()) {
        const int reached = slowest();
        const int ph0 = (int)(issued % KM_RING);
        const bool near = expect > 0 && issued + big > (int64_t)expect - 2;
        const int len = (near || ph0 != 0) ? small : big;
        if (issued < cap && issued - (reached + 1) <= len) {
            if (use_graph) {
                hipGraphExec_t exec = nullptr;
                if (int rc = graph_for(len, ph0, &exec)) return rc;
                KM_HIP(ctx, hipGraphLaunch(exec, ctx->stream));
            } else {
                for (int c = 0; c < len; ++c) enqueue(ctx->stream, (ph0 + c) % KM_RING);
                KM_CHECK_LAUNCH(ctx);
            }
            issued += len;
            continue;
        }
        // waiting: a launch error or a stream that stops making progress ends the solve
        if (reached != seen) {
            seen = reached;
            t_seen = std::chrono::steady_clock::now();
        } else if (++spins % 4096 == 0) {
            const hipError_t q = hipStreamQuery(ctx->stream);
            if (q == hipSuccess) {
                // `reached` was read before the query: the device may have run
                // every queued iteration since and be waiting for the next chunk
                std::atomic_thread_fence(std::memory_order_acquire);
                if (all_done() || slowest() != reached) continue;
                // the queue drained with no progress and no reported stop
                break;
            }
            if (q != hipSuccess && q != hipErrorNotReady)
                return km_fail(ctx, KM_EHIP, "solver stream: %s", hipGetErrorString(q));
            if (std::chrono::steady_clock::now() - t_seen > std::chrono::seconds(120))
                return km_fail(ctx, KM_EHIP, "solver made no progress for 120 s");
        }
        std::this_thread::yield();
    }
    // the final state comes from the host words the stopping test wrote
    if (tail)
        if (int rc = (*tail)()) return rc;
    KM_HIP_SYNC(ctx, hipStreamSynchronize(ctx->stream));
    std::atomic_thread_fence(std::memory_order_acquire);
    for (int q = 0; q < nsys; ++q) {
        volatile int* w = hf + KM_HW * q;
        KState& f = final_state[q];
        f = KState{};
        f.done = w[0];
        f.iter = w[1];
        f.status = w[2];
        f.r0 = *reinterpret_cast<volatile double*>(w + 4);
        f.rTr = *reinterpret_cast<volatile double*>(w + 6);
        if (!f.done) return km_fail(ctx, KM_EHIP, "solver stopped without reporting completion");
    }
    return KM_OK;
}


// Grid, physics and boundary data of one case on the device (km_grid, flattened).
struct GridD {
    int nx, ny, east, pin;
    double dx, dy, rho, mu, dt, alpha_u, alpha_p, alpha_m;
    double bcw_n, bcw_t, bce_n, bce_t, bcn_n, bcn_t, bcs_n, bcs_t;
};

inline GridD to_dev(const km_grid* g) {
    GridD d;
    d.nx = g->nx; d.ny = g->ny; d.east = g->east; d.pin = g->pin_pcor;
    d.dx = g->dx; d.dy = g->dy; d.rho = g->rho; d.mu = g->mu; d.dt = g->dt;
    d.alpha_u = g->alpha_u; d.alpha_p = g->alpha_p; d.alpha_m = g->alpha_m;
    d.bcw_n = g->bc[0][0]; d.bcw_t = g->bc[0][1];
    d.bce_n = g->bc[1][0]; d.bce_t = g->bc[1][1];
    d.bcn_n = g->bc[2][0]; d.bcn_t = g->bc[2][1];
    d.bcs_n = g->bc[3][0]; d.bcs_t = g->bc[3][1];
    return d;
}


static inline int ceil_div_i64(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// One case of a batched SIMPLE step (device table entry): its grid data and
// the device fields of a SIMPLESolver (simple.py:27-47).
struct BCase {
    GridD g;
    double *u, *v, *p, *u0, *v0, *umid, *vmid, *pcor, *mdiv;
    double *coef_u, *coef_v, *coef_p, *b_u, *b_v, *b_p;
    const uint8_t* m;
};

// Stages of a batched SIMPLE step that run as one launch over every case.
enum BatchStage { BS_ASM_U, BS_ASM_V, BS_GHOSTS, BS_BLEND, BS_OUTFLOW, BS_ASM_P, BS_CORRECT, BS_MDIV, BS_TIME, BS_COUNT };
// One system of a batched BiCGSTAB solve (device table entry).
struct BSys {
    Tile tl;                 // this system's tiles (one tile shape for the whole batch)
    Flat fl;                 // flat geometry of the init kernel
    const double* coef;
    const double* b;
    double* x;
    double* ws;              // 8 n values: r, r~, p (2 n), v (2 n), s, t
    KState* st;              // this system's scalars
    double* P;               // this system's partials (12 KM_MAX_RED values)
};

// Plain FP64 BiCGSTAB of n independent systems, every iteration one launch of
// each fused kernel over all systems' tiles.  sys / d_sys: host and device
// copies of the table (tl and fl filled by the caller for tile TI x TJ);
// tstarts / fstarts: device prefix sums (n + 1) of tl.G and fl.G; agg: the
// aggregate state the host polls; states (host, n) receive each system's
// final scalars (the systems' states are contiguous: sys[c].st = sys[0].st + c).
int batch_bicgstab(km_ctx* ctx, int n, const BSys* sys, const BSys* d_sys, const int* tstarts, const int* fstarts,
                   int64_t tblocks, int64_t fblocks, int TI, int TJ, double eps, int max_iter, bool use_graph,
                   KState* agg, KState* states);

int64_t batch_stage_blocks(int stage, const GridD& g);
int batch_stage_launch(km_ctx* ctx, int stage, const BCase* cases, const int* starts, int n, int64_t blocks,
                       int parts, double* P, hipStream_t s);

// ---------------------------------------------------------------- device side

// XCD-aware workgroup remap: hardware workgroup b runs on XCD b % 8; give each
// XCD a contiguous range of logical tiles so neighbouring rows of a stencil
// share that XCD's L2.  Bijective on [0, G).
// Case of workgroup bx in a batched launch: the last c with starts[c] <= bx
// (starts: n + 1 ascending workgroup offsets).
__device__ __forceinline__ int find_case(const int* __restrict__ starts, int n, int bx) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= bx) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int xcd_remap(int b, int G) {
    int xcd = b % KM_NXCD, local = b / KM_NXCD;
    int per = G / KM_NXCD, rem = G % KM_NXCD;
    return xcd < rem ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// Sum over a 256-thread workgroup in a fixed order; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* sh) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    return r;
}

__device__ __forceinline__ double block_max(double v, double* sh) {
    v = wave_max(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) r = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
    return r;
}


// Sum of v over the workgroup, in a fixed order, returned to every thread.
__device__ __forceinline__ double block_sum_all(double v, double* sh) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    const double r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    __syncthreads();
    return r;
}

// Totals of the partials P[q * KM_MAX_RED + k], k < G, for q < NV, returned
// to every thread of the workgroup.  Fixed order: eight strided accumulators
// per thread (eight loads in flight), a fixed tree, then the workgroup sum.
template <int NV>
__device__ __forceinline__ void sum_partials(const double* __restrict__ P, int G, double (&out)[NV], double* sh) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        const double* base = P + (size_t)q * KM_MAX_RED;
        double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int k = threadIdx.x; k < G; k += 8 * KM_BLOCK) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = (k + u * KM_BLOCK < G) ? base[k + u * KM_BLOCK] : 0.0;
#pragma unroll
            for (int u = 0; u < 8; ++u) acc[u] += v[u];
        }
        const double a = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
        out[q] = block_sum_all(a, sh);
    }
}

// sum_partials with every plane's loads issued before any is reduced and the
// NV workgroup sums sharing one pair of barriers: the same totals, bit for
// bit, for kernels that reduce many planes before short work (it holds 8 NV
// values per thread in registers, which the stencil kernels cannot spare).
// Plane q starts at P + q * stride.
template <int NV>
__device__ __forceinline__ void sum_partials_wide(const double* __restrict__ P, int G, double (&out)[NV],
                                                  int stride = KM_MAX_RED) {
    __shared__ double shp[4 * NV];
    double acc[NV][8];
#pragma unroll
    for (int q = 0; q < NV; ++q)
#pragma unroll
        for (int u = 0; u < 8; ++u) acc[q][u] = 0.0;
    for (int k = threadIdx.x; k < G; k += 8 * KM_BLOCK) {
        double v[NV][8];
#pragma unroll
        for (int q = 0; q < NV; ++q)
#pragma unroll
            for (int u = 0; u < 8; ++u)
                v[q][u] = (k + u * KM_BLOCK < G) ? P[(size_t)q * stride + k + u * KM_BLOCK] : 0.0;
#pragma unroll
        for (int q = 0; q < NV; ++q)
#pragma unroll
            for (int u = 0; u < 8; ++u) acc[q][u] += v[q][u];
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double a = ((acc[q][0] + acc[q][1]) + (acc[q][2] + acc[q][3])) +
                   ((acc[q][4] + acc[q][5]) + (acc[q][6] + acc[q][7]));
        a = wave_sum(a);
        if (q == 0) __syncthreads();   // earlier readers of shp are done
        if (lane == 0) shp[q * 4 + w] = a;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; ++q) out[q] = (shp[q * 4] + shp[q * 4 + 1]) + (shp[q * 4 + 2] + shp[q * 4 + 3]);
    __syncthreads();
}

// sum_partials_wide over planes of their own lengths G[q] at P[q]: each total
// has the bits sum_partials_wide<1>(P[q], G[q]) gives.
template <int NV>
__device__ __forceinline__ void sum_partials_each(const double* const (&P)[NV], const int (&G)[NV], double (&out)[NV]) {
    __shared__ double shp[4 * NV];
    int Gm = 0;
#pragma unroll
    for (int q = 0; q < NV; ++q) Gm = G[q] > Gm ? G[q] : Gm;
    double acc[NV][8];
#pragma unroll
    for (int q = 0; q < NV; ++q)
#pragma unroll
        for (int u = 0; u < 8; ++u) acc[q][u] = 0.0;
    for (int k = threadIdx.x; k < Gm; k += 8 * KM_BLOCK) {
        double v[NV][8];
#pragma unroll
        for (int q = 0; q < NV; ++q)
#pragma unroll
            for (int u = 0; u < 8; ++u) v[q][u] = (k + u * KM_BLOCK < G[q]) ? P[q][k + u * KM_BLOCK] : 0.0;
#pragma unroll
        for (int q = 0; q < NV; ++q)
#pragma unroll
            for (int u = 0; u < 8; ++u) acc[q][u] += v[q][u];
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double a = ((acc[q][0] + acc[q][1]) + (acc[q][2] + acc[q][3])) +
                   ((acc[q][4] + acc[q][5]) + (acc[q][6] + acc[q][7]));
        a = wave_sum(a);
        if (q == 0) __syncthreads();   // earlier readers of shp are done
        if (lane == 0) shp[q * 4 + w] = a;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; ++q) out[q] = (shp[q * 4] + shp[q * 4 + 1]) + (shp[q * 4 + 2] + shp[q * 4 + 3]);
    __syncthreads();
}
