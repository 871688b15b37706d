// Shared definitions of the gfx950 SIMPLE kernels: context, error handling,
// wave64 reductions, the deterministic last-workgroup reduction and the
// XCD-aware tile mapping.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/karman.h"

#define KM_BLOCK 256
#define KM_NXCD 8           // MI355X: 8 XCDs, workgroups dispatched round-robin
#define KM_MAX_PARTIALS (1 << 18)

// Device-resident Krylov scalars.  Written only by the last workgroup of a
// reducing kernel; read by the following kernels and copied to the host when
// the solver polls for completion.
struct KState {
    double rho, rho_1, alpha, omega, beta;
    double r0;       // initial r.r
    double rTr;      // latest r.r
    double eps;      // relative tolerance on sqrt(r.r / r0)
    double aux[8];   // dot results of helper kernels
    int32_t iter, status, done, max_iter;
    uint32_t counter[8];   // arrival counters of the last-workgroup reductions
};

// Geometry of a 2-D tiled stencil launch: TI rows x TJ columns per workgroup.
struct Tile {
    int NX, NY, TI, TJ, tiles_j, G;
};

static inline Tile make_tile(int NX, int NY, int TI, int TJ) {
    Tile t;
    t.NX = NX; t.NY = NY; t.TI = TI; t.TJ = TJ;
    t.tiles_j = (NY + TJ - 1) / TJ;
    int tiles_i = (NX + TI - 1) / TI;
    t.G = tiles_i * t.tiles_j;
    return t;
}

struct GraphKey {
    std::vector<uintptr_t> v;
    bool operator<(const GraphKey& o) const { return v < o.v; }
};

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
    double* h_res = nullptr;       // pinned host results
    std::map<std::tuple<int, int, int>, std::pair<int, int>> tiles;   // (NX, NY, prec) -> (TI, TJ)
    std::map<std::tuple<int, int, int>, std::pair<int, int>> tile_override;
    std::map<GraphKey, hipGraphExec_t> graphs;
};

int km_fail(km_ctx* ctx, int code, const char* fmt, ...);

#define KM_HIP(ctx, expr)                                                               \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return km_fail((ctx), KM_EHIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

#define KM_CHECK_LAUNCH(ctx) KM_HIP(ctx, hipGetLastError())

#define KM_REQUIRE(ctx, cond, ...)                     \
    do {                                               \
        if (!(cond)) return km_fail((ctx), KM_EINVAL, __VA_ARGS__); \
    } while (0)

// Workspace management (km_ctx.hip)
int km_ensure_workspace(km_ctx* ctx, size_t vec_bytes);
void* km_ws_vectors(km_ctx* ctx);
int km_choose_tile(km_ctx* ctx, int NX, int NY, int prec, int* TI, int* TJ);
void km_drop_graphs(km_ctx* ctx);

static inline int ceil_div_i64(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// ---------------------------------------------------------------- device side

// XCD-aware workgroup remap: hardware workgroup b runs on XCD b % 8; give each
// XCD a contiguous range of logical tiles so neighbouring rows of a stencil
// share that XCD's L2.  Bijective on [0, G).
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

__device__ __forceinline__ void store_agent(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double load_agent(const double* p) {
    return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Publish NV per-workgroup partials and report whether this workgroup arrived
// last.  partials is laid out [slot][G]; every thread of the workgroup gets
// the answer.  The arrival counter is reset by the last workgroup.
//
// Only agent-scope atomics carry data between workgroups here: the partials
// are agent-coherent stores and the last workgroup reads them back with
// agent-coherent loads.  The store therefore only has to be complete before
// the counter increment, which the explicit s_waitcnt guarantees.  An
// acquire/release fence at agent scope would be correct too, but on gfx950 it
// writes back and invalidates the XCD's whole L2 in every workgroup, which
// cost ~30 us per reduction kernel.
template <int NV>
__device__ __forceinline__ bool publish_partials(const double (&vals)[NV], int slot_index, int G,
                                                 double* partials, uint32_t* counter) {
    __shared__ bool last;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) store_agent(&partials[(size_t)q * G + slot_index], vals[q]);
        __builtin_amdgcn_s_waitcnt(0);   // vmcnt(0) expcnt(0) lgkmcnt(0): the stores above are done
        uint32_t prev = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        last = (prev == (uint32_t)(G - 1));
        if (last) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    return last;
}

// In the last workgroup: deterministic sum of partials[q*G .. q*G+G) for each
// q (fixed strided order, then the fixed block tree).  Result in thread 0.
template <int NV>
__device__ __forceinline__ void finish_partials(const double* partials, int G, double (&out)[NV], double* sh) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double a = 0.0;
        for (int k = threadIdx.x; k < G; k += KM_BLOCK) a += load_agent(&partials[(size_t)q * G + k]);
        out[q] = block_sum(a, sh);
    }
}
