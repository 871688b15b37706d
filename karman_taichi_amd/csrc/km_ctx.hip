// Context, error reporting and workspace arena of libkarman_hip.so.
#include "km_common.h"
#include "km_unet.h"
#include "km_coarse.h"
#include "km_mg.h"

#include <cstdlib>
#include <cstring>

int km_fail(km_ctx* ctx, int code, const char* fmt, ...) {
    if (ctx) {
        char buf[1024];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}

std::shared_mutex& km_host_mutex() {
    static std::shared_mutex m;
    return m;
}

int km_capture(km_ctx* ctx, const std::function<int(hipStream_t)>& body, hipGraphExec_t* out) {
    KmCaptureLock lk(km_host_mutex());
    hipGraph_t graph = nullptr;
    KM_HIP(ctx, hipStreamBeginCapture(ctx->cap_stream, hipStreamCaptureModeThreadLocal));
    const int rc = body(ctx->cap_stream);
    const hipError_t le = hipGetLastError();
    const hipError_t ce = hipStreamEndCapture(ctx->cap_stream, &graph);
    if (rc != KM_OK || le != hipSuccess || ce != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        if (rc != KM_OK) return rc;
        return km_fail(ctx, KM_EHIP, "graph capture failed: %s / %s", hipGetErrorString(le), hipGetErrorString(ce));
    }
    const hipError_t ie = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) return km_fail(ctx, KM_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(ie));
    return KM_OK;
}

void km_drop_graphs(km_ctx* ctx) {
    KmHostLock lk(km_host_mutex());
    for (auto& kv : ctx->graphs) (void)hipGraphExecDestroy(kv.second);
    ctx->graphs.clear();
}

// Fixed header of the arena: KState slots, partials, small results.  Vector
// storage follows it.  Growing the arena frees the old one, so every captured
// graph (which holds raw pointers) is dropped.
static size_t header_bytes() {
    size_t h = 8 * sizeof(KState);
    h = (h + 255) & ~size_t(255);
    h += KM_MAX_PARTIALS * sizeof(double);
    h += 64 * sizeof(double);
    return (h + 255) & ~size_t(255);
}

int km_ensure_workspace(km_ctx* ctx, size_t vec_bytes) {
    const size_t need = header_bytes() + ((vec_bytes + 255) & ~size_t(255));
    if (ctx->ws && ctx->ws_bytes >= need) return KM_OK;
    KM_HIP(ctx, hipSetDevice(ctx->device));
    if (ctx->ws) {
        KM_HIP_SYNC(ctx, hipStreamSynchronize(ctx->stream));
        km_drop_graphs(ctx);
        KM_HIP_SYNC(ctx, hipFree(ctx->ws));
        ctx->ws = nullptr;
        ctx->ws_bytes = 0;
    }
    size_t bytes = need + need / 4;   // headroom so small growth does not thrash
    hipError_t e;
    {
        KmHostLock lk(km_host_mutex());
        e = hipMalloc(&ctx->ws, bytes);
    }
    if (e != hipSuccess) {
        ctx->ws = nullptr;
        return km_fail(ctx, KM_ENOMEM, "workspace hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    ctx->ws_bytes = bytes;
    ctx->ws_gen++;
    // zero the header once, ordered on the context's stream before any use
    KM_HIP(ctx, hipMemsetAsync(ctx->ws, 0, header_bytes(), ctx->stream));
    char* base = static_cast<char*>(ctx->ws);
    size_t off = 0;
    ctx->st = reinterpret_cast<KState*>(base + off);
    off += (8 * sizeof(KState) + 255) & ~size_t(255);
    ctx->partials = reinterpret_cast<double*>(base + off);
    off += KM_MAX_PARTIALS * sizeof(double);
    ctx->dres = reinterpret_cast<double*>(base + off);
    return KM_OK;
}

void* km_ws_vectors(km_ctx* ctx) { return static_cast<char*>(ctx->ws) + header_bytes(); }

extern "C" {

int km_create(km_ctx** out, int device, void* hip_stream) {
    if (!out) return KM_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return KM_EHIP;
    if (device < 0 || device >= n) return KM_EINVAL;
    km_ctx* ctx = new km_ctx();
    ctx->device = device;
    ctx->stream = static_cast<hipStream_t>(hip_stream);
    int rc = KM_OK;
    {
        KmHostLock lk(km_host_mutex());
        void* flag = nullptr;
        if (hipSetDevice(device) != hipSuccess) {
            rc = KM_EHIP;
        } else if (hipHostMalloc(reinterpret_cast<void**>(&ctx->h_st), 4 * sizeof(KState), hipHostMallocDefault) !=
                       hipSuccess ||
                   hipHostMalloc(reinterpret_cast<void**>(&ctx->h_res), 128 * sizeof(double),
                                 hipHostMallocDefault) != hipSuccess) {
            rc = KM_ENOMEM;
        } else if (hipHostMalloc(&flag, KM_HW_WORDS * sizeof(int), hipHostMallocCoherent | hipHostMallocMapped) !=
                       hipSuccess ||
                   hipHostGetDevicePointer(reinterpret_cast<void**>(&ctx->d_flag), flag, 0) != hipSuccess) {
            if (flag) (void)hipHostFree(flag);
            rc = KM_ENOMEM;
        } else {
            ctx->h_flag = static_cast<volatile int*>(flag);
            if (hipStreamCreateWithFlags(&ctx->cap_stream, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&ctx->ev[0], hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&ctx->ev[1], hipEventDisableTiming) != hipSuccess)
                rc = KM_EHIP;
        }
    }
    // the no-allocation entry point's scratch address on this device
    if (rc == KM_OK && km_resolve_scratch(ctx) != KM_OK) rc = KM_EHIP;
    if (rc != KM_OK) {
        ctx->stream = nullptr;   // nothing was enqueued
        km_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return KM_OK;
}

int km_destroy(km_ctx* ctx) {
    if (!ctx) return KM_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) {
        KmHostLock lk(km_host_mutex());
        (void)hipStreamSynchronize(ctx->stream);
    }
    km_drop_graphs(ctx);
    km_net_free(ctx);
    km_batch_free(ctx);
    km_coarse_free(ctx);
    km_mg_free(ctx);
    KmHostLock lk(km_host_mutex());
    if (ctx->ws) (void)hipFree(ctx->ws);
    if (ctx->h_st) (void)hipHostFree(ctx->h_st);
    if (ctx->h_res) (void)hipHostFree(ctx->h_res);
    if (ctx->h_flag) (void)hipHostFree(const_cast<int*>(ctx->h_flag));
    if (ctx->cap_stream) (void)hipStreamDestroy(ctx->cap_stream);
    for (auto& e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    delete ctx;
    return KM_OK;
}

int km_set_stream(km_ctx* ctx, void* hip_stream) {
    if (!ctx) return KM_EINVAL;
    ctx->stream = static_cast<hipStream_t>(hip_stream);
    return KM_OK;
}

const char* km_last_error(const km_ctx* ctx) { return ctx ? ctx->err.c_str() : "NULL context"; }

int km_synchronize(km_ctx* ctx) {
    if (!ctx) return KM_EINVAL;
    KM_HIP_SYNC(ctx, hipStreamSynchronize(ctx->stream));
    return KM_OK;
}

int km_set_tile(km_ctx* ctx, int NX, int NY, int prec, int ti, int tj) {
    if (!ctx) return KM_EINVAL;
    auto key = std::make_tuple(NX, NY, prec);
    if (ti == 0 && tj == 0) {
        ctx->tile_override.erase(key);
        ctx->tiles.erase(key);
        ctx->tuned.erase(key);
        return KM_OK;
    }
    KM_REQUIRE(ctx, tile_ok(ti, tj),
               "tile %dx%d is invalid: at most 1024 cells, and (ti+2)(tj+2) within (cells per thread + 2) x 256",
               ti, tj);
    ctx->tile_override[key] = {ti, tj};
    return KM_OK;
}

int km_set_autotune(km_ctx* ctx, int on) {
    if (!ctx) return KM_EINVAL;
    ctx->autotune = on ? 1 : 0;
    return KM_OK;
}

int km_get_tile(km_ctx* ctx, int NX, int NY, int prec, int* ti, int* tj) {
    if (!ctx || !ti || !tj) return KM_EINVAL;
    return km_choose_tile(ctx, NX, NY, prec, ti, tj);
}

}  // extern "C"
