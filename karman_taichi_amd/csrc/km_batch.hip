// Batched SIMPLE step: many independent channel cases of uneven grid size and
// Reynolds number advanced together.  Every stage of the outer iteration is
// one launch over all cases (each workgroup finds its case in a prefix table
// and runs the single-case kernel body), the three BiCGSTAB solves run all
// systems in the same fused iteration kernels with per-system scalars, and
// the stage sequences between the solves are captured once as hipGraphs.
// One tile shape serves the whole batch; it is chosen by timing candidates on
// the batch's own first systems.
#include "km_common.h"

#include <cstring>

struct km_batch {
    int n = 0;
    std::vector<km_case> cases;        // the uploaded table (change detection)
    char* arena = nullptr;
    size_t bytes = 0;
    BCase* d_cases = nullptr;
    int* d_stage = nullptr;            // BS_COUNT x (n + 1) workgroup prefix sums
    int64_t stage_blocks[BS_COUNT] = {};
    BSys* d_sys[3] = {};
    int* d_tstart[3] = {};
    int* d_fstart[3] = {};
    int64_t tblocks[3] = {}, fblocks[3] = {};
    std::vector<BSys> h_sys[3];
    KState* d_states = nullptr;        // 3 n system states + 3 aggregates
    double* P = nullptr;               // n x 12 KM_MAX_RED partials
    double* vec = nullptr;             // per case 8 x (largest system) work values
    int TI = 0, TJ = 0;
    bool tuned = false;
    std::vector<KState> h_states;
    hipGraphExec_t stage_graph[3] = {nullptr, nullptr, nullptr};
};

namespace {

bool same_table(const km_batch* B, int n, const km_case* cases) {
    return B && B->n == n && std::memcmp(B->cases.data(), cases, sizeof(km_case) * n) == 0;
}

void drop_stage_graphs(km_batch* B) {
    KmHostLock lk(km_host_mutex());
    for (auto& g : B->stage_graph)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
}

int64_t sys_cells(const km_grid& g, int f) {
    const int64_t nx = g.nx, ny = g.ny;
    return f == 0 ? (nx + 3) * (ny + 2) : f == 1 ? (nx + 2) * (ny + 3) : (nx + 2) * (ny + 2);
}

// Fill the solve tables for tile TI x TJ and upload them.
int set_tile(km_ctx* ctx, km_batch* B, int TI, int TJ) {
    const int n = B->n;
    std::vector<int> ts(n + 1), fs(n + 1);
    for (int f = 0; f < 3; ++f) {
        ts[0] = fs[0] = 0;
        for (int c = 0; c < n; ++c) {
            BSys& S = B->h_sys[f][c];
            const km_case& K = B->cases[c];
            const int NX = f == 0 ? K.g.nx + 3 : K.g.nx + 2;
            const int NY = f == 1 ? K.g.ny + 3 : K.g.ny + 2;
            S.tl = make_tile(NX, NY, TI, TJ);
            S.fl = make_flat((int64_t)NX * NY);
            ts[c + 1] = ts[c] + S.tl.G;
            fs[c + 1] = fs[c] + S.fl.G;
        }
        B->tblocks[f] = ts[n];
        B->fblocks[f] = fs[n];
        KM_HIP_SYNC(ctx, hipMemcpy(B->d_sys[f], B->h_sys[f].data(), sizeof(BSys) * n, hipMemcpyHostToDevice));
        KM_HIP_SYNC(ctx, hipMemcpy(B->d_tstart[f], ts.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
        KM_HIP_SYNC(ctx, hipMemcpy(B->d_fstart[f], fs.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
    }
    B->TI = TI;
    B->TJ = TJ;
    km_drop_graphs(ctx);   // captured iteration chunks hold the old geometry
    return KM_OK;
}

size_t a256(size_t v) { return (v + 255) & ~size_t(255); }

// Allocate and upload everything a batch needs (once per case table).
int prepare(km_ctx* ctx, int n, const km_case* cases) {
    km_batch* B = ctx->batch;
    if (same_table(B, n, cases)) return KM_OK;
    KM_HIP(ctx, hipSetDevice(ctx->device));
    KM_HIP_SYNC(ctx, hipStreamSynchronize(ctx->stream));
    if (!B) B = ctx->batch = new km_batch();
    drop_stage_graphs(B);
    km_drop_graphs(ctx);
    if (B->arena) {
        KM_HIP_SYNC(ctx, hipFree(B->arena));
        B->arena = nullptr;
    }
    B->n = n;
    B->cases.assign(cases, cases + n);
    B->tuned = false;
    int64_t vec = 0;
    for (int c = 0; c < n; ++c) {
        int64_t m = 0;
        for (int f = 0; f < 3; ++f) m = std::max(m, sys_cells(cases[c].g, f));
        vec += 8 * m;
    }
    const size_t sz_cases = a256(sizeof(BCase) * n), sz_stage = a256(sizeof(int) * BS_COUNT * (n + 1));
    const size_t sz_sys = a256(sizeof(BSys) * n), sz_starts = a256(sizeof(int) * (n + 1));
    const size_t sz_states = a256(sizeof(KState) * (3 * n + 3));
    const size_t sz_P = a256(sizeof(double) * (size_t)n * 12 * KM_MAX_RED);
    const size_t sz_vec = a256(sizeof(double) * (size_t)vec);
    const size_t total = sz_cases + sz_stage + 3 * (sz_sys + 2 * sz_starts) + sz_states + sz_P + sz_vec;
    void* p = nullptr;
    hipError_t e;
    {
        KmHostLock lk(km_host_mutex());
        e = hipMalloc(&p, total);
    }
    if (e != hipSuccess) return km_fail(ctx, KM_ENOMEM, "batch hipMalloc(%zu) failed: %s", total, hipGetErrorString(e));
    B->arena = static_cast<char*>(p);
    B->bytes = total;
    char* q = B->arena;
    B->d_cases = reinterpret_cast<BCase*>(q); q += sz_cases;
    B->d_stage = reinterpret_cast<int*>(q); q += sz_stage;
    for (int f = 0; f < 3; ++f) {
        B->d_sys[f] = reinterpret_cast<BSys*>(q); q += sz_sys;
        B->d_tstart[f] = reinterpret_cast<int*>(q); q += sz_starts;
        B->d_fstart[f] = reinterpret_cast<int*>(q); q += sz_starts;
    }
    B->d_states = reinterpret_cast<KState*>(q); q += sz_states;
    B->P = reinterpret_cast<double*>(q); q += sz_P;
    B->vec = reinterpret_cast<double*>(q);
    KM_HIP_SYNC(ctx, hipMemset(B->d_states, 0, sz_states));
    // case table and stage prefix sums
    std::vector<BCase> hc(n);
    std::vector<int> st(BS_COUNT * (n + 1));
    for (int s = 0; s < BS_COUNT; ++s) st[s * (n + 1)] = 0;
    for (int c = 0; c < n; ++c) {
        const km_case& K = cases[c];
        BCase& C = hc[c];
        C.g = to_dev(&K.g);
        C.u = K.u; C.v = K.v; C.p = K.p; C.u0 = K.u0; C.v0 = K.v0; C.umid = K.u_mid; C.vmid = K.v_mid;
        C.pcor = K.pcor; C.mdiv = K.mdiv; C.coef_u = K.coef_u; C.coef_v = K.coef_v; C.coef_p = K.coef_p;
        C.b_u = K.b_u; C.b_v = K.b_v; C.b_p = K.b_p; C.m = K.mask;
        for (int s = 0; s < BS_COUNT; ++s) {
            const int64_t nb = batch_stage_blocks(s, C.g);
            const int64_t next = st[s * (n + 1) + c] + nb;
            KM_REQUIRE(ctx, next < ((int64_t)1 << 31), "batch stage %d too large", s);
            st[s * (n + 1) + c + 1] = (int)next;
        }
    }
    for (int s = 0; s < BS_COUNT; ++s) B->stage_blocks[s] = st[s * (n + 1) + n];
    KM_HIP_SYNC(ctx, hipMemcpy(B->d_cases, hc.data(), sizeof(BCase) * n, hipMemcpyHostToDevice));
    KM_HIP_SYNC(ctx, hipMemcpy(B->d_stage, st.data(), sizeof(int) * st.size(), hipMemcpyHostToDevice));
    // solve tables: u (coef_u, b_u -> u_mid), v (-> v_mid), p' (-> pcor)
    double* vp = B->vec;
    for (int f = 0; f < 3; ++f) B->h_sys[f].assign(n, BSys{});
    for (int c = 0; c < n; ++c) {
        const km_case& K = cases[c];
        int64_t m = 0;
        for (int f = 0; f < 3; ++f) m = std::max(m, sys_cells(K.g, f));
        for (int f = 0; f < 3; ++f) {
            BSys& S = B->h_sys[f][c];
            S.coef = f == 0 ? K.coef_u : f == 1 ? K.coef_v : K.coef_p;
            S.b = f == 0 ? K.b_u : f == 1 ? K.b_v : K.b_p;
            S.x = f == 0 ? K.u_mid : f == 1 ? K.v_mid : K.pcor;
            S.ws = vp;                      // the three solves run one after another
            S.st = B->d_states + f * n + c;
            S.P = B->P + (size_t)c * 12 * KM_MAX_RED;
        }
        vp += 8 * m;
    }
    B->h_states.assign(n, KState{});
    int TI = 2, TJ = 128;
    if (int rc = set_tile(ctx, B, TI, TJ)) return rc;
    return KM_OK;
}

int stage(km_ctx* ctx, km_batch* B, int s, int parts, hipStream_t q) {
    return batch_stage_launch(ctx, s, B->d_cases, B->d_stage + s * (B->n + 1), B->n, B->stage_blocks[s], parts, B->P,
                              q);
}

// Stage sequences between the solves: 0 = assemble u, v and ghosts; 1 =
// blend, outlet fix, assemble p'; 2 = corrections and mdiv.
int run_stages(km_ctx* ctx, km_batch* B, int which, bool use_graph) {
    auto enqueue = [&](hipStream_t q) -> int {
        if (which == 0) {
            if (int rc = stage(ctx, B, BS_ASM_U, 0, q)) return rc;
            if (int rc = stage(ctx, B, BS_ASM_V, 0, q)) return rc;
            return stage(ctx, B, BS_GHOSTS, 0, q);
        }
        if (which == 1) {
            if (int rc = stage(ctx, B, BS_BLEND, 0, q)) return rc;
            if (int rc = stage(ctx, B, BS_OUTFLOW, 0, q)) return rc;
            return stage(ctx, B, BS_ASM_P, 0, q);
        }
        if (int rc = stage(ctx, B, BS_CORRECT, 3, q)) return rc;
        return stage(ctx, B, BS_MDIV, 0, q);
    };
    if (!use_graph) return enqueue(ctx->stream);
    if (!B->stage_graph[which]) {
        if (int rc = km_capture(ctx, enqueue, &B->stage_graph[which])) return rc;
    }
    KM_HIP(ctx, hipGraphLaunch(B->stage_graph[which], ctx->stream));
    return KM_OK;
}

int solve(km_ctx* ctx, km_batch* B, int f, double eps, int max_iter, bool use_graph) {
    return batch_bicgstab(ctx, B->n, B->h_sys[f].data(), B->d_sys[f], B->d_tstart[f], B->d_fstart[f], B->tblocks[f],
                          B->fblocks[f], B->TI, B->TJ, eps, max_iter, use_graph, B->d_states + 3 * B->n + f,
                          B->h_states.data());
}

// Time 4 + 16 batched iterations (stopping test off) of the systems of solve
// f for each candidate tile and keep the fastest.
int autotune(km_ctx* ctx, km_batch* B, int f) {
    std::vector<std::pair<int, int>> cand;
    for (int tj : {32, 64, 128, 256})
        for (int ti : {1, 2, 4, 8, 16})
            if (tile_ok(ti, tj) && ti * tj >= 128) cand.push_back({ti, tj});
    hipEvent_t e0, e1;
    KM_HIP_SYNC(ctx, hipEventCreate(&e0));
    KM_HIP_SYNC(ctx, hipEventCreate(&e1));
    double best = 1e300;
    std::pair<int, int> pick = {B->TI, B->TJ};
    for (const auto& c : cand) {
        if (int rc = set_tile(ctx, B, c.first, c.second)) return rc;
        KM_HIP(ctx, hipEventRecord(e0, ctx->stream));
        if (int rc = batch_bicgstab(ctx, B->n, B->h_sys[f].data(), B->d_sys[f], B->d_tstart[f], B->d_fstart[f],
                                    B->tblocks[f], B->fblocks[f], c.first, c.second, 0.0, 16, false,
                                    B->d_states + 3 * B->n + f, B->h_states.data()))
            return rc;
        KM_HIP(ctx, hipEventRecord(e1, ctx->stream));
        KM_HIP_SYNC(ctx, hipEventSynchronize(e1));
        float ms = 0.f;
        KM_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
            best = ms;
            pick = c;
        }
    }
    { KmHostLock lk_(km_host_mutex()); (void)hipEventDestroy(e0); }
    { KmHostLock lk_(km_host_mutex()); (void)hipEventDestroy(e1); }
    B->tuned = true;
    return set_tile(ctx, B, pick.first, pick.second);
}

void fill(km_solve_info* info, const KState& s) {
    info->iterations = s.iter;
    info->status = s.status;
    info->refinements = 0;
    info->fallback = 0;
    info->rel_resid = s.r0 > 0 ? sqrt(s.rTr / s.r0) : 0.0;
    info->rel_resid_true = -1.0;   // not recomputed in batch mode
    info->bnorm = s.r0 > 0 ? sqrt(s.r0) : 0.0;
}

int check_cases(km_ctx* ctx, int n, const km_case* cases) {
    KM_REQUIRE(ctx, n >= 1 && cases, "empty batch");
    for (int c = 0; c < n; ++c) {
        const km_case& K = cases[c];
        KM_REQUIRE(ctx, K.g.nx >= 4 && K.g.ny >= 3, "case %d: grid too small", c);
        KM_REQUIRE(ctx, K.g.dx > 0 && K.g.dy > 0 && K.g.dt > 0 && K.g.rho > 0, "case %d: invalid grid data", c);
        KM_REQUIRE(ctx, K.g.east == KM_EAST_WALL || K.g.east == KM_EAST_OUTFLOW, "case %d: unknown east kind", c);
        KM_REQUIRE(ctx, K.u && K.v && K.p && K.u0 && K.v0 && K.u_mid && K.v_mid && K.pcor && K.mdiv && K.coef_u &&
                            K.coef_v && K.coef_p && K.b_u && K.b_v && K.b_p && K.mask,
                    "case %d: NULL field", c);
    }
    return KM_OK;
}

}  // namespace

void km_batch_free(km_ctx* ctx) {
    if (!ctx || !ctx->batch) return;
    drop_stage_graphs(ctx->batch);
    KmHostLock lk(km_host_mutex());
    if (ctx->batch->arena) (void)hipFree(ctx->batch->arena);
    delete ctx->batch;
    ctx->batch = nullptr;
}

extern "C" {

int km_batch_step(km_ctx* ctx, int n, const km_case* cases, double eps_mom, double eps_p, int max_iter, int flags,
                  km_solve_info* info) {
    if (!ctx) return KM_EINVAL;
    if (int rc = check_cases(ctx, n, cases)) return rc;
    if (int rc = km_ensure_workspace(ctx, 0)) return rc;
    if (int rc = prepare(ctx, n, cases)) return rc;
    km_batch* B = ctx->batch;
    const bool use_graph = !(flags & KM_FLAG_NO_GRAPH);
    if (max_iter <= 0) max_iter = 1 << 30;
    if (int rc = run_stages(ctx, B, 0, use_graph)) return rc;
    if (!B->tuned && ctx->autotune)
        if (int rc = autotune(ctx, B, 0)) return rc;
    if (flags & KM_FLAG_SMOOTH) {
        // u and v case by case with the smoothed flexible solve, both systems
        // of a case in the same launches (km_fbicgstab2)
        for (int c = 0; c < n; ++c) {
            const km_case& K = B->cases[c];
            km_solve_info iu{}, iv{};
            const int nx = K.g.nx, ny = K.g.ny;
            const int rc = km_fbicgstab2(ctx, nx + 3, ny + 2, K.coef_u, K.b_u, K.u_mid, nx + 2, ny + 3, K.coef_v,
                                         K.b_v, K.v_mid, eps_mom, max_iter,
                                         KM_FLAG_SMOOTH | (flags & KM_FLAG_NO_GRAPH), &iu, &iv);
            if (rc < 0) return rc;
            if (info) {
                info[3 * c] = iu;
                info[3 * c + 1] = iv;
            }
        }
    } else {
        for (int f = 0; f < 2; ++f) {
            if (int rc = solve(ctx, B, f, eps_mom, max_iter, use_graph)) return rc;
            if (info)
                for (int c = 0; c < n; ++c) fill(&info[3 * c + f], B->h_states[c]);
        }
    }
    if (int rc = run_stages(ctx, B, 1, use_graph)) return rc;
    if (flags & KM_FLAG_COARSE) {
        // p' case by case with the two-level preconditioned flexible solve
        // (its coarse level is rebuilt per case shape); the assembled p'
        // systems are symmetric bit for bit (three-plane smoothing tiles)
        for (int c = 0; c < n; ++c) {
            const km_case& K = B->cases[c];
            km_solve_info si{};
            const int rc = km_fbicgstab(ctx, K.g.nx + 2, K.g.ny + 2, K.coef_p, K.b_p, K.pcor, eps_p, max_iter,
                                        KM_FLAG_COARSE | KM_FLAG_SYMMETRIC | (flags & KM_FLAG_NO_GRAPH), nullptr,
                                        &si);
            if (rc < 0) return rc;
            if (info) info[3 * c + 2] = si;
        }
    } else {
        if (int rc = solve(ctx, B, 2, eps_p, max_iter, use_graph)) return rc;
        if (info)
            for (int c = 0; c < n; ++c) fill(&info[3 * c + 2], B->h_states[c]);
    }
    if (int rc = run_stages(ctx, B, 2, use_graph)) return rc;
    return KM_OK;
}

int km_batch_time_forward(km_ctx* ctx, int n, const km_case* cases) {
    if (!ctx) return KM_EINVAL;
    if (int rc = check_cases(ctx, n, cases)) return rc;
    if (int rc = km_ensure_workspace(ctx, 0)) return rc;
    if (int rc = prepare(ctx, n, cases)) return rc;
    return stage(ctx, ctx->batch, BS_TIME, 0, ctx->stream);
}

int km_batch_get_tile(km_ctx* ctx, int* ti, int* tj) {
    if (!ctx || !ti || !tj) return KM_EINVAL;
    KM_REQUIRE(ctx, ctx->batch != nullptr, "no batch prepared");
    *ti = ctx->batch->TI;
    *tj = ctx->batch->TJ;
    return KM_OK;
}

}  // extern "C"
