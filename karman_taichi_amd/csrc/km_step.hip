// One SIMPLE outer iteration as one host call: SIMPLESolver.step
// (simple.py:270-281; the outlet fix between the momentum blend and the p'
// assembly as karman_taichi.py:645-646).  The stages are the same C-ABI
// calls a caller would make one by one; issuing them from here leaves no
// interpreter round trip between a solve's end and the next stage's launch,
// and the next stage's kernels reach the stream while the GPU still runs
// the previous ones.
#include "km_common.h"

namespace {

// the two momentum systems: both with the smoothed flexible solve in shared
// launches (km_fbicgstab2), or km_bicgstab each (simple.py:246-248)
int momentum(km_ctx* ctx, const km_case* c, double eps, int flags, km_solve_info* iu, km_solve_info* iv) {
    const int nx = c->g.nx, ny = c->g.ny;
    if (flags & (KM_FLAG_SMOOTH | KM_FLAG_COARSE))
        return km_fbicgstab2(ctx, nx + 3, ny + 2, c->coef_u, c->b_u, c->u_mid, nx + 2, ny + 3, c->coef_v, c->b_v,
                             c->v_mid, eps, 0, flags, iu, iv);
    if (int rc = km_bicgstab(ctx, nx + 3, ny + 2, c->coef_u, c->b_u, c->u_mid, eps, 0, flags, iu)) return rc;
    return km_bicgstab(ctx, nx + 2, ny + 3, c->coef_v, c->b_v, c->v_mid, eps, 0, flags, iv);
}

// solve_momentum_eqn (simple.py:241-250), the blend, the outlet fix and the
// p' assembly (simple.py:252-253): the step up to the p' solve
int before_pcor(km_ctx* ctx, const km_case* c, double eps_mom, int mom_flags, km_solve_info* info) {
    KM_REQUIRE(ctx, !(mom_flags & KM_FLAG_COARSE), "km_simple_step: the momentum systems take no coarse level");
    KM_REQUIRE(ctx, !(mom_flags & KM_FLAG_X0), "km_simple_step: the momentum solves start from zero (simple.py:246)");
    const km_grid* g = &c->g;
    if (int rc = km_assemble_u(ctx, g, c->u, c->v, c->p, c->u0, c->mask, c->coef_u, c->b_u)) return rc;
    if (int rc = km_assemble_v(ctx, g, c->u, c->v, c->p, c->v0, c->mask, c->coef_v, c->b_v)) return rc;
    if (int rc = km_apply_ghosts(ctx, g, c->u, c->v)) return rc;
    if (int rc = momentum(ctx, c, eps_mom, mom_flags, &info[0], &info[1])) return rc;
    if (int rc = km_blend(ctx, g, c->u, c->v, c->u_mid, c->v_mid)) return rc;
    if (g->east == KM_EAST_OUTFLOW)
        if (int rc = km_outflow_fix(ctx, g, c->u)) return rc;
    return km_assemble_p(ctx, g, c->u, c->v, c->coef_u, c->coef_v, c->mask, c->coef_p, c->b_p, c->mdiv);
}

// the p' solve's flags: the assembled p' system is symmetric bit for bit
// (aW(i, j) and aE(i - 1, j) are the same expression; a Dirichlet cell zeroes
// both sides), which the two-level solve's three-plane kernels use
int pcor_flags(int p_flags) { return p_flags | ((p_flags & KM_FLAG_COARSE) ? KM_FLAG_SYMMETRIC : 0); }

int solve_pcor(km_ctx* ctx, const km_case* c, double eps_p, int p_flags, km_solve_info* info) {
    const int nx = c->g.nx, ny = c->g.ny;
    return p_flags & (KM_FLAG_COARSE | KM_FLAG_SMOOTH)
               ? km_fbicgstab(ctx, nx + 2, ny + 2, c->coef_p, c->b_p, c->pcor, eps_p, 0, pcor_flags(p_flags), nullptr,
                              info)
               : km_bicgstab(ctx, nx + 2, ny + 2, c->coef_p, c->b_p, c->pcor, eps_p, 0, p_flags, info);
}

// correct_pressure + correct_velocity (simple.py:218-230), then mdiv
int after_pcor(km_ctx* ctx, const km_case* c) {
    const km_grid* g = &c->g;
    if (int rc = km_correct(ctx, g, c->p, c->u, c->v, c->pcor, c->coef_u, c->coef_v, c->mask, 3)) return rc;
    return km_mdiv(ctx, g, c->u, c->v, c->mdiv, nullptr);
}

}  // namespace

extern "C" {

int km_simple_step(km_ctx* ctx, const km_case* c, double eps_mom, double eps_p, int mom_flags, int p_flags,
                   km_solve_info* info) {
    if (!ctx) return KM_EINVAL;
    KM_REQUIRE(ctx, c && info, "km_simple_step: NULL pointer");
    if (int rc = before_pcor(ctx, c, eps_mom, mom_flags, info)) return rc;
    if (int rc = solve_pcor(ctx, c, eps_p, p_flags, &info[2])) return rc;
    return after_pcor(ctx, c);
}

int km_simple_step_n(km_ctx* ctx, int n, const km_case* cases, double eps_mom, double eps_p, const int* mom_flags,
                     const int* p_flags, km_solve_info* info) {
    if (!ctx) return KM_EINVAL;
    KM_REQUIRE(ctx, n >= 0 && n <= KM_MAX_GROUP_SOLVES, "km_simple_step_n: %d cases (at most %d)", n,
               KM_MAX_GROUP_SOLVES);
    if (n == 0) return KM_OK;
    KM_REQUIRE(ctx, cases && mom_flags && p_flags && info, "km_simple_step_n: NULL pointer");
    for (int k = 0; k < n; ++k)
        if (int rc = before_pcor(ctx, &cases[k], eps_mom, mom_flags[k], &info[3 * k])) return rc;
    // the p' solves: cases of the same flags through km_fbicgstab_n (the
    // two-level ones in shared launches), the others one by one
    std::vector<char> done(n, 0);
    std::vector<int> NX, NY;
    std::vector<const double*> coef, b;
    std::vector<double*> x;
    std::vector<km_solve_info> pinfo;
    std::vector<int> which;
    for (int a = 0; a < n; ++a) {
        if (done[a]) continue;
        if (!(p_flags[a] & KM_FLAG_COARSE)) {
            done[a] = 1;
            if (int rc = solve_pcor(ctx, &cases[a], eps_p, p_flags[a], &info[3 * a + 2])) return rc;
            continue;
        }
        NX.clear(); NY.clear(); coef.clear(); b.clear(); x.clear(); which.clear();
        for (int k = a; k < n; ++k)
            if (!done[k] && p_flags[k] == p_flags[a]) {
                done[k] = 1;
                const km_case* c = &cases[k];
                NX.push_back(c->g.nx + 2);
                NY.push_back(c->g.ny + 2);
                coef.push_back(c->coef_p);
                b.push_back(c->b_p);
                x.push_back(c->pcor);
                which.push_back(k);
            }
        pinfo.assign(which.size(), km_solve_info{});
        if (int rc = km_fbicgstab_n(ctx, (int)which.size(), NX.data(), NY.data(), coef.data(), b.data(), x.data(),
                                    eps_p, 0, pcor_flags(p_flags[a]), pinfo.data()))
            return rc;
        for (size_t j = 0; j < which.size(); ++j) info[3 * which[j] + 2] = pinfo[j];
    }
    for (int k = 0; k < n; ++k)
        if (int rc = after_pcor(ctx, &cases[k])) return rc;
    return KM_OK;
}

}  // extern "C"
