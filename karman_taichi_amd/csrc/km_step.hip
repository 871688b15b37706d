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

}  // namespace

extern "C" {

int km_simple_step(km_ctx* ctx, const km_case* c, double eps_mom, double eps_p, int mom_flags, int p_flags,
                   km_solve_info* info) {
    if (!ctx) return KM_EINVAL;
    KM_REQUIRE(ctx, c && info, "km_simple_step: NULL pointer");
    KM_REQUIRE(ctx, !(mom_flags & KM_FLAG_COARSE), "km_simple_step: the momentum systems take no coarse level");
    KM_REQUIRE(ctx, !(mom_flags & KM_FLAG_X0), "km_simple_step: the momentum solves start from zero (simple.py:246)");
    const km_grid* g = &c->g;
    const int nx = g->nx, ny = g->ny;
    // solve_momentum_eqn (simple.py:241-250)
    if (int rc = km_assemble_u(ctx, g, c->u, c->v, c->p, c->u0, c->mask, c->coef_u, c->b_u)) return rc;
    if (int rc = km_assemble_v(ctx, g, c->u, c->v, c->p, c->v0, c->mask, c->coef_v, c->b_v)) return rc;
    if (int rc = km_apply_ghosts(ctx, g, c->u, c->v)) return rc;
    if (int rc = momentum(ctx, c, eps_mom, mom_flags, &info[0], &info[1])) return rc;
    if (int rc = km_blend(ctx, g, c->u, c->v, c->u_mid, c->v_mid)) return rc;
    if (g->east == KM_EAST_OUTFLOW)
        if (int rc = km_outflow_fix(ctx, g, c->u)) return rc;
    // solve_pcorrection_eqn (simple.py:252-255)
    if (int rc = km_assemble_p(ctx, g, c->u, c->v, c->coef_u, c->coef_v, c->mask, c->coef_p, c->b_p, c->mdiv))
        return rc;
    // the assembled p' system is symmetric bit for bit (aW(i, j) and
    // aE(i - 1, j) are the same expression; a Dirichlet cell zeroes both sides)
    const int rc = p_flags & (KM_FLAG_COARSE | KM_FLAG_SMOOTH)
                       ? km_fbicgstab(ctx, nx + 2, ny + 2, c->coef_p, c->b_p, c->pcor, eps_p, 0,
                                      p_flags | ((p_flags & KM_FLAG_COARSE) ? KM_FLAG_SYMMETRIC : 0), nullptr, &info[2])
                       : km_bicgstab(ctx, nx + 2, ny + 2, c->coef_p, c->b_p, c->pcor, eps_p, 0, p_flags, &info[2]);
    if (rc) return rc;
    // correct_pressure + correct_velocity (simple.py:218-230), then mdiv
    if (int rc2 = km_correct(ctx, g, c->p, c->u, c->v, c->pcor, c->coef_u, c->coef_v, c->mask, 3)) return rc2;
    return km_mdiv(ctx, g, c->u, c->v, c->mdiv, nullptr);
}

}  // extern "C"
