/*
 * karman.h -- C-ABI of libkarman_hip.so, the MI355X (gfx950) SIMPLE hot path.
 *
 * Plain C: pointers and sizes only, no torch types.  Every entry point
 * replaces one reference interface (cited as file:line of the reference
 * repository houkensjtu/karman_taichi).
 *
 * Layout contract
 *   Fields are FP64 device arrays in the reference's padded shapes, i-major
 *   with j contiguous (what ti.field.to_numpy() returns):
 *     u, u0, u_mid : (nx+3) x (ny+2)      simple.py:27,30,33
 *     v, v0, v_mid : (nx+2) x (ny+3)      simple.py:28,31,34
 *     p, pcor, mdiv, mask : (nx+2) x (ny+2)   simple.py:36-40
 *   A coefficient array is SoA: 5 planes of the field's shape, plane k is
 *   0 aP, 1 aW, 2 aE, 3 aN, 4 aS (simple.py:42-47, comments :81-92).
 *   The obstacle mask is uint8 (1 = solid cell), karman_taichi.py:49.
 *
 * Conventions
 *   Return value: KM_OK (0) or a solver status (> 0) or an error (< 0).
 *   A context is bound to one device and one HIP stream; calls enqueue work
 *   on that stream and return without synchronising, except those that
 *   return host values (km_dot, km_mdiv_stats, the solvers' km_solve_info).
 *   The caller owns every field; the context owns its solver workspace,
 *   captured graphs and tile choices.  A context is not thread-safe.
 */
#ifndef KARMAN_H
#define KARMAN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status and error codes */
#define KM_OK 0
#define KM_BREAKDOWN 1   /* rho == 0 or <r~,v> == 0 or omega == 0 (bicgsolver.py:96-98, karman_taichi.py:451-454) */
#define KM_MAXITER 2     /* iteration cap reached (cgsolver.py:14 caps at NX*NY) */
#define KM_NONFINITE 3   /* NaN/Inf residual */
#define KM_STALL 4       /* preconditioned solve stalled (the plain fallback then ran) */
#define KM_EINVAL (-1)
#define KM_EHIP (-2)
#define KM_ENOMEM (-3)

/* vector/coefficient precision of a Krylov solve */
#define KM_F64 0
#define KM_F32 1

/* km_bicgstab / km_cg flags */
#define KM_FLAG_X0 1          /* use x as the initial guess (default: x0 = 0, bicgsolver.py:21-35) */
#define KM_FLAG_MIXED 2       /* FP32 matvec/axpy inner solves, FP64 scalars and FP64 true-residual refinement */
#define KM_FLAG_NO_GRAPH 4    /* launch kernels directly instead of replaying captured graphs */
#define KM_FLAG_COARSE 8      /* km_fbicgstab: two-level preconditioner (aggregation coarse correction first) */
#define KM_FLAG_SMOOTH 16     /* km_fbicgstab: red-black Gauss-Seidel sweeps as the preconditioner (km_set_momentum_sweeps), no coarse level */
#define KM_FLAG_GCR 64        /* km_fbicgstab: flexible GCR (restarted every 16 directions) as the outer method instead
                               * of flexible BiCGSTAB: ||r|| never increases whatever the preconditioner does, so a
                               * non-linear network preconditioner cannot break its recurrences (no counterpart in the
                               * reference) */
#define KM_FLAG_JACOBI 128    /* km_fbicgstab with KM_FLAG_GCR and KM_FLAG_COARSE: without a network, one Jacobi step
                               * after the coarse correction instead of the red-black SOR sweeps; with a network, one
                               * direction z1 + net(r1) followed by a Jacobi step instead of the two directions
                               * M(r) (the two-level preconditioner: SOR sweeps from z1) and z1 + net(r1) */
#define KM_FLAG_MG 256       /* km_fbicgstab: the preconditioner is one multigrid V-cycle (km_mg_vcycle) in place of
                               * Jacobi; FP64 vectors, no coarse level, smoothing or network.  A comparator for the
                               * two-level and learned preconditioners in the spirit of benchmark/mgpcg.py:91-135
                               * (the reference's V-cycle is for a constant-coefficient Poisson operator) */
#define KM_FLAG_SYMMETRIC 512 /* km_fbicgstab with KM_FLAG_COARSE: the caller asserts that the system is symmetric bit
                               * for bit, aW(i, j) = aE(i - 1, j) and aS(i, j) = aN(i, j - 1), as the assembled p'
                               * systems are (compute_coef_p, simple.py:143-170); the two-level smoothing kernels then
                               * read the aP, aE and aN planes only, with results identical to the five-plane kernels.
                               * Results are undefined for a system that is not (no counterpart in the reference) */
#define KM_FLAG_TRUE 32       /* km_bicgstab: after the reference's stop on the recursive residual, check the true
                               * residual ||b - A x|| / ||b|| and restart from x (r = b - A x) while it is above eps
                               * and each restart at least halves it (at most 8 restarts; no counterpart in the
                               * reference, which stops on the recursive residual, bicgsolver.py:128-131) */

/* east boundary kinds */
#define KM_EAST_WALL 0        /* simple.py:184-186, :214-216 (with the bicgsimp.py:196 fix) */
#define KM_EAST_OUTFLOW 1     /* zero gradient + outlet flux fix, karman_taichi.py:180-183, :586-614 */

typedef struct km_ctx km_ctx;

/* Grid, physics and boundary data of a SIMPLESolver (simple.py:11-39). */
typedef struct km_grid {
    int32_t nx, ny;
    int32_t east;        /* KM_EAST_WALL or KM_EAST_OUTFLOW */
    int32_t pin_pcor;    /* 1: p' = 0 at cell (1,1), bicgsimp.py:146-151 */
    double dx, dy, rho, mu, dt;
    double alpha_u, alpha_p, alpha_m;
    double bc[4][2];     /* sides w, e, n, s; [normal, tangential] as simple.py:39 */
} km_grid;

typedef struct km_solve_info {
    int32_t iterations;     /* Krylov iterations (inner iterations summed in mixed mode) */
    int32_t status;         /* KM_OK, KM_BREAKDOWN, KM_MAXITER, KM_NONFINITE, KM_STALL */
    int32_t refinements;    /* mixed mode: FP64 refinement sweeps */
    int32_t fallback;       /* learned mode: 1 if the plain solver finished the solve */
    double rel_resid;       /* recursive sqrt(r.r / r0.r0) (bicgsolver.py:130) */
    double rel_resid_true;  /* ||b - A x|| / ||b|| recomputed in FP64 */
    double bnorm;           /* ||b|| */
} km_solve_info;

/* ---- context (replaces ti.init, simple.py:7) ---- */
int km_create(km_ctx** out, int device, void* hip_stream);
int km_destroy(km_ctx* ctx);
int km_set_stream(km_ctx* ctx, void* hip_stream);
const char* km_last_error(const km_ctx* ctx);
int km_synchronize(km_ctx* ctx);
/* Tile shape (rows x columns of one workgroup) of the stencil kernels for an
 * NX x NY system; ti = tj = 0 clears the override and re-enables autotuning. */
int km_set_tile(km_ctx* ctx, int NX, int NY, int prec, int ti, int tj);
int km_get_tile(km_ctx* ctx, int NX, int NY, int prec, int* ti, int* tj);
/* Tile autotuning (on by default): the first solve of a shape times each valid
 * tile candidate on that system and keeps the fastest (no reference
 * counterpart; the reference's only tuning knob is ti.block_dim,
 * benchmark/para_vs_seri.py:27). */
int km_set_autotune(km_ctx* ctx, int on);

/* ---- discretisation (SIMPLESolver kernels) ---- */
/* compute_coef_u + the u half of set_bc (simple.py:77-92, :177-207) + obstacle rows */
int km_assemble_u(km_ctx* ctx, const km_grid* g, const double* u, const double* v, const double* p,
                  const double* u0, const uint8_t* mask, double* coef_u, double* b_u);
/* compute_coef_v + the v half of set_bc (simple.py:95-110, :188-216) + obstacle rows */
int km_assemble_v(km_ctx* ctx, const km_grid* g, const double* u, const double* v, const double* p,
                  const double* v0, const uint8_t* mask, double* coef_v, double* b_v);
/* ghost/boundary velocities written by set_bc (simple.py:181, :186, :192, :196) */
int km_apply_ghosts(km_ctx* ctx, const km_grid* g, double* u, double* v);
/* update_velocity: u = a_m u* + (1 - a_m) u (simple.py:232-238) */
int km_blend(km_ctx* ctx, const km_grid* g, double* u, double* v, const double* u_mid, const double* v_mid);
/* zero-gradient outlet + flux fix (karman_taichi.py:180-183, correct_conserv :586-614) */
int km_outflow_fix(km_ctx* ctx, const km_grid* g, double* u);
/* compute_coef_p (simple.py:143-170): writes coef_p, b_p and mdiv */
int km_assemble_p(km_ctx* ctx, const km_grid* g, const double* u, const double* v, const double* coef_u,
                  const double* coef_v, const uint8_t* mask, double* coef_p, double* b_p, double* mdiv);
/* correct_pressure (parts bit 1) and/or correct_velocity (parts bit 2) in one
 * launch (simple.py:218-230, karman_taichi.py:507-542) */
int km_correct(km_ctx* ctx, const km_grid* g, double* p, double* u, double* v, const double* pcor,
               const double* coef_u, const double* coef_v, const uint8_t* mask, int parts);
/* compute_mdiv (simple.py:137-141); stats (may be NULL) receives {max|mdiv|, sum mdiv^2}
 * over interior cells (bicgsimp.py:114-122 returns the max) and makes the call synchronous. */
int km_mdiv(km_ctx* ctx, const km_grid* g, const double* u, const double* v, double* mdiv, double* stats);
/* time_forward: u0 <- u, v0 <- v (karman_taichi.py:617-622) */
int km_time_forward(km_ctx* ctx, const km_grid* g, const double* u, const double* v, double* u0, double* v0);

/* ---- Krylov (CGSolver / BICGSolver) ---- */
/* y = A x over the whole padded NX x NY field (bicgsolver.py:65-76, cgsolver.py:56-66) */
int km_stencil_apply(km_ctx* ctx, int NX, int NY, int prec, const void* coef, const void* x, void* y);
/* <a, b> in FP64 over n elements (CGSolver.reduce, cgsolver.py:49-54); synchronous */
int km_dot(km_ctx* ctx, int64_t n, int prec, const void* a, const void* b, double* out);
/* r = b - A x and ||r||^2 (cgsolver.py:121-131 intent); rnorm2 may be NULL (then asynchronous) */
int km_residual(km_ctx* ctx, int NX, int NY, const double* coef, const double* b, const double* x,
                double* r, double* rnorm2);
/* BICGSolver.solve (bicgsolver.py:88-136): FP64 coefficients, right-hand side and solution. */
int km_bicgstab(km_ctx* ctx, int NX, int NY, const double* coef, const double* b, double* x,
                double eps, int max_iter, int flags, km_solve_info* info);
/* km_bicgstab without any allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under stream capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The same plain BiCGSTAB
 * kernels as km_bicgstab, launched directly (no graphs), with the solver's
 * state, partial sums and eight Krylov vectors in a device array the library
 * reserves once per GPU (KM_NOALLOC_SCRATCH_BYTES = 32 MiB; at most
 * KM_NOALLOC_MAX_CELLS = NX*NY cells, e.g. the 1282 x 258 p' system of a
 * 1280 x 256 grid).  Larger systems and KM_FLAG_MIXED return KM_EINVAL.  The
 * tile is the context's cached or override choice, else the default split
 * (no timing).  Every call on one GPU shares that array: calls that overlap
 * on one GPU (several contexts or streams at once) race. */
#define KM_NOALLOC_SCRATCH_BYTES (32u << 20)
#define KM_NOALLOC_MAX_CELLS 521144
int km_bicgstab_noalloc(km_ctx* ctx, int NX, int NY, const double* coef, const double* b, double* x,
                        double eps, int max_iter, int flags, km_solve_info* info);
/* z = M(r): one multigrid V-cycle from z = 0 on the hierarchy of the NX x NY
 * system coef (built on the call; levels, if not null, receives its depth).
 * Levels: 2 x 2 aggregates of interior cells with the Galerkin operator
 * P^T A P, down to at most 400 interior cells; per level one red-black SOR
 * sweep (red first), the residual summed into the aggregates, the next level,
 * the correction prolonged by injection with weight 1.5, one sweep black
 * first; 50 sweeps on the coarsest level.  The V-cycle of
 * benchmark/mgpcg.py:112-135 (red-black smoothing :118-123, summed
 * restriction :103-110, injection :112-115, 50 bottom sweeps :127-128),
 * restated for the variable-coefficient p' operator; the preconditioner
 * KM_FLAG_MG applies.  Synchronises the context's stream. */
int km_mg_vcycle(km_ctx* ctx, int NX, int NY, const double* coef, const double* r, double* z, int* levels);
/* CGSolver.solve (cgsolver.py:83-114). */
int km_cg(km_ctx* ctx, int NX, int NY, const double* coef, const double* b, double* x,
          double eps, int max_iter, int flags, km_solve_info* info);

/* ---- learned p' preconditioner ----
 * A 4-level U-Net of 3x3 convolutions (16/32/64/128 channels, BF16 weights and
 * activations on the matrix cores, FP32 accumulation) maps (p' residual, aP,
 * identity-row mask) to a p' correction.  It fills the right-preconditioner
 * hooks update_phat / update_shat of BICGSolver (bicgsolver.py:48-58, identity
 * there; karman_taichi.py:399,419 plugs Jacobi into the same hooks).  The
 * network itself has no counterpart in the reference.
 *
 * Parameters are one caller-owned FP32 device vector of km_unet_param_count
 * values: for each layer e0 e1 e2 e3 d2 d1 d0, W[cout][3][3][cin] then
 * b[cout]; then the 1x1 output layer w[16], b[1]. */
int km_unet_param_count(int64_t* n);
/* z = M(r) for one system (FP64 coefficients and vectors; the network runs in
 * BF16/FP32).  Asynchronous. */
int km_unet_apply(km_ctx* ctx, const float* params, int NX, int NY, const double* coef, const double* r, double* z);
/* Training step without the update: the batch mean of a relative loss of
 * z = M(r) over B systems and its gradient.  corr == NULL: the residual loss
 * ||r - A z||^2 / ||r||^2; corr = the converged corrections c (A c = r): the
 * A-norm error loss (c - z)^T A (c - z) / c^T r.  coef: B x 5 planes of
 * NX x NY FP32; r, corr: B x NX x NY FP32; grad: param_count FP32.  loss
 * (may be NULL) receives the batch mean and makes the call synchronous.
 * B <= 64. */
int km_unet_loss_grad(km_ctx* ctx, const float* params, int B, int NX, int NY, const float* coef, const float* r,
                      const float* corr, float* grad, double* loss);
/* Adam (Kingma & Ba 2015) on FP32 vectors of n values; step counts from 1. */
int km_adam_step(km_ctx* ctx, int64_t n, float* params, const float* grad, float* m, float* v, int step, double lr,
                 double b1, double b2, double eps);
/* BICGSolver.solve (bicgsolver.py:88-136) with M = the network in the
 * p^ / s^ hooks: flexible BiCGSTAB with FP64 scalars inside FP64 iterative
 * refinement, judged only by the true residual ||b - A x|| / ||b|| < eps.
 * Krylov vectors are FP64, or FP32 with KM_FLAG_MIXED (then each inner sweep
 * stops at 1e-5 at most).  Without KM_FLAG_X0 the first iterate is the
 * network's warm start x0 = M(b) (kept if it lowers the residual).  A sweep
 * that does not halve the true residual is taken back and plain FP64
 * BiCGSTAB finishes (info->fallback = 1).  params == NULL selects the Jacobi
 * preconditioner D^-1 instead (karman_taichi.py:399,419), with no warm start.
 * KM_FLAG_COARSE makes the preconditioner two-level: z1 = P A_c^-1 P^T r on
 * B x B aggregates of free cells (Galerkin A_c, its inverse held across
 * solves, km_set_coarse_update), then the network on r - A z1, added to z1 --
 * or, without a network, two red-black SOR sweeps (omega 1.3) on A z = r from
 * z1, run as a fused four-launch iteration.  KM_FLAG_SMOOTH: one or two
 * red-black Gauss-Seidel sweeps (km_set_momentum_sweeps) on A z = r from
 * zero as the preconditioner (no network,
 * no coarse level; the momentum systems), a fused three-launch iteration.
 * FP64 solves from x = 0 with either fused iteration solve straight into x. */
int km_fbicgstab(km_ctx* ctx, int NX, int NY, const double* coef, const double* b, double* x, double eps,
                 int max_iter, int flags, const float* params, km_solve_info* info);
/* The u and v momentum solves of one SIMPLE step (simple.py:247-248, which
 * calls BICGSolver.solve on each in turn): two independent km_fbicgstab
 * solves with the same eps, max_iter and flags, results and info as two
 * separate calls.  With KM_FLAG_SMOOTH and none of KM_FLAG_MIXED, KM_FLAG_X0,
 * KM_FLAG_COARSE the two systems' first sweeps iterate in the same launches
 * (each workgroup serves one system; each system stops on its own test);
 * otherwise the two solves run one after the other. */
int km_fbicgstab2(km_ctx* ctx, int NX0, int NY0, const double* coef0, const double* b0, double* x0, int NX1, int NY1,
                  const double* coef1, const double* b1, double* x1, double eps, int max_iter, int flags,
                  km_solve_info* info0, km_solve_info* info1);
/* n independent km_fbicgstab solves (no network) with the same eps, max_iter
 * and flags: results and info[k] as n separate calls in order.  With
 * KM_FLAG_COARSE | KM_FLAG_SYMMETRIC and none of KM_FLAG_X0, KM_FLAG_MIXED,
 * KM_FLAG_SMOOTH, KM_FLAG_GCR, KM_FLAG_JACOBI, KM_FLAG_MG (the p' systems of
 * several cases) every system keeps its own coarse level and smoothing tile,
 * and the first sweeps of systems of the same tile shape, up to four at a
 * time, iterate in the same launches, each workgroup serving one system and
 * each system stopping on its own test; otherwise the solves run one after
 * the other.  n <= KM_MAX_GROUP_SOLVES; the systems must not share x or
 * coefficients.  No counterpart in the reference (one case per process). */
#define KM_MAX_GROUP_SOLVES 256
int km_fbicgstab_n(km_ctx* ctx, int n, const int* NX, const int* NY, const double* const* coef,
                   const double* const* b, double* const* x, double eps, int max_iter, int flags,
                   km_solve_info* info);

/* ---- batched cases (no counterpart in the reference, which runs one case) ----
 * Many independent channel cases of uneven size and Reynolds number advance
 * together: each stage of the SIMPLE outer iteration (simple.py:270-281) is
 * one launch over every case, the u, v and p' BiCGSTAB solves run all cases'
 * systems in the same fused iteration kernels (per-case scalars and stopping
 * tests), the stage sequences between the solves are captured as hipGraphs,
 * and one tile shape, timed on the batch's own systems, serves all cases.
 * Each case holds the fields of one SIMPLESolver (the layout contract above). */
typedef struct km_case {
    km_grid g;
    double *u, *v, *p, *u0, *v0, *u_mid, *v_mid, *pcor, *mdiv;
    double *coef_u, *coef_v, *coef_p, *b_u, *b_v, *b_p;
    const uint8_t* mask;
} km_case;
/* One outer iteration of every case (plain FP64 solves to eps_mom / eps_p,
 * x0 = 0 as in SIMPLESolver.step).  info (may be NULL) receives 3 entries per
 * case (u, v, p'); rel_resid_true is not recomputed (-1) for the batched
 * solves.  flags KM_FLAG_COARSE: each case's p' system is solved by the
 * two-level flexible BiCGSTAB of km_fbicgstab instead (case by case; its
 * info carries the true residual); KM_FLAG_SMOOTH: each case's u and v
 * systems by km_fbicgstab2 with KM_FLAG_SMOOTH (case by case). */
int km_batch_step(km_ctx* ctx, int n, const km_case* cases, double eps_mom, double eps_p, int max_iter, int flags,
                  km_solve_info* info);
/* time_forward of every case (karman_taichi.py:617-622). */
int km_batch_time_forward(km_ctx* ctx, int n, const km_case* cases);
/* SIMPLESolver.step of one case (simple.py:270-281, the outlet fix of
 * karman_taichi.py:645-646 after the momentum blend) as one call: assembly,
 * the u and v solves (mom_flags with KM_FLAG_SMOOTH: km_fbicgstab2, else
 * km_bicgstab on each), blend, outlet fix, p' assembly, the p' solve
 * (p_flags with KM_FLAG_COARSE or KM_FLAG_SMOOTH: km_fbicgstab without a
 * network, else km_bicgstab), both corrections and mdiv.  The same kernels
 * in the same order as those calls made one by one; info receives u, v, p'.
 * eps and flags as the single solves (max_iter = NX * NY). */
int km_simple_step(km_ctx* ctx, const km_case* c, double eps_mom, double eps_p, int mom_flags, int p_flags,
                   km_solve_info* info);
/* km_simple_step of n cases (each case's own flags, info 3 entries per case):
 * every case's stages up to its p' system, then the p' solves -- the cases
 * with the same p_flags including KM_FLAG_COARSE together through
 * km_fbicgstab_n (their first sweeps in shared launches), the others one by
 * one -- then every case's corrections.  Each case's results are those of its
 * own km_simple_step.  n <= KM_MAX_GROUP_SOLVES.  No counterpart in the
 * reference (one case per process). */
int km_simple_step_n(km_ctx* ctx, int n, const km_case* cases, double eps_mom, double eps_p, const int* mom_flags,
                     const int* p_flags, km_solve_info* info);
/* The batch's tile shape (chosen on its first step). */
int km_batch_get_tile(km_ctx* ctx, int* ti, int* tj);

/* Aggregate size of the coarse level (0: automatic, the smallest of 8, 12, 16, 24,
 * 32, 48, ..., 256 giving at most 2048 coarse unknowns).  No counterpart in the
 * reference. */
int km_set_coarse(km_ctx* ctx, int block);
/* z1 = P A_c^-1 P^T r and r1 = r - A z1 for one FP64 system (sets the coarse
 * level up first); the training data of the two-level network and tests use it. */
int km_coarse_correct(km_ctx* ctx, int NX, int NY, const double* coef, const double* r, double* z1, double* r1);
/* How the coarse level follows a changing system.  every = 1: a coarse level
 * of the same shape as the last one updates the held A_c^-1 by one
 * Newton-Schulz step X (2 I - A_c X) when ||I - A_c X||_inf < 0.5, instead of
 * factoring A_c again; 0: always factor; k > 1 (default 32): do that on every
 * k-th set-up and, on the set-ups between, reuse the held level (free-cell
 * mask, A_c^-1) unchanged when the coefficient array is the same one as last
 * time (a SIMPLE solver's p' system across outer iterations).  The flexible
 * solve stays exact with a stale coarse inverse; on the 1280 x 256 channel
 * its iteration count does not change.  A caller that rewrites obstacle cells
 * in place should use every <= 1.  No counterpart in the reference. */
int km_set_coarse_update(km_ctx* ctx, int every);
/* Aggregate size and coarse unknowns of the last coarse level. */
int km_coarse_info(km_ctx* ctx, int* block, int* unknowns);
/* z = M(r) for one FP64 system: the two-level preconditioner of the flexible
 * solve with KM_FLAG_COARSE and no network, z1 = P A_c^-1 P^T r then two
 * red-black SOR sweeps (red cells (i + j even) first, omega 1.3) on A z = r.
 * It fills the p^ / s^ hooks of the reference's BiCGSTAB (bicgsolver.py:48-58,
 * karman_taichi.py:399,419 pass a Jacobi scaling there); sets the coarse level
 * up first (always current: no lagged level here). */
int km_twolevel_precond(km_ctx* ctx, int NX, int NY, const double* coef, const double* r, double* z);
/* Tile shape (rows, columns) of the fused two-level smoothing kernels, chosen
 * by timing on the first KM_FLAG_COARSE solve of the coarse level's system.
 * No counterpart in the reference. */
int km_twolevel_tile(km_ctx* ctx, int* ti, int* tj);
/* Threads per workgroup of that tile shape (256 or 512). */
int km_twolevel_tile_threads(km_ctx* ctx, int* threads);
/* Pin the smoothing tile of the coarse level's system to candidate `which`
 * (0 .. 8; KM_EINVAL past the last), or clear the choice (which < 0) so the
 * next solve times the candidates again.  Needs a coarse level.  For tests
 * and tuning studies; no counterpart in the reference. */
int km_twolevel_set_tile(km_ctx* ctx, int which);
/* How the fused two-level kernels combine the previous kernel's per-workgroup
 * partial sums: every consumer workgroup sums them itself, or one workgroup
 * sums them first (an extra small launch) when partials x consumer workgroups
 * is large (C5-sized grids).  mode -1 automatic (default), 0 never, 1 always;
 * the totals are the same bits either way.  Needs a coarse level (one
 * KM_FLAG_COARSE solve first).  No counterpart in the reference. */
int km_set_coarse_totals(km_ctx* ctx, int mode);
/* The same choice for the plain BiCGSTAB iteration (km_bicgstab and the
 * mixed path's FP32 inner solves): P, S and R read the previous kernel's
 * totals from a one-workgroup launch instead of each summing its partials.
 * mode -1 automatic (default: R's totals for P when P's workgroups x R's
 * partials exceed 4e6, e.g. at 4096 x 820), 0 never, 1 all three; the same
 * bits either way.  No counterpart in the reference. */
int km_set_plain_totals(km_ctx* ctx, int mode);
/* Red-black Gauss-Seidel sweeps of KM_FLAG_SMOOTH's preconditioner: 0
 * automatic (two for FP64 systems of at least 200,000 cells, else one;
 * default), 1 or 2.  Two sweeps save an iteration on the 1280 x 256
 * channel's momentum systems; where several contexts' streams share the GPU
 * one sweep is faster.  No counterpart in the reference. */
int km_set_momentum_sweeps(km_ctx* ctx, int sweeps);
/* Launch schedule of the fused two-level iteration: captured graphs of `big`
 * iterations (a multiple of 4; 0: 8) until `lead` iterations before the
 * shortest of the last eight solves of this system shape, then graphs of
 * `small` (1, 2 or 4; 0: 2), so that few launches are left in flight when
 * the stopping test ends the solve.  Defaults 8, 2, 2.  Needs a coarse level.
 * No counterpart in the reference. */
int km_set_coarse_chunks(km_ctx* ctx, int big, int small, int lead);

/* ---- measurement hook (no counterpart in the reference) ----
 * Average device time in microseconds of one launch of a solver kernel on an
 * NX x NY system, from HIP events around `reps` launches on the context's
 * stream.  Used by the benchmark to report achieved bandwidth of the kernel
 * that dominates the solve; rocprofv3 --kernel-trace reports the same kernels. */
#define KM_KERNEL_BICG_A 0   /* P: stopping test, beta; p update + stencil (v = A p) + <r~, v> */
#define KM_KERNEL_BICG_B 1   /* S: alpha; s update + stencil (t = A s) + <t, s>, <t, t> */
#define KM_KERNEL_BICG_R 2   /* R: omega; x, r updates (flat) + <r, r>, <r, r~> */
#define KM_KERNEL_STENCIL 3  /* plain y = A x */
/* the fused two-level flexible iteration (KM_FLAG_COARSE, no network) */
#define KM_KERNEL_TL_PV 4    /* p^ = M(p) (prolongation + 2 red-black SOR sweeps); v = A p^; <r~, v>; P^T v */
#define KM_KERNEL_TL_CV 5    /* stopping test, alpha; coarse Y(v) = A_c^-1 P^T v, Y(s) */
#define KM_KERNEL_TL_ST 6    /* s = r - alpha v; s^ = M(s); t = A s^; <t, s>, <t, t>, <s, r~>, <t, r~>; P^T t */
#define KM_KERNEL_TL_CT 7    /* omega, beta; coarse Y(t), Y(r), Y(p); flat x, r, p updates, <r, r> */
#define KM_KERNEL_BACK_TO_BACK 0x100   /* or-ed in: time `reps` identical back-to-back launches between two events */
#define KM_KERNEL_SYMMETRIC 0x200      /* or-ed in with KM_KERNEL_TL_*: the three-plane tile kernels of a symmetric
                                        * system (KM_FLAG_SYMMETRIC), as the step's p' solve launches them */
int km_time_kernel(km_ctx* ctx, int which, int NX, int NY, const double* coef, const double* b, int reps,
                   double* avg_us);

#ifdef __cplusplus
}
#endif
#endif /* KARMAN_H */
