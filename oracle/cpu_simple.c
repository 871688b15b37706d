/*
 * cpu_simple.c -- C/OpenMP FP64 restatement of the SIMPLE hot path.
 * TEST INFRASTRUCTURE: the timed CPU baseline and a second oracle.
 *
 * It keeps the reference's kernel decomposition, so that its timing stands
 * in for the Taichi CPU path (Taichi 0.9.0 is not installable here): one
 * OpenMP parallel loop per @ti.kernel with the same loop ranges
 * (simple.py:77-238), and the unfused BiCGSTAB iteration of
 * bicgsolver.py:88-136 -- 14 passes over the padded field and 5 reductions
 * per iteration, p_hat = p and s_hat = s as explicit copies, Ashat copied to t.
 * The algebra and boundary/obstacle semantics are those of oracle/fvm.py
 * (same evaluation order; build with -ffp-contract=off).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int32_t nx, ny, east, pin;
    double dx, dy, rho, mu, dt, alpha_u, alpha_p, alpha_m;
    double bc[4][2]; /* w, e, n, s x {normal, tangential} */
} og_grid;

enum { OG_OK = 0, OG_BREAKDOWN = 1, OG_MAXITER = 2, OG_NONFINITE = 3 };

#define IX(i, j, NY) ((size_t)(i) * (size_t)(NY) + (size_t)(j))

/* ------------------------------------------------------------ Krylov */

/* compute_Ap / compute_As (bicgsolver.py:65-76): whole padded range,
 * out-of-array neighbours read as zero */
static void stencil(int NX, int NY, const double* c, const double* x, double* y) {
    const size_t N = (size_t)NX * NY;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NX; ++i)
        for (int j = 0; j < NY; ++j) {
            const size_t k = IX(i, j, NY);
            const double xw = i > 0 ? x[k - NY] : 0.0, xe = i < NX - 1 ? x[k + NY] : 0.0;
            const double xn = j < NY - 1 ? x[k + 1] : 0.0, xs = j > 0 ? x[k - 1] : 0.0;
            y[k] = c[k] * x[k] + c[N + k] * xw + c[2 * N + k] * xe + c[3 * N + k] * xn + c[4 * N + k] * xs;
        }
}

/* CGSolver.reduce (cgsolver.py:49-54) */
static double reduce(size_t n, const double* a, const double* b) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (long k = 0; k < (long)n; ++k) s += a[k] * b[k];
    return s;
}

static void copy(size_t n, const double* a, double* b) {
#pragma omp parallel for schedule(static)
    for (long k = 0; k < (long)n; ++k) b[k] = a[k];
}

/* BICGSolver.solve (bicgsolver.py:88-136), kernel by kernel. */
int og_bicgstab(int NX, int NY, const double* coef, const double* b, double* x, double eps, int max_iter,
                int* iters) {
    const size_t n = (size_t)NX * NY;
    double* w = (double*)malloc(9 * n * sizeof(double));
    if (!w) return -1;
    double *r = w, *rt = w + n, *p = w + 2 * n, *ph = w + 3 * n, *Ap = w + 4 * n, *s = w + 5 * n, *sh = w + 6 * n,
           *t = w + 7 * n, *As = w + 8 * n;
    if (max_iter <= 0) max_iter = (int)n;
    /* init (bicgsolver.py:21-35) */
#pragma omp parallel for schedule(static)
    for (long k = 0; k < (long)n; ++k) {
        r[k] = b[k];
        rt[k] = b[k];
        Ap[k] = 0.0;
        As[k] = 0.0;
        p[k] = 0.0;
        x[k] = 0.0;
    }
    double omega = 1.0, alpha = 1.0, beta = 1.0, rho = 0.0, rho_1 = 1.0;
    const double r0 = reduce(n, r, r);
    int status = OG_MAXITER, it = 0;
    if (r0 == 0.0) { status = OG_OK; goto done; }
    for (int i = 0; i < max_iter; ++i) {
        rho = reduce(n, r, rt);
        if (rho == 0.0) { status = OG_BREAKDOWN; break; }
        if (i == 0) {
            copy(n, r, p);
        } else {
            beta = (rho / rho_1) * (alpha / omega);
#pragma omp parallel for schedule(static)
            for (long k = 0; k < (long)n; ++k) p[k] = r[k] + beta * (p[k] - omega * Ap[k]);
        }
        copy(n, p, ph);                 /* update_phat */
        stencil(NX, NY, coef, ph, Ap);  /* compute_Ap */
        const double rv = reduce(n, rt, Ap);
        if (rv == 0.0) { status = OG_BREAKDOWN; break; }
        alpha = rho / rv;
#pragma omp parallel for schedule(static)
        for (long k = 0; k < (long)n; ++k) s[k] = r[k] - alpha * Ap[k];   /* update_s */
        copy(n, s, sh);                  /* update_shat */
        stencil(NX, NY, coef, sh, As);   /* compute_As */
        copy(n, As, t);                  /* copy(Ashat -> t) */
        const double ts = reduce(n, t, s), tt = reduce(n, t, t);
        omega = tt != 0.0 ? ts / tt : 0.0;
#pragma omp parallel for schedule(static)
        for (long k = 0; k < (long)n; ++k) x[k] = x[k] + (alpha * ph[k] + omega * sh[k]);   /* update_x */
#pragma omp parallel for schedule(static)
        for (long k = 0; k < (long)n; ++k) r[k] = s[k] - omega * t[k];                     /* update_r */
        const double rTr = reduce(n, r, r);
        it = i + 1;
        if (!isfinite(rTr)) { status = OG_NONFINITE; break; }
        if (sqrt(rTr / r0) < eps) { status = OG_OK; break; }
        if (omega == 0.0) { status = OG_BREAKDOWN; break; }
        rho_1 = rho;
    }
done:
    free(w);
    if (iters) *iters = it;
    return status;
}

/* ------------------------------------------------------------ discretisation */

static double bcv(const og_grid* g, int side, int k) { return g->bc[side][k]; }

/* compute_coef_u + obstacle + set_bc u part (simple.py:77-92, :177-207) */
static void assemble_u(const og_grid* g, const double* u, const double* v, const double* p, const double* u0,
                       const uint8_t* m, double* coef, double* bo) {
    const int nx = g->nx, ny = g->ny, NX = nx + 3, NY = ny + 2, NYV = ny + 3, NYP = ny + 2;
    const size_t N = (size_t)NX * NY;
    const double dx = g->dx, dy = g->dy, dt = g->dt, rho = g->rho, mu = g->mu;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NX; ++i)
        for (int j = 0; j < NY; ++j) {
            const size_t k = IX(i, j, NY);
            double ap = 0, aw = 0, ae = 0, an = 0, as = 0, b = 0;
            if (i >= 2 && i <= nx && j >= 1 && j <= ny) {
                const double uP = u[k], uW = u[k - NY], uE = u[k + NY];
                const double vSW = v[IX(i - 1, j, NYV)], vSE = v[IX(i, j, NYV)];
                const double vNW = v[IX(i - 1, j + 1, NYV)], vNE = v[IX(i, j + 1, NYV)];
                aw = -(mu * dy / dx + 0.5 * rho * 0.5 * (uP + uW) * dy);
                ae = -(mu * dy / dx - 0.5 * rho * 0.5 * (uP + uE) * dy);
                an = -(mu * dx / dy - 0.5 * rho * 0.5 * (vNW + vNE) * dx);
                as = -(mu * dx / dy + 0.5 * rho * 0.5 * (vSW + vSE) * dx);
                ap = -(aw + ae + an + as) + rho * 0.5 * (uP + uE) * dy - rho * 0.5 * (uP + uW) * dy +
                     rho * 0.5 * (vNW + vNE) * dx - rho * 0.5 * (vSW + vSE) * dx + rho * dx * dy / dt;
                b = (p[IX(i - 1, j, NYP)] - p[IX(i, j, NYP)]) * dy + rho * dx * dy / dt * u0[k];
                const double wall_x = 2 * mu * dx / dy;
                const size_t cP = IX(i, j, NYP), cW = cP - NYP;
                const int fs = m[cW] | m[cP];
                if (m[cW - NYP]) aw = 0.0;
                if (m[cP + NYP]) ae = 0.0;
                if (m[cW - 1] | m[cP - 1]) { ap = ap + (as + wall_x); as = 0.0; }
                if (m[cW + 1] | m[cP + 1]) { ap = ap + (an + wall_x); an = 0.0; }
                if (i == 2) { b = b + -aw * bcv(g, 0, 0); aw = 0.0; }
                if (i == nx) {
                    if (g->east == 1) ap = ap + ae;
                    else b = b + -ae * bcv(g, 1, 0);
                    ae = 0.0;
                }
                if (j == 1) { b = b + 2 * mu * bcv(g, 3, 1) * dx / dy; ap = ap + (as + 2 * mu * dx / dy); as = 0.0; }
                if (j == ny) { b = b + 2 * mu * bcv(g, 2, 1) * dx / dy; ap = ap + (an + 2 * mu * dx / dy); an = 0.0; }
                if (fs) { ap = 1.0; aw = ae = an = as = b = 0.0; }
            }
            coef[k] = ap; coef[N + k] = aw; coef[2 * N + k] = ae; coef[3 * N + k] = an; coef[4 * N + k] = as;
            bo[k] = b;
        }
}

/* compute_coef_v + obstacle + set_bc v part (simple.py:95-110, :188-216) */
static void assemble_v(const og_grid* g, const double* u, const double* v, const double* p, const double* v0,
                       const uint8_t* m, double* coef, double* bo) {
    const int nx = g->nx, ny = g->ny, NX = nx + 2, NY = ny + 3, NYU = ny + 2, NYP = ny + 2;
    const size_t N = (size_t)NX * NY;
    const double dx = g->dx, dy = g->dy, dt = g->dt, rho = g->rho, mu = g->mu;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NX; ++i)
        for (int j = 0; j < NY; ++j) {
            const size_t k = IX(i, j, NY);
            double ap = 0, aw = 0, ae = 0, an = 0, as = 0, b = 0;
            if (i >= 1 && i <= nx && j >= 2 && j <= ny) {
                const double vP = v[k], vS = v[k - 1], vN = v[k + 1];
                const double uWS = u[IX(i, j - 1, NYU)], uWN = u[IX(i, j, NYU)];
                const double uES = u[IX(i + 1, j - 1, NYU)], uEN = u[IX(i + 1, j, NYU)];
                aw = -(mu * dy / dx + 0.5 * rho * 0.5 * (uWN + uWS) * dy);
                ae = -(mu * dy / dx - 0.5 * rho * 0.5 * (uES + uEN) * dy);
                an = -(mu * dx / dy - 0.5 * rho * 0.5 * (vN + vP) * dx);
                as = -(mu * dx / dy + 0.5 * rho * 0.5 * (vS + vP) * dx);
                ap = -(aw + ae + an + as) + rho * 0.5 * (uES + uEN) * dy - rho * 0.5 * (uWN + uWS) * dy +
                     rho * 0.5 * (vN + vP) * dx - rho * 0.5 * (vS + vP) * dx + rho * dx * dy / dt;
                b = (p[IX(i, j - 1, NYP)] - p[IX(i, j, NYP)]) * dx + rho * dx * dy / dt * v0[k];
                const double wall_y = 2 * mu * dy / dx;
                const size_t cP = IX(i, j, NYP), cS = cP - 1;
                const int fs = m[cS] | m[cP];
                if (m[cS - 1]) as = 0.0;
                if (m[cP + 1]) an = 0.0;
                if (m[cS - NYP] | m[cP - NYP]) { ap = ap + (aw + wall_y); aw = 0.0; }
                if (m[cS + NYP] | m[cP + NYP]) { ap = ap + (ae + wall_y); ae = 0.0; }
                if (j == 2) { b = b + -as * bcv(g, 3, 0); as = 0.0; }
                if (j == ny) { b = b + -an * bcv(g, 2, 0); an = 0.0; }
                if (i == 1) { b = b + 2 * mu * bcv(g, 0, 1) * dy / dx; ap = ap + (aw + 2 * mu * dy / dx); aw = 0.0; }
                if (i == nx) {
                    if (g->east == 1) ap = ap + ae;
                    else { b = b + 2 * mu * bcv(g, 1, 1) * dy / dx; ap = ap + (ae + 2 * mu * dy / dx); }
                    ae = 0.0;
                }
                if (fs) { ap = 1.0; aw = ae = an = as = b = 0.0; }
            }
            coef[k] = ap; coef[N + k] = aw; coef[2 * N + k] = ae; coef[3 * N + k] = an; coef[4 * N + k] = as;
            bo[k] = b;
        }
}

static int dirc(const og_grid* g, const uint8_t* m, int i, int j) {
    if (m[IX(i, j, g->ny + 2)]) return 1;
    if (g->east == 1 && i == g->nx && j >= 1 && j <= g->ny) return 1;
    if (g->pin && i == 1 && j == 1) return 1;
    return 0;
}

/* compute_coef_p (simple.py:143-170) */
static void assemble_p(const og_grid* g, const double* u, const double* v, const double* apu, const double* apv,
                       const uint8_t* m, double* coef, double* bo, double* mdiv) {
    const int nx = g->nx, ny = g->ny, NX = nx + 2, NY = ny + 2, NYU = ny + 2, NYV = ny + 3;
    const size_t N = (size_t)NX * NY;
    const double dx = g->dx, dy = g->dy, rho = g->rho;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NX; ++i)
        for (int j = 0; j < NY; ++j) {
            const size_t k = IX(i, j, NY);
            double ap = 0, aw = 0, ae = 0, an = 0, as = 0, b = 0, md = 0;
            if (i >= 1 && i <= nx && j >= 1 && j <= ny) {
                const size_t ku = IX(i, j, NYU), kv = IX(i, j, NYV);
                md = rho * (u[ku] - u[ku + NYU]) * dy + rho * (v[kv] - v[kv + 1]) * dx;
                if (!(i == 1 || m[k - NY])) aw = -rho * dy * dy / apu[ku];
                if (!(i == nx || m[k + NY])) ae = -rho * dy * dy / apu[ku + NYU];
                if (!(j == ny || m[k + 1])) an = -rho * dx * dx / apv[kv + 1];
                if (!(j == 1 || m[k - 1])) as = -rho * dx * dx / apv[kv];
                ap = -(aw + ae + an + as);
                if (dirc(g, m, i - 1, j)) aw = 0.0;
                if (dirc(g, m, i + 1, j)) ae = 0.0;
                if (dirc(g, m, i, j + 1)) an = 0.0;
                if (dirc(g, m, i, j - 1)) as = 0.0;
                b = md;
                if (dirc(g, m, i, j)) { ap = 1.0; aw = ae = an = as = b = 0.0; }
            }
            coef[k] = ap; coef[N + k] = aw; coef[2 * N + k] = ae; coef[3 * N + k] = an; coef[4 * N + k] = as;
            bo[k] = b;
            mdiv[k] = md;
        }
}

/* One SIMPLE outer iteration, same order as oracle.fvm.simple_step.
 * iters[3] receives the BiCGSTAB iteration counts for u, v, p'. */
int og_simple_step(const og_grid* g, double* u, double* v, double* p, const double* u0, const double* v0,
                   const uint8_t* m, double* pcor, double* mdiv, double eps_mom, double eps_p, int max_iter,
                   int* iters) {
    const int nx = g->nx, ny = g->ny;
    const size_t Nu = (size_t)(nx + 3) * (ny + 2), Nv = (size_t)(nx + 2) * (ny + 3), Np = (size_t)(nx + 2) * (ny + 2);
    double* w = (double*)malloc((7 * Nu + 7 * Nv + 7 * Np) * sizeof(double));
    if (!w) return -1;
    double *cu = w, *bu = cu + 5 * Nu, *um = bu + Nu, *cv = um + Nu, *bv = cv + 5 * Nv, *vm = bv + Nv;
    double *cp = vm + Nv, *bp = cp + 5 * Np;
    const int NYU = ny + 2, NYV = ny + 3, NYP = ny + 2;
    assemble_u(g, u, v, p, u0, m, cu, bu);
    assemble_v(g, u, v, p, v0, m, cv, bv);
    for (int j = 1; j <= ny; ++j) {   /* set_bc ghosts (simple.py:181, :186) */
        u[IX(1, j, NYU)] = bcv(g, 0, 0);
        if (g->east != 1) u[IX(nx + 1, j, NYU)] = bcv(g, 1, 0);
    }
    for (int i = 1; i <= nx; ++i) {   /* (simple.py:192, :196) */
        v[IX(i, 1, NYV)] = bcv(g, 3, 0);
        v[IX(i, ny + 1, NYV)] = bcv(g, 2, 0);
    }
    int st = og_bicgstab(nx + 3, ny + 2, cu, bu, um, eps_mom, max_iter, &iters[0]);
    int st2 = og_bicgstab(nx + 2, ny + 3, cv, bv, vm, eps_mom, max_iter, &iters[1]);
    const double a = g->alpha_m;   /* update_velocity (simple.py:232-238) */
#pragma omp parallel for schedule(static)
    for (int i = 2; i <= nx; ++i)
        for (int j = 1; j <= ny; ++j) u[IX(i, j, NYU)] = a * um[IX(i, j, NYU)] + (1 - a) * u[IX(i, j, NYU)];
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= nx; ++i)
        for (int j = 2; j <= ny; ++j) v[IX(i, j, NYV)] = a * vm[IX(i, j, NYV)] + (1 - a) * v[IX(i, j, NYV)];
    if (g->east == 1) {   /* outlet fix (karman_taichi.py:180-183, :586-614) */
        double fin = 0.0, fout = 0.0;
        for (int j = 1; j <= ny; ++j) u[IX(nx + 1, j, NYU)] = u[IX(nx, j, NYU)];
        for (int j = 1; j <= ny; ++j) { fin += u[IX(1, j, NYU)]; fout += u[IX(nx + 1, j, NYU)]; }
        const double shift = (fin - fout) / ny;
        for (int j = 1; j <= ny; ++j) u[IX(nx + 1, j, NYU)] += shift;
    }
    /* apu / apv are plane 0 of the momentum coefficients */
    assemble_p(g, u, v, cu, cv, m, cp, bp, mdiv);
    int st3 = og_bicgstab(nx + 2, ny + 2, cp, bp, pcor, eps_p, max_iter, &iters[2]);
    /* correct_pressure, correct_velocity (simple.py:218-230) */
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= nx; ++i)
        for (int j = 1; j <= ny; ++j)
            if (!m[IX(i, j, NYP)]) p[IX(i, j, NYP)] = p[IX(i, j, NYP)] + g->alpha_p * pcor[IX(i, j, NYP)];
#pragma omp parallel for schedule(static)
    for (int i = 2; i <= nx; ++i)
        for (int j = 1; j <= ny; ++j) {
            const size_t kc = IX(i, j, NYP), ku = IX(i, j, NYU);
            if (!(m[kc - NYP] | m[kc])) u[ku] = u[ku] + g->alpha_u * (pcor[kc - NYP] - pcor[kc]) * g->dy / cu[ku];
        }
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= nx; ++i)
        for (int j = 2; j <= ny; ++j) {
            const size_t kc = IX(i, j, NYP), kv = IX(i, j, NYV);
            if (!(m[kc - 1] | m[kc])) v[kv] = v[kv] + g->alpha_u * (pcor[kc - 1] - pcor[kc]) * g->dx / cv[kv];
        }
    /* compute_mdiv (simple.py:137-141) */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx + 2; ++i)
        for (int j = 0; j < ny + 2; ++j) {
            double md = 0.0;
            if (i >= 1 && i <= nx && j >= 1 && j <= ny)
                md = g->rho * (u[IX(i, j, NYU)] - u[IX(i + 1, j, NYU)]) * g->dy +
                     g->rho * (v[IX(i, j, NYV)] - v[IX(i, j + 1, NYV)]) * g->dx;
            mdiv[IX(i, j, NYP)] = md;
        }
    free(w);
    return st != OG_OK ? st : (st2 != OG_OK ? st2 : st3);
}

int og_num_threads(void) {
#ifdef _OPENMP
    extern int omp_get_max_threads(void);
    return omp_get_max_threads();
#else
    return 1;
#endif
}
