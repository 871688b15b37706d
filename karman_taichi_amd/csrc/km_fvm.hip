// Staggered-grid finite-volume kernels of the SIMPLE step for gfx950.
//
// One thread per padded array element, flat i-major index (j contiguous), so
// every load and store of the centre element is a coalesced HBM sweep.  The
// formulas keep the reference's algebraic form and evaluation order
// (simple.py:77-238) and the library is built with -ffp-contract=off, so the
// coefficients are bit-identical to an IEEE FP64 evaluation of the same
// expressions.  Boundary conditions (set_bc, simple.py:172-216) and the
// obstacle mask are applied in the same kernel as the assembly instead of a
// second pass over the boundary rows.
#include "km_common.h"

namespace {

// ---------------------------------------------------------------- u momentum
// Faces i in [2, nx], j in [1, ny] (simple.py:80); every other padded entry
// gets an all-zero row and b = 0.
__device__ __forceinline__ void assemble_u_body(GridD g, const double* __restrict__ u, const double* __restrict__ v,
             const double* __restrict__ p, const double* __restrict__ u0,
             const uint8_t* __restrict__ m, double* __restrict__ coef, double* __restrict__ bout, int bx) {
    const int NX = g.nx + 3, NY = g.ny + 2;           // u shape
    const int NYV = g.ny + 3, NYP = g.ny + 2;         // v and p/mask row pitch
    const int64_t N = (int64_t)NX * NY;
    const int64_t k = (int64_t)bx * KM_BLOCK + threadIdx.x;
    if (k >= N) return;
    const int i = (int)(k / NY), j = (int)(k - (int64_t)i * NY);
    double ap = 0.0, aw = 0.0, ae = 0.0, an = 0.0, as = 0.0, b = 0.0;
    if (i >= 2 && i <= g.nx && j >= 1 && j <= g.ny) {
        const double nx_unused = 0.0; (void)nx_unused;
        const double dx = g.dx, dy = g.dy, dt = g.dt, rho = g.rho, mu = g.mu;
        const double uP = u[k], uW = u[k - NY], uE = u[k + NY];
        const double vSW = v[(int64_t)(i - 1) * NYV + j], vSE = v[(int64_t)i * NYV + j];
        const double vNW = v[(int64_t)(i - 1) * NYV + j + 1], vNE = v[(int64_t)i * NYV + j + 1];
        aw = -(mu * dy / dx + 0.5 * rho * 0.5 * (uP + uW) * dy);
        ae = -(mu * dy / dx - 0.5 * rho * 0.5 * (uP + uE) * dy);
        an = -(mu * dx / dy - 0.5 * rho * 0.5 * (vNW + vNE) * dx);
        as = -(mu * dx / dy + 0.5 * rho * 0.5 * (vSW + vSE) * dx);
        ap = -(aw + ae + an + as)
             + rho * 0.5 * (uP + uE) * dy
             - rho * 0.5 * (uP + uW) * dy
             + rho * 0.5 * (vNW + vNE) * dx
             - rho * 0.5 * (vSW + vSE) * dx
             + rho * dx * dy / dt;
        b = (p[(int64_t)(i - 1) * NYP + j] - p[(int64_t)i * NYP + j]) * dy + rho * dx * dy / dt * u0[k];

        // obstacle (karman_taichi.py:220-248): normal neighbours that are
        // solid faces drop out, tangential neighbours across the surface get
        // the wall shear of simple.py:198-207
        const double wall_x = 2 * mu * dx / dy;
        const int64_t cP = (int64_t)i * NYP + j, cW = cP - NYP;
        const bool face_solid = m[cW] | m[cP];
        if (m[cW - NYP]) aw = 0.0;
        if (m[cP + NYP]) ae = 0.0;
        if (m[cW - 1] | m[cP - 1]) { ap = ap + (as + wall_x); as = 0.0; }
        if (m[cW + 1] | m[cP + 1]) { ap = ap + (an + wall_x); an = 0.0; }

        // set_bc, u part (simple.py:177-207): west/east folds, then walls
        if (i == 2) { b = b + -aw * g.bcw_n; aw = 0.0; }
        if (i == g.nx) {
            if (g.east == KM_EAST_OUTFLOW) ap = ap + ae;
            else b = b + -ae * g.bce_n;
            ae = 0.0;
        }
        if (j == 1) { b = b + 2 * mu * g.bcs_t * dx / dy; ap = ap + (as + 2 * mu * dx / dy); as = 0.0; }
        if (j == g.ny) { b = b + 2 * mu * g.bcn_t * dx / dy; ap = ap + (an + 2 * mu * dx / dy); an = 0.0; }

        if (face_solid) { ap = 1.0; aw = ae = an = as = b = 0.0; }
    }
    coef[k] = ap;
    coef[N + k] = aw;
    coef[2 * N + k] = ae;
    coef[3 * N + k] = an;
    coef[4 * N + k] = as;
    bout[k] = b;
}

__global__ void __launch_bounds__(KM_BLOCK)
k_assemble_u(GridD g, const double* __restrict__ u, const double* __restrict__ v,
             const double* __restrict__ p, const double* __restrict__ u0,
             const uint8_t* __restrict__ m, double* __restrict__ coef, double* __restrict__ bout) {
    assemble_u_body(g, u, v, p, u0, m, coef, bout, blockIdx.x);
}


// ---------------------------------------------------------------- v momentum
// Faces i in [1, nx], j in [2, ny] (simple.py:98).
__device__ __forceinline__ void assemble_v_body(GridD g, const double* __restrict__ u, const double* __restrict__ v,
             const double* __restrict__ p, const double* __restrict__ v0,
             const uint8_t* __restrict__ m, double* __restrict__ coef, double* __restrict__ bout, int bx) {
    const int NX = g.nx + 2, NY = g.ny + 3;           // v shape
    const int NYU = g.ny + 2, NYP = g.ny + 2;
    const int64_t N = (int64_t)NX * NY;
    const int64_t k = (int64_t)bx * KM_BLOCK + threadIdx.x;
    if (k >= N) return;
    const int i = (int)(k / NY), j = (int)(k - (int64_t)i * NY);
    double ap = 0.0, aw = 0.0, ae = 0.0, an = 0.0, as = 0.0, b = 0.0;
    if (i >= 1 && i <= g.nx && j >= 2 && j <= g.ny) {
        const double dx = g.dx, dy = g.dy, dt = g.dt, rho = g.rho, mu = g.mu;
        const double vP = v[k], vS = v[k - 1], vN = v[k + 1];
        const double uWS = u[(int64_t)i * NYU + j - 1], uWN = u[(int64_t)i * NYU + j];
        const double uES = u[(int64_t)(i + 1) * NYU + j - 1], uEN = u[(int64_t)(i + 1) * NYU + j];
        aw = -(mu * dy / dx + 0.5 * rho * 0.5 * (uWN + uWS) * dy);
        ae = -(mu * dy / dx - 0.5 * rho * 0.5 * (uES + uEN) * dy);
        an = -(mu * dx / dy - 0.5 * rho * 0.5 * (vN + vP) * dx);
        as = -(mu * dx / dy + 0.5 * rho * 0.5 * (vS + vP) * dx);
        ap = -(aw + ae + an + as)
             + rho * 0.5 * (uES + uEN) * dy
             - rho * 0.5 * (uWN + uWS) * dy
             + rho * 0.5 * (vN + vP) * dx
             - rho * 0.5 * (vS + vP) * dx
             + rho * dx * dy / dt;
        b = (p[(int64_t)i * NYP + j - 1] - p[(int64_t)i * NYP + j]) * dx + rho * dx * dy / dt * v0[k];

        const double wall_y = 2 * mu * dy / dx;
        const int64_t cP = (int64_t)i * NYP + j, cS = cP - 1;
        const bool face_solid = m[cS] | m[cP];
        if (m[cS - 1]) as = 0.0;
        if (m[cP + 1]) an = 0.0;
        if (m[cS - NYP] | m[cP - NYP]) { ap = ap + (aw + wall_y); aw = 0.0; }
        if (m[cS + NYP] | m[cP + NYP]) { ap = ap + (ae + wall_y); ae = 0.0; }

        // set_bc, v part (simple.py:188-196, :209-216 with bicgsimp.py:196)
        if (j == 2) { b = b + -as * g.bcs_n; as = 0.0; }
        if (j == g.ny) { b = b + -an * g.bcn_n; an = 0.0; }
        if (i == 1) { b = b + 2 * mu * g.bcw_t * dy / dx; ap = ap + (aw + 2 * mu * dy / dx); aw = 0.0; }
        if (i == g.nx) {
            if (g.east == KM_EAST_OUTFLOW) ap = ap + ae;
            else { b = b + 2 * mu * g.bce_t * dy / dx; ap = ap + (ae + 2 * mu * dy / dx); }
            ae = 0.0;
        }

        if (face_solid) { ap = 1.0; aw = ae = an = as = b = 0.0; }
    }
    coef[k] = ap;
    coef[N + k] = aw;
    coef[2 * N + k] = ae;
    coef[3 * N + k] = an;
    coef[4 * N + k] = as;
    bout[k] = b;
}

__global__ void __launch_bounds__(KM_BLOCK)
k_assemble_v(GridD g, const double* __restrict__ u, const double* __restrict__ v,
             const double* __restrict__ p, const double* __restrict__ v0,
             const uint8_t* __restrict__ m, double* __restrict__ coef, double* __restrict__ bout) {
    assemble_v_body(g, u, v, p, v0, m, coef, bout, blockIdx.x);
}


// Ghost velocities written by set_bc (simple.py:181, :186, :192, :196).
__device__ __forceinline__ void ghosts_body(GridD g, double* __restrict__ u, double* __restrict__ v, int bx) {
    const int t = bx * KM_BLOCK + threadIdx.x;
    const int NYU = g.ny + 2, NYV = g.ny + 3;
    if (t < g.ny) {
        const int j = t + 1;
        u[(int64_t)1 * NYU + j] = g.bcw_n;
        if (g.east != KM_EAST_OUTFLOW) u[(int64_t)(g.nx + 1) * NYU + j] = g.bce_n;
    }
    if (t < g.nx) {
        const int i = t + 1;
        v[(int64_t)i * NYV + 1] = g.bcs_n;
        v[(int64_t)i * NYV + g.ny + 1] = g.bcn_n;
    }
}

__global__ void __launch_bounds__(KM_BLOCK)
k_ghosts(GridD g, double* __restrict__ u, double* __restrict__ v) {
    ghosts_body(g, u, v, blockIdx.x);
}


// update_velocity (simple.py:232-238): u faces i in [2,nx] x [1,ny], then v
// faces i in [1,nx] x [2,ny], as one flat launch.
__device__ __forceinline__ void blend_body(GridD g, double* __restrict__ u, double* __restrict__ v,
        const double* __restrict__ um, const double* __restrict__ vm, int bx) {
    const int64_t t = (int64_t)bx * KM_BLOCK + threadIdx.x;
    const int64_t Nu = (int64_t)(g.nx - 1) * g.ny, Nv = (int64_t)g.nx * (g.ny - 1);
    const double a = g.alpha_m;
    if (t < Nu) {
        const int i = 2 + (int)(t / g.ny), j = 1 + (int)(t % g.ny);
        const int64_t k = (int64_t)i * (g.ny + 2) + j;
        u[k] = a * um[k] + (1 - a) * u[k];
    } else if (t < Nu + Nv) {
        const int64_t q = t - Nu;
        const int i = 1 + (int)(q / (g.ny - 1)), j = 2 + (int)(q % (g.ny - 1));
        const int64_t k = (int64_t)i * (g.ny + 3) + j;
        v[k] = a * vm[k] + (1 - a) * v[k];
    }
}

__global__ void __launch_bounds__(KM_BLOCK)
k_blend(GridD g, double* __restrict__ u, double* __restrict__ v,
        const double* __restrict__ um, const double* __restrict__ vm) {
    blend_body(g, u, v, um, vm, blockIdx.x);
}


// Zero-gradient outlet and flux fix (karman_taichi.py:180-183, :586-614).
// One workgroup; the column sums run sequentially in one lane so that they
// are the same sums, in the same order, as the reference's serial loops
// (the columns are staged in LDS by the whole workgroup first).
__device__ __forceinline__ void outflow_body(const GridD& g, double* __restrict__ u) {
    constexpr int kStageCells = 1024;
    const int NYU = g.ny + 2;
    double* uin = u + (int64_t)1 * NYU;
    double* ulast = u + (int64_t)g.nx * NYU;
    double* uout = u + (int64_t)(g.nx + 1) * NYU;
    __shared__ double shift;
    __shared__ double sin_[kStageCells], sout[kStageCells];
    double fin = 0.0, fout = 0.0;   // thread 0's running sums
    for (int j0 = 1; j0 <= g.ny; j0 += kStageCells) {
        const int cnt = g.ny - j0 + 1 < kStageCells ? g.ny - j0 + 1 : kStageCells;
        __syncthreads();   // thread 0 is done with the previous chunk
        for (int t = threadIdx.x; t < cnt; t += KM_BLOCK) {
            const double ul = ulast[j0 + t];
            uout[j0 + t] = ul;   // zero gradient
            sin_[t] = uin[j0 + t];
            sout[t] = ul;
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int t = 0; t < cnt; ++t) {
                fin += sin_[t];
                fout += sout[t];
            }
    }
    if (threadIdx.x == 0) shift = (fin - fout) / g.ny;
    __syncthreads();
    for (int j = 1 + threadIdx.x; j <= g.ny; j += KM_BLOCK) uout[j] += shift;
}

__global__ void __launch_bounds__(KM_BLOCK) k_outflow(GridD g, double* __restrict__ u) { outflow_body(g, u); }

__device__ __forceinline__ bool dirichlet_cell(const GridD& g, const uint8_t* m, int i, int j) {
    if (m[(int64_t)i * (g.ny + 2) + j]) return true;
    if (g.east == KM_EAST_OUTFLOW && i == g.nx && j >= 1 && j <= g.ny) return true;
    if (g.pin && i == 1 && j == 1) return true;
    return false;
}

// compute_coef_p (simple.py:143-170) with the guards, mask and p' Dirichlet
// cells; writes mdiv over the interior and zero on the padding.
__device__ __forceinline__ void assemble_p_body(GridD g, const double* __restrict__ u, const double* __restrict__ v,
             const double* __restrict__ apu, const double* __restrict__ apv,
             const uint8_t* __restrict__ m, double* __restrict__ coef, double* __restrict__ bout,
             double* __restrict__ mdiv, int bx) {
    const int NX = g.nx + 2, NY = g.ny + 2, NYU = g.ny + 2, NYV = g.ny + 3;
    const int64_t N = (int64_t)NX * NY;
    const int64_t k = (int64_t)bx * KM_BLOCK + threadIdx.x;
    if (k >= N) return;
    const int i = (int)(k / NY), j = (int)(k - (int64_t)i * NY);
    double ap = 0.0, aw = 0.0, ae = 0.0, an = 0.0, as = 0.0, b = 0.0, md = 0.0;
    if (i >= 1 && i <= g.nx && j >= 1 && j <= g.ny) {
        const double dx = g.dx, dy = g.dy, rho = g.rho;
        const int64_t ku = (int64_t)i * NYU + j, kv = (int64_t)i * NYV + j;
        md = rho * (u[ku] - u[ku + NYU]) * dy + rho * (v[kv] - v[kv + 1]) * dx;
        if (!(i == 1 || m[k - NY])) aw = -rho * dy * dy / apu[ku];
        if (!(i == g.nx || m[k + NY])) ae = -rho * dy * dy / apu[ku + NYU];
        if (!(j == g.ny || m[k + 1])) an = -rho * dx * dx / apv[kv + 1];
        if (!(j == 1 || m[k - 1])) as = -rho * dx * dx / apv[kv];
        ap = -(aw + ae + an + as);
        if (dirichlet_cell(g, m, i - 1, j)) aw = 0.0;
        if (dirichlet_cell(g, m, i + 1, j)) ae = 0.0;
        if (dirichlet_cell(g, m, i, j + 1)) an = 0.0;
        if (dirichlet_cell(g, m, i, j - 1)) as = 0.0;
        b = md;
        if (dirichlet_cell(g, m, i, j)) { ap = 1.0; aw = ae = an = as = b = 0.0; }
    }
    coef[k] = ap;
    coef[N + k] = aw;
    coef[2 * N + k] = ae;
    coef[3 * N + k] = an;
    coef[4 * N + k] = as;
    bout[k] = b;
    mdiv[k] = md;
}

__global__ void __launch_bounds__(KM_BLOCK)
k_assemble_p(GridD g, const double* __restrict__ u, const double* __restrict__ v,
             const double* __restrict__ apu, const double* __restrict__ apv,
             const uint8_t* __restrict__ m, double* __restrict__ coef, double* __restrict__ bout,
             double* __restrict__ mdiv) {
    assemble_p_body(g, u, v, apu, apv, m, coef, bout, mdiv, blockIdx.x);
}


// correct_pressure + correct_velocity (simple.py:218-230): p cells, then u
// faces, then v faces as one flat launch; solid cells/faces are skipped
// (karman_taichi.py:515, :527, :536).
__device__ __forceinline__ void correct_body(GridD g, double* __restrict__ p, double* __restrict__ u, double* __restrict__ v,
          const double* __restrict__ pc, const double* __restrict__ apu, const double* __restrict__ apv,
          const uint8_t* __restrict__ m, int parts, int bx) {
    const int64_t t = (int64_t)bx * KM_BLOCK + threadIdx.x;
    const int nx = g.nx, ny = g.ny, NYP = ny + 2;
    const int64_t Np = (int64_t)nx * ny, Nu = (int64_t)(nx - 1) * ny, Nv = (int64_t)nx * (ny - 1);
    if (t < Np) {
        if (!(parts & 1)) return;
        const int i = 1 + (int)(t / ny), j = 1 + (int)(t % ny);
        const int64_t k = (int64_t)i * NYP + j;
        if (!m[k]) p[k] = p[k] + g.alpha_p * pc[k];
    } else if (!(parts & 2)) {
        return;
    } else if (t < Np + Nu) {
        const int64_t q = t - Np;
        const int i = 2 + (int)(q / ny), j = 1 + (int)(q % ny);
        const int64_t kc = (int64_t)i * NYP + j, ku = (int64_t)i * (ny + 2) + j;
        if (!(m[kc - NYP] | m[kc]))
            u[ku] = u[ku] + g.alpha_u * (pc[kc - NYP] - pc[kc]) * g.dy / apu[ku];
    } else if (t < Np + Nu + Nv) {
        const int64_t q = t - Np - Nu;
        const int i = 1 + (int)(q / (ny - 1)), j = 2 + (int)(q % (ny - 1));
        const int64_t kc = (int64_t)i * NYP + j, kv = (int64_t)i * (ny + 3) + j;
        if (!(m[kc - 1] | m[kc]))
            v[kv] = v[kv] + g.alpha_u * (pc[kc - 1] - pc[kc]) * g.dx / apv[kv];
    }
}

__global__ void __launch_bounds__(KM_BLOCK)
k_correct(GridD g, double* __restrict__ p, double* __restrict__ u, double* __restrict__ v,
          const double* __restrict__ pc, const double* __restrict__ apu, const double* __restrict__ apv,
          const uint8_t* __restrict__ m, int parts) {
    correct_body(g, p, u, v, pc, apu, apv, m, parts, blockIdx.x);
}


// compute_mdiv (simple.py:137-141) with per-workgroup partials of max|mdiv|
// and sum mdiv^2 over the interior (bicgsimp.py:114-122 returns the max).
__device__ __forceinline__ void mdiv_body(GridD g, Flat fl, const double* __restrict__ u, const double* __restrict__ v, double* __restrict__ mdiv,
       double* __restrict__ P, int bx) {
    const int NY = g.ny + 2, NYU = g.ny + 2, NYV = g.ny + 3;
    __shared__ double sh[4];
    double mx = 0.0, sq = 0.0;
    const int64_t k0 = (int64_t)bx * fl.chunk;
    const int64_t k1 = k0 + fl.chunk < fl.n ? k0 + fl.chunk : fl.n;
    for (int64_t k = k0 + threadIdx.x; k < k1; k += KM_BLOCK) {
        const int i = (int)(k / NY), j = (int)(k - (int64_t)i * NY);
        double md = 0.0;
        if (i >= 1 && i <= g.nx && j >= 1 && j <= g.ny) {
            const int64_t ku = (int64_t)i * NYU + j, kv = (int64_t)i * NYV + j;
            md = g.rho * (u[ku] - u[ku + NYU]) * g.dy + g.rho * (v[kv] - v[kv + 1]) * g.dx;
        }
        mdiv[k] = md;
        mx = fmax(mx, fabs(md));
        sq += md * md;
    }
    const double bmx = block_max(mx, sh);
    const double bsq = block_sum_all(sq, sh);
    if (threadIdx.x == 0) {
        P[bx] = bmx;
        P[KM_MAX_RED + bx] = bsq;
    }
}

__global__ void __launch_bounds__(KM_BLOCK)
k_mdiv(GridD g, Flat fl, const double* __restrict__ u, const double* __restrict__ v, double* __restrict__ mdiv,
       double* __restrict__ P) {
    mdiv_body(g, fl, u, v, mdiv, P, blockIdx.x);
}


// One workgroup: res[0] = max of the max partials, res[1] = sum of the sums.
__global__ void __launch_bounds__(KM_BLOCK)
k_mdiv_finish(const double* __restrict__ P, int G, double* __restrict__ res) {
    __shared__ double sh[4];
    double mx = 0.0;
    for (int k = threadIdx.x; k < G; k += KM_BLOCK) mx = fmax(mx, P[k]);
    mx = block_max(mx, sh);
    double s[1];
    sum_partials<1>(P + KM_MAX_RED, G, s, sh);
    if (threadIdx.x == 0) {
        res[0] = mx;
        res[1] = s[0];
    }
}

__device__ __forceinline__ void time_forward_body(int64_t Nu, int64_t Nv, const double* __restrict__ u, const double* __restrict__ v,
               double* __restrict__ u0, double* __restrict__ v0, int bx) {
    const int64_t t = (int64_t)bx * KM_BLOCK + threadIdx.x;
    if (t < Nu) u0[t] = u[t];
    else if (t < Nu + Nv) v0[t - Nu] = v[t - Nu];
}

__global__ void __launch_bounds__(KM_BLOCK)
k_time_forward(int64_t Nu, int64_t Nv, const double* __restrict__ u, const double* __restrict__ v,
               double* __restrict__ u0, double* __restrict__ v0) {
    time_forward_body(Nu, Nv, u, v, u0, v0, blockIdx.x);
}


// ---------------------------------------------------------------- batched stages
// One launch per stage over every case of a batch: workgroup bx belongs to
// case find_case(starts, n, bx) and runs that case's body with its local
// workgroup index, so each case sees exactly the single-case kernel.
#define KB_CASE                                          \
    const int c_ = find_case(starts, n, blockIdx.x);     \
    const BCase& C = cs[c_];                             \
    const int bx = blockIdx.x - starts[c_];

__global__ void __launch_bounds__(KM_BLOCK) kb_assemble_u(const BCase* __restrict__ cs, const int* __restrict__ starts, int n) {
    KB_CASE
    assemble_u_body(C.g, C.u, C.v, C.p, C.u0, C.m, C.coef_u, C.b_u, bx);
}
__global__ void __launch_bounds__(KM_BLOCK) kb_assemble_v(const BCase* __restrict__ cs, const int* __restrict__ starts, int n) {
    KB_CASE
    assemble_v_body(C.g, C.u, C.v, C.p, C.v0, C.m, C.coef_v, C.b_v, bx);
}
__global__ void __launch_bounds__(KM_BLOCK) kb_ghosts(const BCase* __restrict__ cs, const int* __restrict__ starts, int n) {
    KB_CASE
    ghosts_body(C.g, C.u, C.v, bx);
}
__global__ void __launch_bounds__(KM_BLOCK) kb_blend(const BCase* __restrict__ cs, const int* __restrict__ starts, int n) {
    KB_CASE
    blend_body(C.g, C.u, C.v, C.umid, C.vmid, bx);
}
__global__ void __launch_bounds__(KM_BLOCK) kb_outflow(const BCase* __restrict__ cs, const int* __restrict__ starts, int n) {
    KB_CASE
    (void)bx;
    if (C.g.east == KM_EAST_OUTFLOW) outflow_body(C.g, C.u);
}
__global__ void __launch_bounds__(KM_BLOCK) kb_assemble_p(const BCase* __restrict__ cs, const int* __restrict__ starts, int n) {
    KB_CASE
    assemble_p_body(C.g, C.u, C.v, C.coef_u, C.coef_v, C.m, C.coef_p, C.b_p, C.mdiv, bx);
}
__global__ void __launch_bounds__(KM_BLOCK)
kb_correct(const BCase* __restrict__ cs, const int* __restrict__ starts, int n, int parts) {
    KB_CASE
    correct_body(C.g, C.p, C.u, C.v, C.pcor, C.coef_u, C.coef_v, C.m, parts, bx);
}
// mdiv of every case; partials of case c go to P + c * 2 * KM_MAX_RED
__global__ void __launch_bounds__(KM_BLOCK)
kb_mdiv(const BCase* __restrict__ cs, const int* __restrict__ starts, int n, double* __restrict__ P) {
    KB_CASE
    const Flat fl = make_flat((int64_t)(C.g.nx + 2) * (C.g.ny + 2));
    mdiv_body(C.g, fl, C.u, C.v, C.mdiv, P + (int64_t)c_ * 2 * KM_MAX_RED, bx);
}
__global__ void __launch_bounds__(KM_BLOCK) kb_time_forward(const BCase* __restrict__ cs, const int* __restrict__ starts, int n) {
    KB_CASE
    const int64_t Nu = (int64_t)(C.g.nx + 3) * (C.g.ny + 2), Nv = (int64_t)(C.g.nx + 2) * (C.g.ny + 3);
    time_forward_body(Nu, Nv, C.u, C.v, C.u0, C.v0, bx);
}
#undef KB_CASE

int check_grid(km_ctx* ctx, const km_grid* g) {
    KM_REQUIRE(ctx, g != nullptr, "grid is NULL");
    KM_REQUIRE(ctx, g->nx >= 4 && g->ny >= 3, "grid too small (nx=%d ny=%d; need nx>=4, ny>=3)", g->nx, g->ny);
    KM_REQUIRE(ctx, (int64_t)(g->nx + 3) * (g->ny + 3) < (int64_t)1 << 31, "grid too large");
    KM_REQUIRE(ctx, g->dx > 0 && g->dy > 0 && g->dt > 0 && g->rho > 0, "dx, dy, dt and rho must be positive");
    KM_REQUIRE(ctx, g->east == KM_EAST_WALL || g->east == KM_EAST_OUTFLOW, "unknown east boundary kind %d", g->east);
    return KM_OK;
}

}  // namespace

// Workgroups one case needs in a stage (the single-case launch geometry).
int64_t batch_stage_blocks(int stage, const GridD& g) {
    const int64_t nx = g.nx, ny = g.ny;
    switch (stage) {
        case BS_ASM_U: return (int64_t)ceil_div_i64((nx + 3) * (ny + 2), KM_BLOCK);
        case BS_ASM_V: return (int64_t)ceil_div_i64((nx + 2) * (ny + 3), KM_BLOCK);
        case BS_GHOSTS: return (int64_t)ceil_div_i64(nx > ny ? nx : ny, KM_BLOCK);
        case BS_BLEND: return (int64_t)ceil_div_i64((nx - 1) * ny + nx * (ny - 1), KM_BLOCK);
        case BS_OUTFLOW: return 1;
        case BS_ASM_P: return (int64_t)ceil_div_i64((nx + 2) * (ny + 2), KM_BLOCK);
        case BS_CORRECT: return (int64_t)ceil_div_i64(nx * ny + (nx - 1) * ny + nx * (ny - 1), KM_BLOCK);
        case BS_MDIV: return make_flat((nx + 2) * (ny + 2)).G;
        case BS_TIME: return (int64_t)ceil_div_i64((nx + 3) * (ny + 2) + (nx + 2) * (ny + 3), KM_BLOCK);
    }
    return 0;
}

int batch_stage_launch(km_ctx* ctx, int stage, const BCase* cs, const int* starts, int n, int64_t blocks, int parts,
                       double* P, hipStream_t s) {
    KM_REQUIRE(ctx, blocks > 0 && blocks < ((int64_t)1 << 31), "batched launch of %lld workgroups", (long long)blocks);
    const dim3 grid((unsigned)blocks), blk(KM_BLOCK);
    switch (stage) {
        case BS_ASM_U: hipLaunchKernelGGL(kb_assemble_u, grid, blk, 0, s, cs, starts, n); break;
        case BS_ASM_V: hipLaunchKernelGGL(kb_assemble_v, grid, blk, 0, s, cs, starts, n); break;
        case BS_GHOSTS: hipLaunchKernelGGL(kb_ghosts, grid, blk, 0, s, cs, starts, n); break;
        case BS_BLEND: hipLaunchKernelGGL(kb_blend, grid, blk, 0, s, cs, starts, n); break;
        case BS_OUTFLOW: hipLaunchKernelGGL(kb_outflow, grid, blk, 0, s, cs, starts, n); break;
        case BS_ASM_P: hipLaunchKernelGGL(kb_assemble_p, grid, blk, 0, s, cs, starts, n); break;
        case BS_CORRECT: hipLaunchKernelGGL(kb_correct, grid, blk, 0, s, cs, starts, n, parts); break;
        case BS_MDIV: hipLaunchKernelGGL(kb_mdiv, grid, blk, 0, s, cs, starts, n, P); break;
        case BS_TIME: hipLaunchKernelGGL(kb_time_forward, grid, blk, 0, s, cs, starts, n); break;
        default: return km_fail(ctx, KM_EINVAL, "unknown batch stage %d", stage);
    }
    KM_CHECK_LAUNCH(ctx);
    return KM_OK;
}

extern "C" {

int km_assemble_u(km_ctx* ctx, const km_grid* g, const double* u, const double* v, const double* p,
                  const double* u0, const uint8_t* mask, double* coef_u, double* b_u) {
    if (!ctx) return KM_EINVAL;
    if (int rc = check_grid(ctx, g)) return rc;
    KM_REQUIRE(ctx, u && v && p && u0 && mask && coef_u && b_u, "km_assemble_u: NULL pointer");
    const int64_t N = (int64_t)(g->nx + 3) * (g->ny + 2);
    hipLaunchKernelGGL(k_assemble_u, dim3(ceil_div_i64(N, KM_BLOCK)), dim3(KM_BLOCK), 0, ctx->stream,
                       to_dev(g), u, v, p, u0, mask, coef_u, b_u);
    KM_CHECK_LAUNCH(ctx);
    return KM_OK;
}

int km_assemble_v(km_ctx* ctx, const km_grid* g, const double* u, const double* v, const double* p,
                  const double* v0, const uint8_t* mask, double* coef_v, double* b_v) {
    if (!ctx) return KM_EINVAL;
    if (int rc = check_grid(ctx, g)) return rc;
    KM_REQUIRE(ctx, u && v && p && v0 && mask && coef_v && b_v, "km_assemble_v: NULL pointer");
    const int64_t N = (int64_t)(g->nx + 2) * (g->ny + 3);
    hipLaunchKernelGGL(k_assemble_v, dim3(ceil_div_i64(N, KM_BLOCK)), dim3(KM_BLOCK), 0, ctx->stream,
                       to_dev(g), u, v, p, v0, mask, coef_v, b_v);
    KM_CHECK_LAUNCH(ctx);
    return KM_OK;
}

int km_apply_ghosts(km_ctx* ctx, const km_grid* g, double* u, double* v) {
    if (!ctx) return KM_EINVAL;
    if (int rc = check_grid(ctx, g)) return rc;
    KM_REQUIRE(ctx, u && v, "km_apply_ghosts: NULL pointer");
    const int n = g->nx > g->ny ? g->nx : g->ny;
    hipLaunchKernelGGL(k_ghosts, dim3(ceil_div_i64(n, KM_BLOCK)), dim3(KM_BLOCK), 0, ctx->stream, to_dev(g), u, v);
    KM_CHECK_LAUNCH(ctx);
    return KM_OK;
}

int km_blend(km_ctx* ctx, const km_grid* g, double* u, double* v, const double* u_mid, const double* v_mid) {
    if (!ctx) return KM_EINVAL;
    if (int rc = check_grid(ctx, g)) return rc;
    KM_REQUIRE(ctx, u && v && u_mid && v_mid, "km_blend: NULL pointer");
    const int64_t n = (int64_t)(g->nx - 1) * g->ny + (int64_t)g->nx * (g->ny - 1);
    hipLaunchKernelGGL(k_blend, dim3(ceil_div_i64(n, KM_BLOCK)), dim3(KM_BLOCK), 0, ctx->stream,
                       to_dev(g), u, v, u_mid, v_mid);
    KM_CHECK_LAUNCH(ctx);
    return KM_OK;
}

int km_outflow_fix(km_ctx* ctx, const km_grid* g, double* u) {
    if (!ctx) return KM_EINVAL;
    if (int rc = check_grid(ctx, g)) return rc;
    KM_REQUIRE(ctx, u, "km_outflow_fix: NULL pointer");
    hipLaunchKernelGGL(k_outflow, dim3(1), dim3(KM_BLOCK), 0, ctx->stream, to_dev(g), u);
    KM_CHECK_LAUNCH(ctx);
    return KM_OK;
}

int km_assemble_p(km_ctx* ctx, const km_grid* g, const double* u, const double* v, const double* coef_u,
                  const double* coef_v, const uint8_t* mask, double* coef_p, double* b_p, double* mdiv) {
    if (!ctx) return KM_EINVAL;
    if (int rc = check_grid(ctx, g)) return rc;
    KM_REQUIRE(ctx, u && v && coef_u && coef_v && mask && coef_p && b_p && mdiv, "km_assemble_p: NULL pointer");
    const int64_t N = (int64_t)(g->nx + 2) * (g->ny + 2);
    hipLaunchKernelGGL(k_assemble_p, dim3(ceil_div_i64(N, KM_BLOCK)), dim3(KM_BLOCK), 0, ctx->stream,
                       to_dev(g), u, v, coef_u, coef_v, mask, coef_p, b_p, mdiv);
    KM_CHECK_LAUNCH(ctx);
    return KM_OK;
}

int km_correct(km_ctx* ctx, const km_grid* g, double* p, double* u, double* v, const double* pcor,
               const double* coef_u, const double* coef_v, const uint8_t* mask, int parts) {
    if (!ctx) return KM_EINVAL;
    if (int rc = check_grid(ctx, g)) return rc;
    KM_REQUIRE(ctx, p && u && v && pcor && coef_u && coef_v && mask, "km_correct: NULL pointer");
    KM_REQUIRE(ctx, parts >= 1 && parts <= 3, "km_correct: parts must be 1 (p), 2 (u, v) or 3");
    const int64_t n = (int64_t)g->nx * g->ny + (int64_t)(g->nx - 1) * g->ny + (int64_t)g->nx * (g->ny - 1);
    hipLaunchKernelGGL(k_correct, dim3(ceil_div_i64(n, KM_BLOCK)), dim3(KM_BLOCK), 0, ctx->stream,
                       to_dev(g), p, u, v, pcor, coef_u, coef_v, mask, parts);
    KM_CHECK_LAUNCH(ctx);
    return KM_OK;
}

int km_mdiv(km_ctx* ctx, const km_grid* g, const double* u, const double* v, double* mdiv, double* stats) {
    if (!ctx) return KM_EINVAL;
    if (int rc = check_grid(ctx, g)) return rc;
    KM_REQUIRE(ctx, u && v && mdiv, "km_mdiv: NULL pointer");
    const Flat fl = make_flat((int64_t)(g->nx + 2) * (g->ny + 2));
    if (int rc = km_ensure_workspace(ctx, 0)) return rc;
    double* P = ctx->partials + KM_PX;
    hipLaunchKernelGGL(k_mdiv, dim3(fl.G), dim3(KM_BLOCK), 0, ctx->stream, to_dev(g), fl, u, v, mdiv, P);
    KM_CHECK_LAUNCH(ctx);
    if (stats) {
        hipLaunchKernelGGL(k_mdiv_finish, dim3(1), dim3(KM_BLOCK), 0, ctx->stream, (const double*)P, fl.G, ctx->dres);
        KM_CHECK_LAUNCH(ctx);
        KM_HIP(ctx, hipMemcpyAsync(ctx->h_res, ctx->dres, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        KM_HIP_SYNC(ctx, hipStreamSynchronize(ctx->stream));
        stats[0] = ctx->h_res[0];
        stats[1] = ctx->h_res[1];
    }
    return KM_OK;
}

int km_time_forward(km_ctx* ctx, const km_grid* g, const double* u, const double* v, double* u0, double* v0) {
    if (!ctx) return KM_EINVAL;
    if (int rc = check_grid(ctx, g)) return rc;
    KM_REQUIRE(ctx, u && v && u0 && v0, "km_time_forward: NULL pointer");
    const int64_t Nu = (int64_t)(g->nx + 3) * (g->ny + 2), Nv = (int64_t)(g->nx + 2) * (g->ny + 3);
    hipLaunchKernelGGL(k_time_forward, dim3(ceil_div_i64(Nu + Nv, KM_BLOCK)), dim3(KM_BLOCK), 0, ctx->stream,
                       Nu, Nv, u, v, u0, v0);
    KM_CHECK_LAUNCH(ctx);
    return KM_OK;
}

}  // extern "C"
