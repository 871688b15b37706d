"""NumPy FP64 restatement of the implicit 5-point stencil solvers (oracle, test-only).

Vectors are padded 2-D arrays (NX, NY); the matrix is the SoA coefficient
array (5, NX, NY), k = 0 aP, 1 aW, 2 aE, 3 aN, 4 aS.  The stencil runs over
the whole padded range (cgsolver.py:13 ``N_ext = 0``); neighbours outside the
array read as zero (the reference reads out of bounds there and relies on
zero coefficients, cgsolver.py:59-60 -- parity unpinned for non-zero halos).

Status codes match the C-ABI (include/karman.h).
"""
from __future__ import annotations

import math

import numpy as np

OK = 0
BREAKDOWN = 1
MAXITER = 2
NONFINITE = 3
STALL = 4


def stencil(coef, x):
    """y = c0 x + c1 x_W + c2 x_E + c3 x_N + c4 x_S (bicgsolver.py:65-76,
    cgsolver.py:56-66), summed left to right as the reference writes it."""
    xp = np.zeros((x.shape[0] + 2, x.shape[1] + 2), dtype=x.dtype)
    xp[1:-1, 1:-1] = x
    return (coef[0] * x + coef[1] * xp[:-2, 1:-1] + coef[2] * xp[2:, 1:-1]
            + coef[3] * xp[1:-1, 2:] + coef[4] * xp[1:-1, :-2])


def dot(a, b):
    """<a, b> over the whole padded field (cgsolver.py:49-54)."""
    return float(np.dot(a.ravel(), b.ravel()))


def to_csr(coef):
    """The same operator as a scipy CSR matrix over the flattened padded field."""
    import scipy.sparse as sp
    _, NX, NY = coef.shape
    n = NX * NY
    idx = np.arange(n).reshape(NX, NY)
    rows, cols, vals = [idx.ravel()], [idx.ravel()], [coef[0].ravel()]
    for k, (di, dj) in zip((1, 2, 3, 4), ((-1, 0), (1, 0), (0, 1), (0, -1))):
        ii, jj = np.meshgrid(np.arange(NX), np.arange(NY), indexing='ij')
        ni, nj = ii + di, jj + dj
        ok = (ni >= 0) & (ni < NX) & (nj >= 0) & (nj < NY)
        rows.append(idx[ok])
        cols.append(idx[ni[ok], nj[ok]])
        vals.append(coef[k][ok])
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    return A


def direct_solve(coef, b):
    """Exact solution by sparse LU over the rows that are not identically zero
    (the padded halo rows are decoupled 0 = 0 equations, left at zero)."""
    import scipy.sparse.linalg as spla
    A = to_csr(coef)
    live = np.flatnonzero(np.abs(A).sum(axis=1).A.ravel() != 0)
    x = np.zeros(b.size)
    if live.size:
        sub = A[live][:, live].tocsc()
        x[live] = spla.spsolve(sub, b.ravel()[live])
    return x.reshape(b.shape)


class Info(dict):
    pass


def bicgstab(coef, b, eps=1e-8, x0=None, max_iter=None, precond=None, history=None, rho_form="direct"):
    """BiCGSTAB of bicgsolver.py:88-136 (van der Vorst, right-preconditioned hooks).

    x0 = 0 and r~ = r = b (bicgsolver.py:21-35) unless ``x0`` is given (then
    r = b - A x0).  Stop when sqrt(r.r / b.b) < eps with the recursive residual
    (bicgsolver.py:130, initial_rTr is r0.r0 which equals b.b for x0 = 0).
    ``precond`` is the M^-1 hook of update_phat/update_shat (bicgsolver.py:48-58;
    identity when None).  Guards: rho == 0 -> BREAKDOWN (bicgsolver.py:96-98),
    <r~, v> == 0 -> BREAKDOWN, <t, t> == 0 -> omega = 0 and, unless that
    converged, BREAKDOWN (the omega == 0 check of karman_taichi.py:451-454);
    non-finite residual -> NONFINITE.  max_iter defaults to NX*NY
    (cgsolver.py:14).

    ``rho_form``: "direct" forms rho = <r, r~> as bicgsolver.py:95 does;
    "split" forms it from the previous iteration's vectors as
    <s, r~> - omega <t, r~>, which equals <r, r~> in exact arithmetic (r = s -
    omega t) and is what a fused GPU iteration can reduce before r exists.
    """
    NX, NY = b.shape
    steps = NX * NY if max_iter is None else max_iter
    M = (lambda z: z) if precond is None else precond
    if x0 is None:
        x = np.zeros_like(b)
        r = b.copy()
    else:
        x = x0.copy()
        r = b - stencil(coef, x)
    r_tld = r.copy()
    p = np.zeros_like(b)
    v = np.zeros_like(b)
    omega = alpha = rho_1 = 1.0
    initial_rTr = dot(r, r)
    info = Info(iterations=0, status=OK, rel_resid=0.0, rho=[], alpha=[], omega=[])
    if initial_rTr == 0.0:
        info["rel_resid_true"] = 0.0
        return x, info
    rTr = initial_rTr
    status = MAXITER
    it = 0
    if rho_form not in ("direct", "split"):
        raise ValueError(rho_form)
    s_rt = t_rt = 0.0
    for i in range(steps):
        rho = dot(r, r_tld) if (rho_form == "direct" or i == 0) else s_rt - omega * t_rt
        if rho == 0.0:
            status = BREAKDOWN
            break
        if i == 0:
            p = r.copy()
        else:
            beta = (rho / rho_1) * (alpha / omega)
            p = r + beta * (p - omega * v)
        p_hat = M(p)
        v = stencil(coef, p_hat)
        rv = dot(r_tld, v)
        if rv == 0.0:
            status = BREAKDOWN
            break
        alpha = rho / rv
        s = r - alpha * v
        s_hat = M(s)
        t = stencil(coef, s_hat)
        tt = dot(t, t)
        omega = dot(t, s) / tt if tt != 0.0 else 0.0
        if rho_form == "split":
            s_rt, t_rt = dot(s, r_tld), dot(t, r_tld)
        x = x + (alpha * p_hat + omega * s_hat)
        r = s - omega * t
        rTr = dot(r, r)
        it = i + 1
        info["rho"].append(rho)
        info["alpha"].append(alpha)
        info["omega"].append(omega)
        if history is not None:
            history.append(math.sqrt(max(rTr, 0.0)))
        if not math.isfinite(rTr):
            status = NONFINITE
            break
        if math.sqrt(rTr / initial_rTr) < eps:
            status = OK
            break
        if omega == 0.0:
            status = BREAKDOWN
            break
        rho_1 = rho
    info["iterations"] = it
    info["status"] = status
    info["rel_resid"] = math.sqrt(rTr / initial_rTr) if math.isfinite(rTr) else float("nan")
    bb = dot(b, b)
    rt = b - stencil(coef, x)
    info["rel_resid_true"] = math.sqrt(dot(rt, rt) / bb) if bb > 0 else 0.0
    return x, info


def cg(coef, b, eps=1e-8, max_iter=None):
    """Conjugate gradient of cgsolver.py:83-114 (x0 = 0, p0 = r0 = b,
    stop when |r| < eps |r0|)."""
    NX, NY = b.shape
    steps = NX * NY if max_iter is None else max_iter
    x = np.zeros_like(b)
    r = b.copy()
    p = r.copy()
    initial_rTr = dot(r, r)
    old = initial_rTr
    info = Info(iterations=0, status=OK)
    if initial_rTr == 0.0:
        info["rel_resid"] = 0.0
        info["rel_resid_true"] = 0.0
        return x, info
    status = MAXITER
    it = 0
    new = old
    for i in range(steps):
        Ap = stencil(coef, p)
        pAp = dot(p, Ap)
        if pAp == 0.0:
            status = BREAKDOWN
            break
        alpha = old / pAp
        x = x + alpha * p
        r = r - alpha * Ap
        new = dot(r, r)
        it = i + 1
        if not math.isfinite(new):
            status = NONFINITE
            break
        if math.sqrt(new) < math.sqrt(initial_rTr) * eps:
            status = OK
            break
        beta = new / old
        p = r + beta * p
        old = new
    info["iterations"] = it
    info["status"] = status
    info["rel_resid"] = math.sqrt(new / initial_rTr) if math.isfinite(new) else float("nan")
    rt = b - stencil(coef, x)
    info["rel_resid_true"] = math.sqrt(dot(rt, rt) / initial_rTr)
    return x, info


def jacobi_precond(coef):
    """M^-1 = diag(A)^-1, the hook commented out at bicgsolver.py:51,57 and
    active at karman_taichi.py:399,419; zero diagonal rows map to zero."""
    d = coef[0]
    inv = np.where(d != 0.0, 1.0 / np.where(d != 0.0, d, 1.0), 0.0)
    return lambda z: inv * z


def fbicgstab(coef, b, eps=1e-8, x0=None, max_iter=None, precond=None, stall_window=200,
              stall_ratio=0.9, fallback=True):
    """Flexible right-preconditioned BiCGSTAB judged only by the true residual.

    The preconditioner may change between applications (a neural network).
    Convergence: ||b - A x|| / ||b|| < eps with the residual recomputed from x.
    Stall: when the best true residual has not dropped by ``stall_ratio`` over
    ``stall_window`` iterations (or the iteration broke down), restart with
    plain BiCGSTAB from the current x.  Returns (x, info) with info['fallback'].
    """
    NX, NY = b.shape
    steps = NX * NY if max_iter is None else max_iter
    bb = dot(b, b)
    info = Info(iterations=0, status=OK, fallback=False)
    if bb == 0.0:
        info["rel_resid_true"] = 0.0
        return np.zeros_like(b), info
    bn = math.sqrt(bb)
    x = np.zeros_like(b) if x0 is None else x0.copy()
    r = b - stencil(coef, x)
    rel = math.sqrt(dot(r, r)) / bn
    if rel < eps:
        info["rel_resid_true"] = rel
        return x, info
    r_tld = r.copy()
    p = np.zeros_like(b)
    v = np.zeros_like(b)
    omega = alpha = rho_1 = 1.0
    best, best_it = rel, 0
    it = 0
    status = MAXITER
    M = (lambda z: z) if precond is None else precond
    for i in range(steps):
        rho = dot(r, r_tld)
        if rho == 0.0 or not math.isfinite(rho):
            status = BREAKDOWN
            break
        if i == 0:
            p = r.copy()
        else:
            beta = (rho / rho_1) * (alpha / omega)
            p = r + beta * (p - omega * v)
        p_hat = M(p)
        v = stencil(coef, p_hat)
        rv = dot(r_tld, v)
        if rv == 0.0 or not math.isfinite(rv):
            status = BREAKDOWN
            break
        alpha = rho / rv
        s = r - alpha * v
        s_hat = M(s)
        t = stencil(coef, s_hat)
        tt = dot(t, t)
        omega = dot(t, s) / tt if tt != 0.0 else 0.0
        x_new = x + (alpha * p_hat + omega * s_hat)
        it = i + 1
        r_true = b - stencil(coef, x_new)
        rel = math.sqrt(dot(r_true, r_true)) / bn
        if not math.isfinite(rel):
            status = NONFINITE
            break
        x = x_new
        r = s - omega * t
        if rel < eps:
            status = OK
            break
        if rel < stall_ratio * best:
            best, best_it = rel, it
        elif it - best_it >= stall_window:
            status = STALL
            break
        if omega == 0.0:
            status = BREAKDOWN
            break
        rho_1 = rho
    info["iterations"] = it
    info["status"] = status
    if status != OK and fallback:
        info["fallback"] = True
        r0 = b - stencil(coef, x)
        d, sub = bicgstab(coef, r0, eps=eps * bn / max(math.sqrt(dot(r0, r0)), 1e-300),
                          max_iter=max_iter)
        x = x + d
        info["fallback_iterations"] = sub["iterations"]
        info["status"] = sub["status"]
    rt = b - stencil(coef, x)
    info["rel_resid_true"] = math.sqrt(dot(rt, rt)) / bn
    return x, info


# ---------------------------------------------------------------- two-level

def fgcr(coef, b, eps=1e-8, max_iter=None, precond=None, restart=16, history=None):
    """Flexible GCR (the generalised conjugate residual method with a
    preconditioner that may change from one application to the next, e.g. a
    neural network) judged by the residual relative to ||b||, x0 = 0.

    Iteration k: z = M(r), q = A z; q and z are orthogonalised against the
    stored q_j (classical Gram-Schmidt, twice) and scaled so that ||q|| = 1;
    alpha = <r, q>; x += alpha z; r -= alpha q.  Each step minimises ||r||
    over the stored directions, so ||r|| never increases.  After ``restart``
    directions the basis is cleared (x and r are kept).  On the recursive
    residual reaching eps the true residual b - A x replaces r and decides.
    With M = identity and no restart the residual norms are those of GMRES.
    """
    NX, NY = b.shape
    steps = NX * NY if max_iter is None else max_iter
    M = (lambda z: z) if precond is None else precond
    bb = dot(b, b)
    info = Info(iterations=0, status=OK)
    x = np.zeros_like(b)
    if bb == 0.0:
        info["rel_resid_true"] = 0.0
        return x, info
    bn = math.sqrt(bb)
    r = b.copy()
    Q, Z = [], []
    status = MAXITER
    it = 0
    while it < steps:
        z = M(r)
        q = stencil(coef, z)
        for _ in range(2):
            beta = [dot(q, qj) for qj in Q]
            for bj, qj, zj in zip(beta, Q, Z):
                q = q - bj * qj
                z = z - bj * zj
        nq = math.sqrt(dot(q, q))
        if not (nq > 0.0) or not math.isfinite(nq):
            status = BREAKDOWN if nq == 0.0 else NONFINITE
            break
        q, z = q / nq, z / nq
        alpha = dot(r, q)
        x = x + alpha * z
        r = r - alpha * q
        it += 1
        rel = math.sqrt(dot(r, r)) / bn
        if history is not None:
            history.append(rel)
        if rel < eps:
            r = b - stencil(coef, x)
            if math.sqrt(dot(r, r)) / bn < eps:
                status = OK
                break
        Q.append(q)
        Z.append(z)
        if len(Q) >= restart:
            Q, Z = [], []
    info["iterations"] = it
    info["status"] = status
    rt = b - stencil(coef, x)
    info["rel_resid_true"] = math.sqrt(dot(rt, rt)) / bn
    return x, info


def coarse_jacobi_precond(coef, B):
    """M(r) = z1 + D^-1 (r - A z1), z1 = P A_c^-1 P^T r: the coarse correction
    followed by one Jacobi step (the coarse correction alone maps every
    residual into the aggregates' piecewise-constant space, so it cannot
    precondition by itself)."""
    P, Ac = aggregation(coef, B)
    Aci = np.linalg.inv(Ac)
    jac = jacobi_precond(coef)

    def M(r):
        z1 = (P @ (Aci @ (P.T @ r.ravel()))).reshape(r.shape)
        return z1 + jac(r - stencil(coef, z1))
    return M


# ------------------------------------------------------------ multigrid V-cycle
# A geometric comparator in the spirit of benchmark/mgpcg.py:91-135 (V-cycle,
# red-black Gauss-Seidel smoothing, summed restriction, injection
# prolongation) for the variable-coefficient 5-point p' operator: each level
# aggregates 2 x 2 interior cells of the level above, the coarse operator is
# the Galerkin product P^T A P with P the indicator of the free cells, so
# it is again a 5-point operator.

def mg_coarsen(coef):
    """(coef_c, free_c): the Galerkin 5-point operator of 2 x 2 aggregates of
    coef's interior (padded (5, nxc + 2, nyc + 2)).  Free fine cells (a
    non-zero link) contribute aP, links inside the aggregate to aP, links
    across an aggregate edge to that side's coarse link.  Coarse cells with
    no free child are identity rows."""
    _, NX, NY = coef.shape
    nx, ny = NX - 2, NY - 2
    nxc, nyc = (nx + 1) // 2, (ny + 1) // 2
    c = np.zeros((5, nxc + 2, nyc + 2))
    free = free_cells(coef)
    anyfree = np.zeros((nxc + 2, nyc + 2), dtype=bool)
    for I in range(1, nxc + 1):
        for J in range(1, nyc + 1):
            ap = aw = ae = an = as_ = 0.0
            hit = False
            for di in (0, 1):
                for dj in (0, 1):
                    i, j = 2 * I - 1 + di, 2 * J - 1 + dj
                    if i > nx or j > ny or not free[i, j]:
                        continue
                    hit = True
                    ap += coef[0, i, j]
                    if di == 1:
                        ap += coef[1, i, j]
                    else:
                        aw += coef[1, i, j]
                    if di == 0 and i + 1 <= nx:
                        ap += coef[2, i, j]
                    else:
                        ae += coef[2, i, j]
                    if dj == 0 and j + 1 <= ny:
                        ap += coef[3, i, j]
                    else:
                        an += coef[3, i, j]
                    if dj == 1:
                        ap += coef[4, i, j]
                    else:
                        as_ += coef[4, i, j]
            if hit:
                c[:, I, J] = (ap, aw, ae, an, as_)
                anyfree[I, J] = True
            else:
                c[0, I, J] = 1.0
    return c, anyfree


def mg_hierarchy(coef, min_cells=400, max_levels=12):
    """Operators of every level, finest first: coarsen until the interior has
    at most ``min_cells`` cells or a side has fewer than 4 cells."""
    levels = [coef]
    while len(levels) < max_levels:
        _, NX, NY = levels[-1].shape
        nx, ny = NX - 2, NY - 2
        if nx * ny <= min_cells or nx < 4 or ny < 4:
            break
        levels.append(mg_coarsen(levels[-1])[0])
    return levels


def mg_restrict(coef_f, r):
    """Sum of the free fine cells' values into their 2 x 2 aggregates."""
    _, NX, NY = coef_f.shape
    nx, ny = NX - 2, NY - 2
    nxc, nyc = (nx + 1) // 2, (ny + 1) // 2
    free = free_cells(coef_f)
    out = np.zeros((nxc + 2, nyc + 2))
    for di in (0, 1):
        for dj in (0, 1):
            sub = (r * free)[1 + di:nx + 1:2, 1 + dj:ny + 1:2]
            out[1:1 + sub.shape[0], 1:1 + sub.shape[1]] += sub
    return out


def mg_prolong_add(coef_f, z, zc, kappa=1.0):
    """z += kappa zc on the free fine cells of each aggregate (injection)."""
    _, NX, NY = coef_f.shape
    nx, ny = NX - 2, NY - 2
    free = free_cells(coef_f)
    up = np.zeros_like(z)
    up[1:nx + 1, 1:ny + 1] = np.repeat(np.repeat(zc[1:-1, 1:-1], 2, axis=0), 2, axis=1)[:nx, :ny]
    return z + np.where(free, kappa * up, 0.0)


def mg_vcycle_precond(coef, pre=1, post=1, coarse_sweeps=50, omega=1.0, min_cells=400, kappa=1.5):
    """M(r): one V-cycle from z = 0 on the aggregation hierarchy of coef:
    ``pre`` red-black SOR sweeps (red first), the residual restricted, the
    next level's V-cycle, the correction prolonged with weight ``kappa``,
    ``post`` sweeps (black first, as mgpcg.py:125-127 reverses the colours);
    ``coarse_sweeps`` sweeps on the coarsest level.  Piecewise-constant
    prolongation makes the Galerkin coarse correction too small by about
    half on smooth errors; kappa = 1.5 over-weights it (measured per-cycle
    contraction on a Poisson problem: 0.67 / 0.71 at 64^2 / 128^2 cells with
    kappa = 1, 0.45 / 0.52 with 1.5)."""
    levels = mg_hierarchy(coef, min_cells)

    def sweep(c, r, z, first):
        d = c[0]
        dinv = np.where(d != 0.0, 1.0 / np.where(d != 0.0, d, 1.0), 0.0)
        red = red_cells(r.shape)
        for colour in ((red, ~red) if first == 0 else (~red, red)):
            z = np.where(colour, z + omega * dinv * (r - stencil(c, z)), z)
        return z

    def vcycle(l, r):
        c = levels[l]
        z = np.zeros_like(r)
        if l == len(levels) - 1:
            for _ in range(coarse_sweeps):
                z = sweep(c, r, z, 0)
            return z
        for _ in range(pre):
            z = sweep(c, r, z, 0)
        rc = mg_restrict(c, r - stencil(c, z))
        z = mg_prolong_add(c, z, vcycle(l + 1, rc), kappa)
        for _ in range(post):
            z = sweep(c, r, z, 1)
        return z

    return lambda r: vcycle(0, r), levels


def coarse_block(nx, ny, req=0):
    """Aggregate edge length: req, else the smallest of 8, 12, 16, 24, ..., 256
    giving at most 2048 coarse unknowns, or on larger grids at most as many as
    4 m^2 <= 21 nx ny allows (at most 8192; the fused solver's choice)."""
    if req > 0:
        return req
    for B in (8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256):
        mJ = -(-ny // B)
        m = (-(-nx // B)) * mJ
        if mJ > 63:   # the fused solver's band limit
            continue
        if m <= 2048 or (m <= 8192 and 4 * m * m <= 21 * nx * ny):
            return B
    return 512


def free_cells(coef):
    """Cells with a non-zero link (identity rows and the halo are not free)."""
    return (coef[1] != 0) | (coef[2] != 0) | (coef[3] != 0) | (coef[4] != 0)


def aggregation(coef, B):
    """P (n x m, indicator of the free cells of each B x B interior aggregate,
    i-major aggregate order) and the Galerkin operator A_c = P^T A P (dense)."""
    import scipy.sparse as sp
    _, NX, NY = coef.shape
    nx, ny = NX - 2, NY - 2
    mI, mJ = -(-nx // B), -(-ny // B)
    ii, jj = np.meshgrid(np.arange(NX), np.arange(NY), indexing="ij")
    agg = (np.clip(ii - 1, 0, None) // B) * mJ + np.clip(jj - 1, 0, None) // B
    rows = np.flatnonzero(free_cells(coef).ravel())
    P = sp.csr_matrix((np.ones(rows.size), (rows, agg.ravel()[rows])), shape=(NX * NY, mI * mJ))
    A = to_csr(coef)
    Ac = (P.T @ A @ P).toarray()
    empty = np.asarray(P.sum(axis=0)).ravel() == 0
    Ac[empty, empty] = 1.0          # aggregates without free cells: identity rows
    return P, Ac


def coarse_correct(coef, r, B):
    """z1 = P A_c^-1 P^T r and r1 = r - A z1."""
    P, Ac = aggregation(coef, B)
    z1 = (P @ np.linalg.solve(Ac, P.T @ r.ravel())).reshape(r.shape)
    return z1, r - stencil(coef, z1)


def red_cells(shape):
    """Red cells of the red-black ordering: (i + j) even."""
    ii, jj = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), indexing="ij")
    return (ii + jj) % 2 == 0


def rbsor_sweep(coef, r, z, omega):
    """One red-black SOR sweep on A z = r: the red cells, then the black cells,
    each z + omega (r - A z) / aP with the latest values of the other colour;
    rows with aP = 0 are left unchanged."""
    d = coef[0]
    dinv = np.where(d != 0.0, 1.0 / np.where(d != 0.0, d, 1.0), 0.0)
    red = red_cells(r.shape)
    for colour in (red, ~red):
        z = np.where(colour, z + omega * dinv * (r - stencil(coef, z)), z)
    return z


def twolevel_precond(coef, B, sweeps=2, omega=1.3):
    """M(r) = z1 + red-black SOR sweeps on A z = r from z1 = P A_c^-1 P^T r:
    the preconditioner of the flexible solve with the coarse level and no
    network."""
    P, Ac = aggregation(coef, B)
    Aci = np.linalg.inv(Ac)

    def M(r):
        z = (P @ (Aci @ (P.T @ r.ravel()))).reshape(r.shape)
        for _ in range(sweeps):
            z = rbsor_sweep(coef, r, z, omega)
        return z
    return M
