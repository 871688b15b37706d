"""Pins of the discretisation oracle (oracle/fvm.py) without Taichi.

* uniform-flow closed form of the momentum links (Appendix A special case);
* an element-by-element loop restatement of simple.py:77-110 / :143-163 on a
  tiny grid (catches slice-offset mistakes of the vectorised oracle);
* p' matrix invariants: symmetry, zero row sums, M-matrix signs, SPD;
* continuity after a corrected step is exact on every interior fluid cell
  (a wrong sign or index in assemble_p / correct breaks it);
* the telescoping global mass balance and the outlet flux fix;
* plane Poiseuille flow (docs/notes.md:33) and Ghia et al. 1982 Re=100
  (data/ghia1982.dat, copied under tests/golden/).
"""
import os

import numpy as np
import pytest

from karman_taichi_amd import cases
from oracle import fvm, krylov

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def small_channel(nx=24, ny=10, d=2, **kw):
    c = cases.Case(name="small", nx=nx, ny=ny, lx=nx / ny, ly=1.0, obstacle_cells=d,
                   obstacle_xc=0.3 * nx / ny, mu=0.02, dt=0.05, **kw)
    return c


def test_uniform_flow_closed_form():
    c = cases.Case(name="uni", nx=12, ny=8, lx=3.0, ly=1.0, mu=0.013, dt=0.02, rho=1.3)
    m = cases.make_mask(c)
    U = 0.8
    u = np.full(c.shape_u, U)
    v = np.zeros(c.shape_v)
    p = np.zeros(c.shape_p)
    cu, _ = fvm.assemble_u(c, u, v, p, u, m)
    cv, _ = fvm.assemble_v(c, u, v, p, v, m)
    Dx, Dy = c.mu * c.dy / c.dx, c.mu * c.dx / c.dy
    F = c.rho * U * c.dy
    a0 = c.rho * c.dx * c.dy / c.dt
    I, J = slice(3, c.nx - 1), slice(2, c.ny)          # faces clear of every boundary row
    for coef in (cu[:, I, J], cv[:, 2:c.nx, 3:c.ny - 1]):
        np.testing.assert_allclose(coef[1], -(Dx + F / 2), rtol=1e-13)
        np.testing.assert_allclose(coef[2], -(Dx - F / 2), rtol=1e-13)
        np.testing.assert_allclose(coef[3], -Dy, rtol=1e-13)
        np.testing.assert_allclose(coef[4], -Dy, rtol=1e-13)
        np.testing.assert_allclose(coef[0], 2 * Dx + 2 * Dy + a0, rtol=1e-13)


def _loop_coef_u(c, u, v, p, u0):
    """simple.py:80-92 plus set_bc (:177-207) written per element, index by index."""
    nx, ny, dx, dy, dt, rho, mu, bc = c.nx, c.ny, c.dx, c.dy, c.dt, c.rho, c.mu, c.bc
    co = np.zeros((5,) + c.shape_u); b = np.zeros(c.shape_u)
    for i in range(2, nx + 1):
        for j in range(1, ny + 1):
            co[1, i, j] = -(mu * dy / dx + 0.5 * rho * 0.5 * (u[i, j] + u[i - 1, j]) * dy)
            co[2, i, j] = -(mu * dy / dx - 0.5 * rho * 0.5 * (u[i, j] + u[i + 1, j]) * dy)
            co[3, i, j] = -(mu * dx / dy - 0.5 * rho * 0.5 * (v[i - 1, j + 1] + v[i, j + 1]) * dx)
            co[4, i, j] = -(mu * dx / dy + 0.5 * rho * 0.5 * (v[i - 1, j] + v[i, j]) * dx)
            co[0, i, j] = (-(co[1, i, j] + co[2, i, j] + co[3, i, j] + co[4, i, j])
                           + rho * 0.5 * (u[i, j] + u[i + 1, j]) * dy - rho * 0.5 * (u[i, j] + u[i - 1, j]) * dy
                           + rho * 0.5 * (v[i - 1, j + 1] + v[i, j + 1]) * dx
                           - rho * 0.5 * (v[i - 1, j] + v[i, j]) * dx + rho * dx * dy / dt)
            b[i, j] = (p[i - 1, j] - p[i, j]) * dy + rho * dx * dy / dt * u0[i, j]
    for j in range(1, ny + 1):
        b[2, j] += -co[1, 2, j] * bc['w'][0]; co[1, 2, j] = 0.0
        b[nx, j] += -co[2, nx, j] * bc['e'][0]; co[2, nx, j] = 0.0
    for i in range(2, nx + 1):
        b[i, 1] += 2 * mu * bc['s'][1] * dx / dy
        co[0, i, 1] += (co[4, i, 1] + 2 * mu * dx / dy); co[4, i, 1] = 0.0
        b[i, ny] += 2 * mu * bc['n'][1] * dx / dy
        co[0, i, ny] += (co[3, i, ny] + 2 * mu * dx / dy); co[3, i, ny] = 0.0
    return co, b


def _loop_coef_v(c, u, v, p, v0):
    nx, ny, dx, dy, dt, rho, mu, bc = c.nx, c.ny, c.dx, c.dy, c.dt, c.rho, c.mu, c.bc
    co = np.zeros((5,) + c.shape_v); b = np.zeros(c.shape_v)
    for i in range(1, nx + 1):
        for j in range(2, ny + 1):
            co[1, i, j] = -(mu * dy / dx + 0.5 * rho * 0.5 * (u[i, j] + u[i, j - 1]) * dy)
            co[2, i, j] = -(mu * dy / dx - 0.5 * rho * 0.5 * (u[i + 1, j - 1] + u[i + 1, j]) * dy)
            co[3, i, j] = -(mu * dx / dy - 0.5 * rho * 0.5 * (v[i, j + 1] + v[i, j]) * dx)
            co[4, i, j] = -(mu * dx / dy + 0.5 * rho * 0.5 * (v[i, j - 1] + v[i, j]) * dx)
            co[0, i, j] = (-(co[1, i, j] + co[2, i, j] + co[3, i, j] + co[4, i, j])
                           + rho * 0.5 * (u[i + 1, j - 1] + u[i + 1, j]) * dy - rho * 0.5 * (u[i, j] + u[i, j - 1]) * dy
                           + rho * 0.5 * (v[i, j + 1] + v[i, j]) * dx - rho * 0.5 * (v[i, j - 1] + v[i, j]) * dx
                           + rho * dx * dy / dt)
            b[i, j] = (p[i, j - 1] - p[i, j]) * dx + rho * dx * dy / dt * v0[i, j]
    for i in range(1, nx + 1):
        b[i, 2] += -co[4, i, 2] * bc['s'][0]; co[4, i, 2] = 0.0
        b[i, ny] += -co[3, i, ny] * bc['n'][0]; co[3, i, ny] = 0.0
    for j in range(2, ny + 1):
        b[1, j] += 2 * mu * bc['w'][1] * dy / dx
        co[0, 1, j] += (co[1, 1, j] + 2 * mu * dy / dx); co[1, 1, j] = 0.0
        b[nx, j] += 2 * mu * bc['e'][1] * dy / dx
        co[0, nx, j] += (co[2, nx, j] + 2 * mu * dy / dx); co[2, nx, j] = 0.0   # bicgsimp.py:196
    return co, b


def test_loop_restatement_wall_bcs():
    c = cases.Case(name="w", nx=6, ny=5, lx=1.2, ly=1.0, mu=0.03, dt=0.1, east=cases.EAST_WALL,
                   bc={'w': [0.7, 0.2], 'e': [0.4, -0.3], 'n': [0.1, 1.0], 's': [-0.2, 0.5]})
    rng = np.random.default_rng(11)
    u, v, p = rng.standard_normal(c.shape_u), rng.standard_normal(c.shape_v), rng.standard_normal(c.shape_p)
    u0, v0 = rng.standard_normal(c.shape_u), rng.standard_normal(c.shape_v)
    m = cases.make_mask(c)
    cu, bu = fvm.assemble_u(c, u, v, p, u0, m)
    lu, lbu = _loop_coef_u(c, u, v, p, u0)
    np.testing.assert_allclose(cu, lu, rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(bu, lbu, rtol=1e-14, atol=1e-15)
    cv, bv = fvm.assemble_v(c, u, v, p, v0, m)
    lv, lbv = _loop_coef_v(c, u, v, p, v0)
    np.testing.assert_allclose(cv, lv, rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(bv, lbv, rtol=1e-14, atol=1e-15)
    # the simple.py:215 quirk differs exactly on the east v column
    cq, _ = fvm.assemble_v(c, u, v, p, v0, m, east_wall_quirk=True)
    assert not np.allclose(cq[0, c.nx, 2:c.ny + 1], cv[0, c.nx, 2:c.ny + 1])
    np.testing.assert_array_equal(cq[0, :c.nx], cv[0, :c.nx])


def test_loop_restatement_pcor():
    c = cases.Case(name="w", nx=7, ny=5, lx=1.4, ly=1.0, east=cases.EAST_WALL)
    rng = np.random.default_rng(3)
    u, v = rng.standard_normal(c.shape_u), rng.standard_normal(c.shape_v)
    apu = 1.0 + rng.random(c.shape_u)
    apv = 1.0 + rng.random(c.shape_v)
    m = cases.make_mask(c)
    cp, bp, md = fvm.assemble_p(c, u, v, apu, apv, m)
    nx, ny, dx, dy, rho = c.nx, c.ny, c.dx, c.dy, c.rho
    for i in range(1, nx + 1):
        for j in range(1, ny + 1):
            mdiv = rho * (u[i, j] - u[i + 1, j]) * dy + rho * (v[i, j] - v[i, j + 1]) * dx
            aw = 0.0 if i == 1 else -rho * dy * dy / apu[i, j]
            ae = 0.0 if i == nx else -rho * dy * dy / apu[i + 1, j]
            an = 0.0 if j == ny else -rho * dx * dx / apv[i, j + 1]
            as_ = 0.0 if j == 1 else -rho * dx * dx / apv[i, j]
            assert md[i, j] == mdiv and bp[i, j] == mdiv
            assert (cp[1, i, j], cp[2, i, j], cp[3, i, j], cp[4, i, j]) == (aw, ae, an, as_)
            assert cp[0, i, j] == -(aw + ae + an + as_)


def test_mask_rows_and_wall_treatment():
    c = small_channel()
    m = cases.make_mask(c)
    i0, i1, j0, j1 = c.obstacle_box()
    rng = np.random.default_rng(5)
    u, v, p = rng.random(c.shape_u), 0.1 * rng.standard_normal(c.shape_v), rng.random(c.shape_p)
    cu, bu = fvm.assemble_u(c, u, v, p, u, m)
    nomask, _ = fvm.assemble_u(c, u, v, p, u, np.zeros_like(m))
    # faces touching the obstacle are identity rows
    for i in range(i0, i1 + 2):
        for j in range(j0, j1 + 1):
            assert cu[0, i, j] == 1.0 and not cu[1:, i, j].any() and bu[i, j] == 0.0
    # faces just above the obstacle: south link replaced by wall shear 2 mu dx/dy
    for i in range(i0, i1 + 2):
        j = j1 + 1
        assert cu[4, i, j] == 0.0
        want = nomask[0, i, j] + (nomask[4, i, j] + 2 * c.mu * c.dx / c.dy)
        assert cu[0, i, j] == pytest.approx(want, rel=1e-14)
    # faces upstream / downstream along the stream-wise line drop the link to the solid face
    assert cu[2, i0 - 1, j0] == 0.0 and cu[1, i1 + 2, j0] == 0.0
    assert cu[0, i0 - 1, j0] == nomask[0, i0 - 1, j0]


def test_pcor_matrix_invariants():
    c = small_channel()
    m = cases.make_mask(c)
    st = fvm.State(c, *cases.initial_fields(c, m), m)
    fvm.simple_step(st, solver="direct")
    coef, b, _ = fvm.assemble_p(c, st.u, st.v, st.coef_u[0], st.coef_v[0], m)
    A = krylov.to_csr(coef)
    assert abs(A - A.T).max() == 0.0
    # in plane form too: aE(i,j) == aW(i+1,j), aN(i,j) == aS(i,j+1) bit for
    # bit, zero past the last row and column
    aw, ae, an, as_ = coef[1], coef[2], coef[3], coef[4]
    assert np.array_equal(ae[:-1], aw[1:]) and not ae[-1].any()
    assert np.array_equal(an[:, :-1], as_[:, 1:]) and not an[:, -1].any()
    off = coef[1:].copy()
    assert (off <= 0).all()
    live = coef[0] != 0
    assert (coef[0][live] > 0).all()
    dirc = fvm.dirichlet_cells(c, m)
    near = np.zeros_like(dirc)
    near[1:-1, 1:-1] = dirc[:-2, 1:-1] | dirc[2:, 1:-1] | dirc[1:-1, :-2] | dirc[1:-1, 2:]
    rows = live & ~dirc & ~near
    rs = coef.sum(axis=0)
    assert np.abs(rs[rows]).max() <= 1e-15 * np.abs(coef[0]).max()
    Ad = A.toarray()
    keep = np.flatnonzero(np.abs(Ad).sum(axis=1) != 0)
    ev = np.linalg.eigvalsh(Ad[np.ix_(keep, keep)])
    assert ev.min() > 0


def test_continuity_exact_after_correction_and_mass_balance():
    c = small_channel()
    m = cases.make_mask(c)
    st = fvm.State(c, *cases.initial_fields(c, m), m)
    for _ in range(3):
        fvm.simple_step(st, solver="direct")
        md = st.mdiv
        dirc = fvm.dirichlet_cells(c, m)
        fluid = np.zeros_like(dirc)
        fluid[1:c.nx + 1, 1:c.ny + 1] = True
        fluid &= ~dirc
        scale = np.abs(st.b_p).max()
        assert np.abs(md[fluid]).max() <= 1e-10 * scale
        # telescoping: sum of cell imbalances = inflow - outflow through the boundary
        inflow = c.rho * c.dy * (st.u[1, 1:c.ny + 1].sum() - st.u[c.nx + 1, 1:c.ny + 1].sum())
        inflow += c.rho * c.dx * (st.v[1:c.nx + 1, 1].sum() - st.v[1:c.nx + 1, c.ny + 1].sum())
        assert md[1:c.nx + 1, 1:c.ny + 1].sum() == pytest.approx(inflow, abs=1e-12)


def test_outflow_fix_balances_flux():
    c = small_channel()
    rng = np.random.default_rng(9)
    u = rng.random(c.shape_u)
    fvm.outflow_fix(c, u)
    assert u[c.nx + 1, 1:c.ny + 1].sum() == pytest.approx(u[1, 1:c.ny + 1].sum(), rel=1e-14)
    d = u[c.nx + 1, 1:c.ny + 1] - u[c.nx, 1:c.ny + 1]
    assert np.ptp(d) < 1e-14    # a uniform shift of the zero-gradient column


def run_steady(c, n_iter, tol=1e-9):
    """Pseudo-transient march to steady state: one SIMPLE step per time step.
    At steady state u = u0, so the transient term drops out."""
    m = cases.make_mask(c)
    st = fvm.State(c, *cases.initial_fields(c, m), m)
    for k in range(n_iter):
        u_old = st.u.copy()
        fvm.simple_step(st, solver="direct")
        fvm.time_forward(st)
        if np.abs(st.u - u_old).max() < tol:
            break
    return st, k


def test_poiseuille_developed_profile():
    """Uniform inlet U in a long channel: downstream u_max -> 1.5 U and
    dp/dx -> -12 mu U / H^2 (plane Poiseuille, the author's check in docs/notes.md:33)."""
    c = cases.Case(name="pois", nx=64, ny=12, lx=8.0, ly=1.0, mu=0.1, dt=0.2, v_perturb=0.0,
                   alpha_m=0.7, alpha_p=0.3)
    st, it = run_steady(c, 400)
    assert it < 399
    i = int(0.75 * c.nx)
    prof = st.u[i, 1:c.ny + 1]
    y = (np.arange(c.ny) + 0.5) * c.dy
    exact = 6.0 * y * (1 - y)
    assert np.abs(prof - exact).max() < 0.03 * 1.5
    dpdx = (st.p[i + 1, 1:c.ny + 1] - st.p[i - 1, 1:c.ny + 1]).mean() / (2 * c.dx)
    assert dpdx == pytest.approx(-12 * c.mu * 1.0, rel=0.05)


def test_ghia_re100_cavity():
    """Lid-driven cavity Re = 100 (simple.py:287-303 setup with the bicgsimp.py
    pin) against the Ghia 1982 centre-line u(y) (data/ghia1982.dat column 2)."""
    c = cases.cavity(32, 100.0)
    c.alpha_m, c.alpha_p, c.dt = 0.7, 0.3, 0.1
    st, it = run_steady(c, 1500, tol=1e-7)
    assert it < 1499
    y_ref, u_ref = np.loadtxt(os.path.join(GOLDEN, "ghia1982.dat"), unpack=True, skiprows=2, usecols=(0, 1))
    # u faces at x = 0.5 are i = nx/2 + 1; cell centres in y
    uc = st.u[c.nx // 2 + 1, 1:c.ny + 1]
    y = (np.arange(c.ny) + 0.5) * c.dy
    u_interp = np.interp(y_ref[1:-1], y, uc)
    err = np.abs(u_interp - u_ref[1:-1]).max()
    assert err < 0.02, err
