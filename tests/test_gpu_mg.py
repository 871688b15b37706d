"""The multigrid V-cycle (km_mg_vcycle, the KM_FLAG_MG preconditioner) through
the C-ABI against the oracle's restatement (oracle.krylov.mg_vcycle_precond,
itself pinned on the CPU: Galerkin coarse operators against scipy's P^T A P,
mesh-independent contraction on a Poisson problem).

Bars: one V-cycle M(r) equal to the oracle's to 1e-14 relative (the kernels
sum each expression in the oracle's order with contraction off, so only the
rounding of the divisions could differ, and both are IEEE) with the same
hierarchy depth; preconditioned flexible GCR reaches the true residual with
an iteration count within two of the oracle's, the solution within
kappa * eps of the LU solution (kappa ~ 3e5 on the C1 p' system); flexible
BiCGSTAB with the V-cycle reaches the true residual in far fewer iterations
than plain BiCGSTAB.
"""
import ctypes

import numpy as np
import pytest
import torch

from karman_taichi_amd import _lib, cases
from karman_taichi_amd._lib import KmSolveInfo, ptr
from oracle import fvm, krylov

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1_pcor():
    c = cases.c1()
    m = cases.make_mask(c)
    st = fvm.State(c, *cases.initial_fields(c, m), m)
    fvm.simple_step(st, solver="direct")
    coef, b, _ = fvm.assemble_p(c, st.u, st.v, st.coef_u[0], st.coef_v[0], m)
    return coef, b, krylov.direct_solve(coef, b)


def poisson(nx, ny):
    """Cell-centred 5-point Laplacian, Dirichlet walls (links to the halo dropped)."""
    c = np.zeros((5, nx + 2, ny + 2))
    c[0, 1:-1, 1:-1] = 4.0
    c[1, 2:-1, 1:-1] = -1.0
    c[2, 1:-2, 1:-1] = -1.0
    c[3, 1:-1, 1:-2] = -1.0
    c[4, 1:-1, 2:-1] = -1.0
    return c


def random_system(NX, NY, seed):
    """Random diagonally dominant 5-point system with a zero halo and a block
    of identity rows (an obstacle) inside."""
    rng = np.random.default_rng(seed)
    c = np.zeros((5, NX, NY))
    c[1:, 1:-1, 1:-1] = -rng.random((4, NX - 2, NY - 2))
    c[1, 1, :] = 0
    c[2, -2, :] = 0
    c[4, :, 1] = 0
    c[3, :, -2] = 0
    i0, j0 = NX // 3, NY // 3
    c[1:, i0:i0 + 3, j0:j0 + 4] = 0
    c[2, i0 - 1, j0:j0 + 4] = 0
    c[1, i0 + 3, j0:j0 + 4] = 0
    c[3, i0:i0 + 3, j0 - 1] = 0
    c[4, i0:i0 + 3, j0 + 4] = 0
    c[0, 1:-1, 1:-1] = -(c[1] + c[2] + c[3] + c[4])[1:-1, 1:-1] + 0.05
    c[0, i0:i0 + 3, j0:j0 + 4] = 1.0
    r = np.zeros((NX, NY))
    r[1:-1, 1:-1] = rng.standard_normal((NX - 2, NY - 2))
    r[i0:i0 + 3, j0:j0 + 4] = 0.0
    return c, r


def vcycle_gpu(dev, coef, r):
    ctx = _lib.context_for(dev)
    ct = torch.from_numpy(coef).to(dev)
    rt = torch.from_numpy(r).to(dev)
    zt = torch.empty_like(rt)
    lv = ctypes.c_int(0)
    ctx.call("km_mg_vcycle", r.shape[0], r.shape[1], ptr(ct), ptr(rt), ptr(zt), ctypes.byref(lv))
    return zt.cpu().numpy(), lv.value


def solve(dev, coef, b, flags, eps=1e-10):
    ctx = _lib.context_for(dev)
    ct = torch.from_numpy(coef).to(dev)
    bt = torch.from_numpy(b).to(dev)
    xt = torch.zeros_like(bt)
    info = KmSolveInfo()
    ctx.call("km_fbicgstab", b.shape[0], b.shape[1], ptr(ct), ptr(bt), ptr(xt), eps, 0, flags, None,
             ctypes.byref(info))
    return xt.cpu().numpy(), info


def resid(coef, x, b):
    return float(np.linalg.norm(b - krylov.stencil(coef, x)) / np.linalg.norm(b))


@pytest.mark.parametrize("system", ["c1_pcor", "poisson_96x40", "random_37x23"])
def test_vcycle_matches_oracle(system, c1_pcor, dev):
    if system == "c1_pcor":
        coef, r = c1_pcor[0], c1_pcor[1]
    elif system == "poisson_96x40":
        coef = poisson(96, 40)
        r = np.zeros((98, 42))
        r[1:-1, 1:-1] = np.random.default_rng(3).standard_normal((96, 40))
    else:
        coef, r = random_system(37, 23, seed=5)
    M, levels = krylov.mg_vcycle_precond(coef)
    zo = M(r)
    zg, depth = vcycle_gpu(dev, coef, r)
    err = float(np.abs(zg - zo).max() / np.abs(zo).max())
    print(f"V-cycle {system}: {depth} levels (oracle {len(levels)}), max |z_gpu - z_oracle| / max |z| = {err:.1e}")
    assert depth == len(levels)
    assert err <= 1e-14


def test_vcycle_twice_on_one_context(c1_pcor, dev):
    """A second system of another shape rebuilds the hierarchy; the first
    shape again gives the first result."""
    coef, r = c1_pcor[0], c1_pcor[1]
    z1, _ = vcycle_gpu(dev, coef, r)
    c2, r2 = random_system(37, 23, seed=5)
    vcycle_gpu(dev, c2, r2)
    z3, _ = vcycle_gpu(dev, coef, r)
    assert np.array_equal(z1, z3)


def test_mg_gcr_matches_oracle(c1_pcor, dev):
    coef, b, xd = c1_pcor
    M, _ = krylov.mg_vcycle_precond(coef)
    _, io = krylov.fgcr(coef, b, eps=1e-10, precond=M, restart=16)
    x, info = solve(dev, coef, b, _lib.KM_FLAG_GCR | _lib.KM_FLAG_MG)
    true = resid(coef, x, b)
    print(f"GCR + V-cycle C1 p': gpu {info.iterations} it, oracle {io['iterations']} it, true {true:.2e}")
    assert info.status == _lib.KM_OK and info.fallback == 0 and true < 1e-10
    assert abs(info.iterations - io["iterations"]) <= 2
    assert np.linalg.norm(x - xd) <= 3e-5 * np.linalg.norm(xd)


def test_mg_fbicgstab(c1_pcor, dev):
    coef, b, xd = c1_pcor
    x, info = solve(dev, coef, b, _lib.KM_FLAG_MG)
    _, plain = solve(dev, coef, b, 0)
    true = resid(coef, x, b)
    print(f"flexible BiCGSTAB + V-cycle C1 p': {info.iterations} it (plain {plain.iterations}), true {true:.2e}")
    assert info.status == _lib.KM_OK and info.fallback == 0 and true < 1e-10
    assert info.iterations * 5 < plain.iterations
    assert np.linalg.norm(x - xd) <= 3e-5 * np.linalg.norm(xd)


def test_mg_c2_pcor(dev):
    """The C2 p' system (1280 x 256 interior): both outer methods with the
    V-cycle reach 1e-8 without the plain fallback."""
    from karman_taichi_amd.simple import SIMPLESolver
    s = SIMPLESolver.from_case(cases.c2(), device=dev)
    s.init()
    s.solve_momentum_eqn(1)
    s.correct_conserv()
    s.compute_coef_p()
    coef, b = s._coef_p.cpu().numpy(), s.b_p.cpu().numpy()
    out = {}
    for name, flags in (("gcr", _lib.KM_FLAG_GCR | _lib.KM_FLAG_MG), ("fbicgstab", _lib.KM_FLAG_MG)):
        x, info = solve(dev, coef, b, flags, eps=1e-8)
        true = resid(coef, x, b)
        out[name] = (info.iterations, true)
        assert info.status == _lib.KM_OK and info.fallback == 0 and true < 1e-8, (name, info.iterations)
    print("C2 p' with the V-cycle (iterations, true residual):", out)


def test_mg_rejects_other_preconditioners(c1_pcor, dev):
    coef, b, _ = c1_pcor
    for flags in (_lib.KM_FLAG_MG | _lib.KM_FLAG_MIXED, _lib.KM_FLAG_MG | _lib.KM_FLAG_COARSE,
                  _lib.KM_FLAG_MG | _lib.KM_FLAG_GCR | _lib.KM_FLAG_MIXED):
        with pytest.raises(_lib.KarmanError, match="KM_FLAG_MG"):
            solve(dev, coef, b, flags)
