"""Flexible GCR (KM_FLAG_GCR) through the C-ABI against the oracle's
restatement (oracle.krylov.fgcr) on the C1 p' system.

Bars: status OK with the true residual below eps, no fallback; the solution
within kappa * eps of the LU solution (kappa ~ 3e5); with a preconditioner
the oracle applies identically (the two-level M(r) agrees to ~4e-14, the
coarse correction + Jacobi step to rounding), iteration counts within two of
the oracle's (GCR minimises ||r|| at every step, so rounding moves its count
far less than BiCGSTAB's); FP32 directions (KM_FLAG_MIXED) reach the same
FP64 true residual; an untrained network (a non-linear preconditioner that
breaks flexible BiCGSTAB) converges.
"""
import ctypes

import numpy as np
import pytest
import torch

from karman_taichi_amd import _lib, cases
from karman_taichi_amd._lib import KmSolveInfo, ptr
from oracle import fvm, krylov, unet

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1_pcor():
    c = cases.c1()
    m = cases.make_mask(c)
    st = fvm.State(c, *cases.initial_fields(c, m), m)
    fvm.simple_step(st, solver="direct")
    coef, b, _ = fvm.assemble_p(c, st.u, st.v, st.coef_u[0], st.coef_v[0], m)
    return coef, b, krylov.direct_solve(coef, b)


def gcr(dev, coef, b, flags, params=None, eps=1e-10):
    ctx = _lib.context_for(dev)
    ct = torch.from_numpy(coef).to(dev)
    bt = torch.from_numpy(b).to(dev)
    xt = torch.zeros_like(bt)
    pt = None if params is None else torch.from_numpy(np.asarray(params, np.float32)).to(dev)
    info = KmSolveInfo()
    ctx.call("km_fbicgstab", b.shape[0], b.shape[1], ptr(ct), ptr(bt), ptr(xt), eps, 0,
             flags | _lib.KM_FLAG_GCR, None if pt is None else ptr(pt), ctypes.byref(info))
    return xt.cpu().numpy(), info


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def resid(coef, x, b):
    return float(np.linalg.norm(b - krylov.stencil(coef, x)) / np.linalg.norm(b))


@pytest.mark.parametrize("variant", ["twolevel", "coarse_jacobi"])
def test_fgcr_matches_oracle(variant, c1_pcor, dev):
    coef, b, xd = c1_pcor
    B = krylov.coarse_block(b.shape[0] - 2, b.shape[1] - 2)
    if variant == "twolevel":
        flags, M = _lib.KM_FLAG_COARSE, krylov.twolevel_precond(coef, B)
    else:
        flags, M = _lib.KM_FLAG_COARSE | _lib.KM_FLAG_JACOBI, krylov.coarse_jacobi_precond(coef, B)
    x, info = gcr(dev, coef, b, flags)
    _, io = krylov.fgcr(coef, b, eps=1e-10, precond=M, restart=16)
    true = resid(coef, x, b)
    print(f"FGCR {variant}: gpu {info.iterations} it ({info.refinements} restarts), oracle {io['iterations']} it, "
          f"true {true:.2e}, err vs LU {rel(x, xd):.2e}")
    assert info.status == _lib.KM_OK and info.fallback == 0 and true < 1e-10
    assert abs(info.iterations - io["iterations"]) <= 2
    assert rel(x, xd) < 3e-5


def test_fgcr_fp32_directions(c1_pcor, dev):
    coef, b, xd = c1_pcor
    x, info = gcr(dev, coef, b, _lib.KM_FLAG_COARSE | _lib.KM_FLAG_MIXED, eps=1e-8)
    true = resid(coef, x, b)
    print(f"FGCR two-level FP32 directions: {info.iterations} it, true {true:.2e}")
    assert info.status == _lib.KM_OK and info.fallback == 0 and true < 1e-8
    assert rel(x, xd) < 3e-3


def test_fgcr_with_untrained_network(c1_pcor, dev):
    """The network after the coarse correction, untrained: the solve ends at
    the requested true residual (flexible GCR's residual never grows; when a
    restart cycle gains less than 10%, plain BiCGSTAB finishes from x)."""
    coef, b, xd = c1_pcor
    x, info = gcr(dev, coef, b, _lib.KM_FLAG_COARSE | _lib.KM_FLAG_MIXED, params=unet.init_params(13), eps=1e-8)
    true = resid(coef, x, b)
    print(f"FGCR coarse + untrained network: {info.iterations} it, fallback {info.fallback}, true {true:.2e}")
    assert info.status == _lib.KM_OK and true < 1e-8


def test_fgcr_zero_rhs(c1_pcor, dev):
    coef, b, _ = c1_pcor
    x, info = gcr(dev, coef, np.zeros_like(b), _lib.KM_FLAG_COARSE)
    assert info.status == _lib.KM_OK and info.iterations == 0 and not x.any()


def test_fgcr_fp32_directions_c2(dev):
    """FP32 residual-space directions on the C2 p' system (the coarse
    correction + one Jacobi step): the true residual reaches 1e-8 without the
    plain fallback (the solution-space directions stay FP64)."""
    from karman_taichi_amd.simple import SIMPLESolver
    s = SIMPLESolver.from_case(cases.c2(), device=dev)
    s.init()
    s.solve_momentum_eqn(1)
    s.correct_conserv()
    s.compute_coef_p()
    coef, b = s._coef_p.cpu().numpy(), s.b_p.cpu().numpy()
    out = {}
    for mixed in (False, True):
        flags = _lib.KM_FLAG_COARSE | _lib.KM_FLAG_JACOBI | (_lib.KM_FLAG_MIXED if mixed else 0)
        x, info = gcr(dev, coef, b, flags, eps=1e-8)
        true = resid(coef, x, b)
        out["fp32" if mixed else "fp64"] = (info.iterations, info.fallback, true)
        assert info.status == _lib.KM_OK and info.fallback == 0 and true < 1e-8, (mixed, info.iterations)
    print("C2 FGCR coarse + Jacobi (iterations, fallback, true residual):", out)
