"""km_bicgstab_noalloc (the entry point that allocates nothing) against
km_bicgstab through the C-ABI: same kernels and tile, so the same bits.

Two calls in a row on one context, the second on a smaller system (the
scratch array keeps the first call's contents), then a system beyond the
scratch limit, which must come back as KM_EINVAL before any launch."""
import ctypes

import numpy as np
import pytest
import torch

from karman_taichi_amd import _lib, cases
from karman_taichi_amd._lib import KmSolveInfo, ptr
from karman_taichi_amd.simple import SIMPLESolver

pytestmark = pytest.mark.gpu


def pcor_system(c, dev):
    s = SIMPLESolver.from_case(c, device=dev)
    s.init()
    s.solve_momentum_eqn(1)
    s.correct_conserv()
    s.compute_coef_p()
    torch.cuda.synchronize(dev)
    return s._coef_p.clone(), s.b_p.clone()


def solve(ctx, name, coef, b, eps):
    NX, NY = b.shape
    x = torch.full_like(b, 3.0)        # overwritten: x0 = 0 without KM_FLAG_X0
    info = KmSolveInfo()
    ctx.call(name, NX, NY, ptr(coef), ptr(b), ptr(x), eps, 0, 0, ctypes.byref(info))
    torch.cuda.synchronize()
    return x, info


def test_noalloc_matches_bicgstab(dev):
    ctx = _lib.context_for(dev)
    for c in (cases.c2(), cases.c1()):      # the larger system first, then a smaller one
        coef, b = pcor_system(c, dev)
        assert b.numel() <= _lib.NOALLOC_MAX_CELLS
        x_ref, i_ref = solve(ctx, "km_bicgstab", coef, b, 1e-8)
        x_na, i_na = solve(ctx, "km_bicgstab_noalloc", coef, b, 1e-8)
        print(f"{c.name}: {i_na.iterations} iterations, true residual {i_na.rel_resid_true:.2e}")
        assert i_na.status == _lib.KM_OK and i_ref.status == _lib.KM_OK
        assert (i_na.iterations, i_na.rel_resid, i_na.rel_resid_true, i_na.bnorm) == \
            (i_ref.iterations, i_ref.rel_resid, i_ref.rel_resid_true, i_ref.bnorm)
        assert torch.equal(x_na, x_ref)
        assert i_na.rel_resid_true < 1e-7


def test_noalloc_rejects_oversized_systems(dev):
    ctx = _lib.context_for(dev)
    NX, NY = 1024, _lib.NOALLOC_MAX_CELLS // 1024 + 1
    dummy = torch.zeros(8, dtype=torch.float64, device=dev)   # never read: the size check comes first
    info = KmSolveInfo()
    with pytest.raises(_lib.KarmanError, match="scratch"):
        ctx.call("km_bicgstab_noalloc", NX, NY, ptr(dummy), ptr(dummy), ptr(dummy), 1e-8, 0, 0,
                 ctypes.byref(info))
    with pytest.raises(_lib.KarmanError):
        ctx.call("km_bicgstab_noalloc", 8, 1, ptr(dummy), ptr(dummy), ptr(dummy), 1e-8, 0, _lib.KM_FLAG_MIXED,
                 ctypes.byref(info))
