"""Group solves and steps: the two-level p' solves of several cases with their
first sweeps in shared launches (km_fbicgstab_n, km_simple_step_n) against
the same systems solved and stepped one by one.

Bars
* km_fbicgstab_n: every system's x bit for bit and its iteration count and
  status equal to its own km_fbicgstab on the same coarse level and smoothing
  tile (each workgroup of a shared launch runs its system's own code);
* km_simple_step_n / BatchedSIMPLE.step_cases(group=...): each case within the
  solve tolerance of the same case stepped alone (its own solver objects, so
  its own tile choice), FP32 and FP64 cases mixed.
"""
import ctypes

import numpy as np
import pytest
import torch

from karman_taichi_amd import _lib, cases
from karman_taichi_amd._lib import KmSolveInfo
from karman_taichi_amd.batch import BatchedSIMPLE
from karman_taichi_amd.simple import SIMPLESolver

pytestmark = pytest.mark.gpu

FLAGS = _lib.KM_FLAG_COARSE | _lib.KM_FLAG_SYMMETRIC


def channel(k, nx, re):
    ny = nx // 5
    return cases.channel(f"g{k}_{nx}", nx, ny, max(2, round(ny / 8)), re, 0.64 * 5.0 / nx, 10, seed=k)


def pcor_system(case, dev):
    """The assembled p' system of the case's first outer step (device tensors)."""
    s = SIMPLESolver.from_case(case, device=dev)
    s.init()
    s.solve_momentum_eqn(1)
    s.correct_conserv()
    s.compute_coef_p()
    torch.cuda.synchronize()
    return s._coef_p.clone(), s.b_p.clone()


def solve_one(ctx, coef, b, flags=FLAGS):
    x = torch.full_like(b, 3.0)
    info = KmSolveInfo()
    ctx.call("km_fbicgstab", b.shape[0], b.shape[1], ctypes.c_void_p(coef.data_ptr()), ctypes.c_void_p(b.data_ptr()),
             ctypes.c_void_p(x.data_ptr()), 1e-8, 0, flags, None, ctypes.byref(info))
    torch.cuda.synchronize()
    return x, info


def solve_group(ctx, systems, flags=FLAGS):
    n = len(systems)
    xs = [torch.full_like(b, 3.0) for _, b in systems]
    NX = (ctypes.c_int * n)(*[b.shape[0] for _, b in systems])
    NY = (ctypes.c_int * n)(*[b.shape[1] for _, b in systems])
    coef = (ctypes.c_void_p * n)(*[c.data_ptr() for c, _ in systems])
    bb = (ctypes.c_void_p * n)(*[b.data_ptr() for _, b in systems])
    xx = (ctypes.c_void_p * n)(*[x.data_ptr() for x in xs])
    info = (KmSolveInfo * n)()
    ctx.call("km_fbicgstab_n", n, NX, NY, coef, bb, xx, 1e-8, 0, flags, info)
    torch.cuda.synchronize()
    return xs, list(info)


@pytest.fixture(scope="module")
def group_systems(dev):
    # three systems of one shape, one of another, one more of the first:
    # a group of four, then a group of one
    return [pcor_system(channel(k, nx, re), dev)
            for k, (nx, re) in enumerate(((128, 260.0), (128, 120.0), (96, 180.0), (128, 200.0), (128, 90.0)))]


def test_group_solve_matches_single_solves(dev, group_systems):
    """Each system's own solve on a private context (its coarse level factored
    afresh each time, the 16 x 32 smoothing tile pinned on every shape), then
    the five together through km_fbicgstab_n: the same bits."""
    idx = dev.index if dev.index is not None else torch.cuda.current_device()
    ctx = _lib.Context(idx, torch.cuda.current_stream(dev).cuda_stream)
    try:
        ctx.call("km_set_coarse_update", 0)
        for coef, b in group_systems:            # levels set up, the tile of each shape pinned
            solve_one(ctx, coef, b)
            ctx.call("km_twolevel_set_tile", 2)
        single = [solve_one(ctx, coef, b) for coef, b in group_systems]
        xs, infos = solve_group(ctx, group_systems)
        for k, ((xa, ia), xb, ib) in enumerate(zip(single, xs, infos)):
            print(f"system {k} {tuple(xa.shape)}: single {ia.iterations} it, group {ib.iterations} it, "
                  f"true resid {ib.rel_resid_true:.2e}")
            assert ia.status == ib.status == _lib.KM_OK
            assert ia.iterations == ib.iterations
            assert ib.rel_resid_true < 1e-8
            assert torch.equal(xa, xb), k
    finally:
        ctx.close()


def test_group_solve_falls_back_and_checks(dev, group_systems):
    """Flags the group does not take (FP32 vectors) run one by one with the
    results of single solves; shared solution vectors are refused."""
    idx = dev.index if dev.index is not None else torch.cuda.current_device()
    ctx = _lib.Context(idx, torch.cuda.current_stream(dev).cuda_stream)
    try:
        ctx.call("km_set_coarse_update", 0)
        two = group_systems[:2]
        flags = FLAGS | _lib.KM_FLAG_MIXED
        for coef, b in two:
            solve_one(ctx, coef, b, flags)
            ctx.call("km_twolevel_set_tile", 2)
        single = [solve_one(ctx, coef, b, flags) for coef, b in two]
        xs, infos = solve_group(ctx, two, flags)
        for (xa, ia), xb, ib in zip(single, xs, infos):
            assert ia.status == ib.status == _lib.KM_OK and ia.iterations == ib.iterations
            assert torch.equal(xa, xb)
        xs, _ = solve_group(ctx, [])
        assert xs == []
        coef, b = group_systems[0]
        x = torch.zeros_like(b)
        NX = (ctypes.c_int * 2)(b.shape[0], b.shape[0])
        NY = (ctypes.c_int * 2)(b.shape[1], b.shape[1])
        cc = (ctypes.c_void_p * 2)(coef.data_ptr(), group_systems[1][0].data_ptr())
        bb = (ctypes.c_void_p * 2)(b.data_ptr(), b.data_ptr())
        xx = (ctypes.c_void_p * 2)(x.data_ptr(), x.data_ptr())
        with pytest.raises(_lib.KarmanError, match="share"):
            ctx.call("km_fbicgstab_n", 2, NX, NY, cc, bb, xx, 1e-8, 0, FLAGS, None)
    finally:
        ctx.close()


def group_batch():
    out = [channel(k, nx, re) for k, (nx, re) in
           enumerate(((64, 90.0), (96, 180.0), (128, 260.0), (80, 140.0), (128, 200.0), (96, 120.0)))]
    for k, c in enumerate(out):
        c.mixed = k in (1, 4)   # two FP32 cases among FP64 ones
    out.append(cases.cavity(24, re=100.0))
    return out


@pytest.mark.parametrize("workers,group", [(1, 4), (2, 3)])
def test_group_steps_match_single_steps(workers, group, dev):
    """step_cases with grouped steps (km_simple_step_n) against every case
    stepped alone in its own precision: every solve OK, fields within the
    solve tolerance, over two outer steps."""
    alone = BatchedSIMPLE(group_batch(), device=dev)
    alone.init()
    grp = BatchedSIMPLE(group_batch(), device=dev)
    grp.init()
    for _ in range(2):
        for c, s in zip(alone.cases, alone.solvers):
            s.mom_smooth = s.p_coarse = True
            s.mixed = c.mixed
            s.step()
        infos = grp.step_cases(workers=workers, group=group)
    torch.cuda.synchronize()
    for c, a, b, inf in zip(alone.cases, alone.solvers, grp.solvers, infos):
        for i in inf:
            assert i.status == _lib.KM_OK and i.rel_resid_true < 1e-8, c.name
        for f in ("u", "v", "p"):
            x, y = getattr(a, f).cpu().numpy(), getattr(b, f).cpu().numpy()
            e = np.linalg.norm(x - y) / max(np.linalg.norm(x), 1e-300)
            print(f"{c.name} ({'FP32' if c.mixed else 'FP64'}) {f}: {e:.2e}")
            assert e < 1e-6, (c.name, f, e)
