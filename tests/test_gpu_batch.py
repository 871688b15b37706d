"""Batched SIMPLE (km_batch_step) against the single-case path.

The batched kernels run the single-case kernel bodies with each case's own
scalars and partial sums, so with the same tile shape every case must come
out bit-identical to stepping it alone -- and the single-case path is in turn
pinned to the oracle (test_gpu_parity.py).  Cases of different size, Reynolds
number, obstacle and east boundary, two outer steps and a time advance.
"""
import numpy as np
import pytest
import torch

from karman_taichi_amd import _lib, cases
from karman_taichi_amd.batch import BatchedSIMPLE
from karman_taichi_amd.simple import SIMPLESolver

pytestmark = pytest.mark.gpu


def small_batch():
    out = []
    for k, (nx, re) in enumerate(((64, 90.0), (96, 180.0), (128, 260.0), (80, 140.0))):
        ny = nx // 5
        out.append(cases.channel(f"b{k}_{nx}", nx, ny, max(2, round(ny / 8)), re, 0.64 * 5.0 / nx, 10, seed=k))
    # a closed cavity (wall east boundary, pinned p') in the same batch
    out.append(cases.cavity(24, re=100.0))
    return out


@pytest.mark.parametrize("graph", [True, False])
def test_batch_matches_single_cases_bit_exact(graph, dev):
    batch = BatchedSIMPLE(small_batch(), device=dev)
    batch.init()
    infos = [batch.step(graph=graph) for _ in range(2)]
    batch.time_forward()
    infos.append(batch.step(graph=graph))
    ti, tj = batch.tile
    print(f"batch tile {ti}x{tj}")
    for k, c in enumerate(batch.cases):
        s = SIMPLESolver.from_case(c, device=dev)
        for shape in (c.shape_u, c.shape_v, c.shape_p):
            s.ctx.call("km_set_tile", shape[0], shape[1], _lib.KM_F64, ti, tj)
        s.init()
        single = []
        for step in range(3):
            if step == 2:
                s.time_forward()
            r = s.step()
            single.append((r["u"], r["v"], r["p"]))
        b = batch.solvers[k]
        for f in ("u", "v", "p", "pcor", "mdiv"):
            np.testing.assert_array_equal(getattr(b, f).cpu().numpy(), getattr(s, f).cpu().numpy(),
                                          err_msg=f"case {c.name} field {f}")
        for step in range(3):
            for f in range(3):
                assert infos[step][k][f].iterations == single[step][f].iterations
                assert infos[step][k][f].status == single[step][f].status == _lib.KM_OK
        for shape in (c.shape_u, c.shape_v, c.shape_p):
            s.ctx.call("km_set_tile", shape[0], shape[1], _lib.KM_F64, 0, 0)
    print("iterations (u, v, p') of the last step:", [tuple(i.iterations for i in x) for x in infos[-1]])


def test_batch_uneven_c3_sample(dev):
    """Eight cases drawn like the 96-case C3 set (nx 256..2048): one step converges everywhere."""
    sample = cases.batched_cases(96, seed=0)[:8]
    batch = BatchedSIMPLE(sample, device=dev)
    batch.init()
    infos = batch.step()
    for c, inf in zip(sample, infos):
        assert all(i.status == _lib.KM_OK for i in inf), c.name
        assert all(i.rel_resid < 1e-8 for i in inf), c.name
    print("C3 sample iterations:", [(c.nx, tuple(i.iterations for i in inf)) for c, inf in zip(sample, infos)])


@pytest.mark.parametrize("smooth", [False, True])
def test_batch_twolevel_p(smooth, dev):
    """km_batch_step with KM_FLAG_COARSE: the p' systems go case by case through
    the two-level flexible BiCGSTAB (and with KM_FLAG_SMOOTH the u, v systems
    through the smoothed pair solve); every such solve reaches 1e-8 in the true
    residual and the fields agree with the plain batch to the solve tolerance."""
    plain = BatchedSIMPLE(small_batch(), device=dev)
    plain.init()
    two = BatchedSIMPLE(small_batch(), device=dev)
    two.init()
    for _ in range(2):
        ip = plain.step()
        it = two.step(coarse=True, smooth=smooth)
    for c, a, b, i2 in zip(plain.cases, plain.solvers, two.solvers, it):
        assert i2[2].status == _lib.KM_OK and i2[2].rel_resid_true < 1e-8, c.name
        if smooth:
            for f in (0, 1):
                assert i2[f].status == _lib.KM_OK and i2[f].rel_resid_true < 1e-8, (c.name, f)
        for f in ("u", "v", "p"):
            x, y = getattr(a, f).cpu().numpy(), getattr(b, f).cpu().numpy()
            e = np.linalg.norm(x - y) / max(np.linalg.norm(x), 1e-300)
            assert e < 1e-6, (c.name, f, e)
    print("two-level p' iterations:", [x[2].iterations for x in it], "plain:", [x[2].iterations for x in ip])


def test_batch_cases_concurrent(dev):
    """BatchedSIMPLE.step_cases (independent SIMPLE steps, two cases at a time
    on their own contexts and streams) against stepping each case alone with
    the same solvers: every solve OK, fields within the solve tolerance."""
    alone = BatchedSIMPLE(small_batch(), device=dev)
    alone.init()
    conc = BatchedSIMPLE(small_batch(), device=dev)
    conc.init()
    for _ in range(2):
        for s in alone.solvers:
            s.mom_smooth = s.p_coarse = True
            s.step()
        infos = conc.step_cases(workers=2)
    torch.cuda.synchronize()
    for c, a, b, inf in zip(alone.cases, alone.solvers, conc.solvers, infos):
        for i in inf:
            assert i.status == _lib.KM_OK and i.rel_resid_true < 1e-8, c.name
        for f in ("u", "v", "p"):
            x, y = getattr(a, f).cpu().numpy(), getattr(b, f).cpu().numpy()
            e = np.linalg.norm(x - y) / max(np.linalg.norm(x), 1e-300)
            assert e < 1e-6, (c.name, f, e)


def test_batch_cases_mixed_precision(dev):
    """step_cases on a batch whose cases alternate FP32 and FP64 Krylov
    vectors: every case against the same case stepped alone in its own
    precision (within the solve tolerance), and, with inner eps 1e-12, one
    FP32 case and one FP64 case against the oracle (exact LU solves) to 1e-9."""
    from oracle import fvm

    def batch_cases():
        out = small_batch()[:4]
        for k, c in enumerate(out):
            c.mixed = k % 2 == 0
        return out

    alone = BatchedSIMPLE(batch_cases(), device=dev)
    alone.init()
    conc = BatchedSIMPLE(batch_cases(), device=dev)
    conc.init()
    for _ in range(2):
        for c, s in zip(alone.cases, alone.solvers):
            s.mom_smooth = s.p_coarse = True
            s.mixed = c.mixed
            s.step()
        infos = conc.step_cases(workers=2)
    torch.cuda.synchronize()
    for c, a, b, inf in zip(alone.cases, alone.solvers, conc.solvers, infos):
        assert b.mixed == c.mixed   # the solver keeps its case's precision
        for i in inf:
            assert i.status == _lib.KM_OK and i.rel_resid_true < 1e-8, c.name
        for f in ("u", "v", "p"):
            x, y = getattr(a, f).cpu().numpy(), getattr(b, f).cpu().numpy()
            e = np.linalg.norm(x - y) / max(np.linalg.norm(x), 1e-300)
            assert e < 1e-6, (c.name, f, e)
    tight = BatchedSIMPLE(batch_cases()[:2], device=dev)
    tight.init()
    for _ in range(2):
        tight.step_cases(workers=2, eps_mom=1e-12, eps_p=1e-12)
    torch.cuda.synchronize()
    for c, s in zip(tight.cases, tight.solvers):
        m = cases.make_mask(c)
        st = fvm.State(c, *cases.initial_fields(c, m), m)
        for _ in range(2):
            fvm.simple_step(st, solver="direct")
        for f in ("u", "v", "p"):
            x = getattr(s, f).cpu().numpy()
            e = np.linalg.norm(x - getattr(st, f)) / np.linalg.norm(getattr(st, f))
            print(f"{c.name} ({'FP32' if c.mixed else 'FP64'} vectors) {f}: {e:.2e} against the oracle")
            assert e <= 1e-9, (c.name, f, e)
