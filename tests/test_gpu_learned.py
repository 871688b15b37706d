"""The trained network (karman_taichi_amd/weights/unet_c2.npz: trained on the
GPU by the library's own forward/backward kernels with Adam, A-norm loss, on
20,016 (residual, correction) pairs harvested from plain GPU runs of the
1280x256 channel at Re 100-250) on the C2 p' systems: iteration counts to
1e-6 and 1e-8 with and without the network (the reference's hooks,
bicgsolver.py:48-58).

Measured and stated, not assumed: every solve must end at eps in the true
residual; the learned solve may reach it through its plain-BiCGSTAB fallback
(the network gives no useful direction on these systems), and the test
reports what it gained or lost against the same solve without the network.
"""
import numpy as np
import pytest
import torch

from karman_taichi_amd import _lib, cases, learned
from karman_taichi_amd.simple import SIMPLESolver
from oracle import unet

pytestmark = pytest.mark.gpu


def test_weights_are_committed_and_trained():
    with np.load(learned.WEIGHTS_DIR / "unet_c2.npz", allow_pickle=False) as d:
        p = d["params"]
    assert p.shape == (unet.param_count(),) and np.isfinite(p).all()
    assert not np.allclose(p, learned.init_params(0))   # not the initial draw


@pytest.fixture(scope="module")
def c2_systems(dev):
    s = SIMPLESolver.from_case(cases.c2(), device=dev)
    s.init()
    out = []
    for k in range(2):
        s.solve_momentum_eqn(1)
        s.correct_conserv()
        s.compute_coef_p()
        out.append((s._coef_p.clone(), s.b_p.clone()))
        s.p_correction_solver.solve(eps=1e-8)
        s._correct()
        s.compute_mdiv()
    return out


def test_c4_iterations_with_and_without_the_network(c2_systems, dev):
    trained = learned.UNetPreconditioner.load(learned.WEIGHTS_DIR / "unet_c2.npz", device=dev)
    untrained = learned.UNetPreconditioner(device=dev, seed=0)
    rows = {}
    for eps in (1e-6, 1e-8):
        for coef, b in c2_systems:
            x = torch.zeros_like(b)
            from karman_taichi_amd.solvers import BICGSolver
            runs = {
                "plain": lambda: BICGSolver(coef, b, x).solve(eps=eps),
                "learned": lambda: learned.fbicgstab(coef, b, x, trained, eps=eps, mixed=True, coarse=True, gcr=True),
                "untrained": lambda: learned.fbicgstab(coef, b, x, untrained, eps=eps, mixed=True, coarse=True,
                                                       gcr=True),
                "no_network": lambda: learned.fbicgstab(coef, b, x, None, eps=eps, mixed=True, coarse=True, gcr=True),
                "coarse_jacobi": lambda: learned.fbicgstab(coef, b, x, None, eps=eps, mixed=True, coarse=True,
                                                           gcr=True, jacobi=True),
                "twolevel_sor": lambda: learned.fbicgstab(coef, b, x, None, eps=eps, coarse=True),
            }
            for name, fn in runs.items():
                x.zero_()
                info = fn()
                assert info.status == _lib.KM_OK, (name, eps, info)
                assert info.rel_resid_true < eps, (name, eps, info.rel_resid_true)
                rows.setdefault((name, eps), []).append((info.iterations, int(info.fallback)))
    for (name, eps), v in rows.items():
        print(f"C2 p' to {eps:g}: {name:13s} (iterations, fallback) {v}")
    for eps in (1e-6, 1e-8):
        gain = [a[0] - b[0] for a, b in zip(rows[("no_network", eps)], rows[("learned", eps)])]
        print(f"to {eps:g}: iterations saved by the network over the same solve without it: {gain}")
