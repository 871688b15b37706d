"""Multi-step SIMPLE runs on the GPU against the FP64 oracle.

The oracle (``oracle.fvm.simple_step(solver="direct")``) solves each of the
three systems of an outer step exactly (sparse LU), so its trajectory is the
discretisation's own; the GPU runs solve to an inner tolerance of 1e-13 (the
stopping test of every path is on the residual relative to the right-hand
side), so the two trajectories differ only by what the solves leave.

Bar (the parity protocol of the SIMPLE loop, simple.py:241-285 with the time
loop of karman_taichi.py:617-622): relative L2 error of u, v and p, and of
mdiv measured against ||b_p|| (the mass imbalance the correction removes),
at most 1e-9 after every compared step.

Paths:
* ``plain``: the reference's algorithm, BiCGSTAB for u, v and p'
  (bicgsolver.py:88-136), one C-ABI call per stage; its stop on the
  recursive residual leaves the C2 p' system's true residual near 2e-9
  (measured), so the C2 run also checks the true residual and restarts from
  x while it is above eps (``true_residual``, KM_FLAG_TRUE);
* ``fast``: the default path (one km_simple_step call per outer step: the
  u/v pair by smoothed flexible BiCGSTAB, p' by the two-level preconditioned
  flexible BiCGSTAB with its coarse inverse held across solves and updated
  every 32nd solve);
* ``mixed_fast``: the default path with FP32 Krylov vectors and FP64
  scalars, true residuals and refinement (KM_FLAG_MIXED).
"""
import numpy as np
import pytest

from karman_taichi_amd import _lib, cases
from karman_taichi_amd.simple import SIMPLESolver
from oracle import fvm

pytestmark = pytest.mark.gpu

EPS = 1e-13
BAR = 1e-9


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def gpu_solver(c, dev, path, true_residual=False):
    s = SIMPLESolver.from_case(c, device=dev)
    s.init()
    s.true_residual = true_residual
    s.eps_mom = s.eps_p = EPS
    s.mom_smooth = s.p_coarse = path != "plain"
    s.mixed = path == "mixed_fast"
    s.native_step = path != "plain"
    return s


def errors(s, st):
    out = {k: rel(getattr(s, k).cpu().numpy(), getattr(st, k)) for k in ("u", "v", "p")}
    out["mdiv"] = float(np.linalg.norm(s.mdiv.cpu().numpy() - st.mdiv) / np.linalg.norm(st.b_p))
    return out


def check(errs, where):
    for k, e in errs.items():
        assert e <= BAR, (where, k, e, errs)


@pytest.fixture(scope="module")
def c1_oracle_run():
    """The oracle's first 10 time steps of C1 (8 outer steps each), with the
    fields after every time step."""
    c = cases.c1()
    m = cases.make_mask(c)
    st = fvm.State(c, *cases.initial_fields(c, m), m)
    snaps = []
    for ts in range(10):
        for k in range(c.n_outer):
            fvm.simple_step(st, solver="direct")
        snaps.append({k: np.array(getattr(st, k)) for k in ("u", "v", "p", "mdiv", "b_p")})
        fvm.time_forward(st)
    return c, snaps


class _Snap:
    def __init__(self, d):
        self.__dict__.update(d)


@pytest.mark.parametrize("path", ["plain", "fast", "mixed_fast"])
def test_c1_first_ten_time_steps(path, c1_oracle_run, dev):
    """C1 (320x64, Re 150, dt 0.01), 10 time steps x 8 outer steps from the
    seeded start; compared after every time step.  80 outer steps take the
    fast paths' coarse level through two inverse updates."""
    c, snaps = c1_oracle_run
    s = gpu_solver(c, dev, path)
    worst = {}
    iters = []
    for ts in range(10):
        for k in range(c.n_outer):
            info = s.step()
            for f in "uvp":
                assert info[f].status == _lib.KM_OK, (ts, k, f, info[f])
            iters.append(tuple(info[f].iterations for f in "uvp"))
        errs = errors(s, _Snap(snaps[ts]))
        worst = {k: max(worst.get(k, 0.0), e) for k, e in errs.items()}
        check(errs, f"time step {ts}")
        s.time_forward()
    print(f"C1 {path}: worst over 10 time steps {worst}; (u, v, p') iterations first/last {iters[0]} {iters[-1]}")


@pytest.fixture(scope="module")
def c2_oracle_run():
    """The oracle's first two outer steps of C2 (1280x256, Re 150, dt
    0.0025) from the seeded start."""
    c = cases.c2()
    m = cases.make_mask(c)
    st = fvm.State(c, *cases.initial_fields(c, m), m)
    snaps = []
    for k in range(2):
        fvm.simple_step(st, solver="direct")
        snaps.append({k: np.array(getattr(st, k)) for k in ("u", "v", "p", "mdiv", "b_p")})
    return c, snaps


@pytest.mark.parametrize("path", ["plain", "fast", "mixed_fast"])
def test_c2_steps_vs_oracle(path, c2_oracle_run, dev):
    """C2 in full FP64 (plain and the default path) and with FP32 vectors
    (mixed_fast): two outer steps, each compared with the oracle."""
    c, snaps = c2_oracle_run
    s = gpu_solver(c, dev, path, true_residual=path == "plain")
    for k, snap in enumerate(snaps):
        info = s.step()
        for f in "uvp":
            assert info[f].status == _lib.KM_OK, (k, f, info[f])
        errs = errors(s, _Snap(snap))
        print(f"C2 {path} outer step {k}: {errs}; iterations {[info[f].iterations for f in 'uvp']}, "
              f"p' true residual {info['p'].rel_resid_true:.2e}")
        check(errs, f"outer step {k}")
