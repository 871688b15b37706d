"""The C/OpenMP oracle (oracle/cpu_simple.c, the timed CPU baseline) agrees
with the NumPy oracle: bit-exact BiCGSTAB-free parts are covered through a
full SIMPLE step at a tight inner tolerance, and BiCGSTAB against LU."""
import numpy as np
import pytest

from karman_taichi_amd import cases
from oracle import cpu, fvm, krylov


@pytest.fixture(scope="module", autouse=True)
def built():
    cpu.build()


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("mk", [lambda: cases.c1(), lambda: cases.cavity(24, 100.0)])
def test_simple_step_matches_numpy_oracle(mk):
    c = mk()
    m = cases.make_mask(c)
    init = cases.initial_fields(c, m)
    a = fvm.State(c, *init, m)
    b = cpu.State(c, *init, m)
    for _ in range(2):
        fvm.simple_step(a, solver="direct")
        st, it = b.step(eps_mom=1e-13, eps_p=1e-13)
        assert st == 0, (st, it)
        for name in ("u", "v", "p"):
            assert rel(getattr(b, name), getattr(a, name)) < 1e-9, name


def test_bicgstab_matches_lu():
    c = cases.c1()
    m = cases.make_mask(c)
    st = fvm.State(c, *cases.initial_fields(c, m), m)
    fvm.simple_step(st, solver="direct")
    coef, bb, _ = fvm.assemble_p(c, st.u, st.v, st.coef_u[0], st.coef_v[0], m)
    x, info = cpu.bicgstab(coef, bb, eps=1e-12)
    assert info["status"] == 0
    xd = krylov.direct_solve(coef, bb)
    assert rel(x, xd) < 1e-7
