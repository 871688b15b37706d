"""MI355X-native SIMPLE hot path of houkensjtu/karman_taichi.

Drop-in Python classes (``SIMPLESolver``, ``BICGSolver``, ``CGSolver``) over
hand-written gfx950 HIP kernels in ``libkarman_hip.so`` (C-ABI:
``include/karman.h``).  Build with ``python -m karman_taichi_amd.build``.
"""
from . import cases  # noqa: F401  (pure host-side set-up, no GPU needed)

__all__ = ["cases", "SIMPLESolver", "BICGSolver", "CGSolver", "SolveInfo"]


def __getattr__(name):
    if name in ("BICGSolver", "CGSolver", "SolveInfo"):
        from . import solvers
        return getattr(solvers, name)
    if name == "SIMPLESolver":
        from .simple import SIMPLESolver
        return SIMPLESolver
    raise AttributeError(name)
