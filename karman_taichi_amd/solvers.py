"""Drop-in ``CGSolver`` / ``BICGSolver`` (cgsolver.py:6-131, bicgsolver.py:6-136).

Same constructors, method names and in-place semantics as the reference:
``solver = BICGSolver(coef, b, x); solver.solve(eps=1e-8, quiet=True)`` writes
the solution into ``x``.  Fields are torch tensors on a ROCm device in the
reference's padded shapes:

* ``coef``: (NX, NY, 5) -- either reference-layout AoS (contiguous) or a
  permuted view of SoA storage (strides (NY, 1, NX*NY), what
  ``SIMPLESolver.coef_*`` hands out) -- or native SoA (5, NX, NY);
* ``b``, ``x``: (NX, NY) contiguous FP64.

Differences, all supersets: ``solve`` returns a :class:`SolveInfo` (the
reference returns None and only prints), accepts ``max_iter``, ``mixed``
(FP32 inner solves with FP64 refinement) and ``x0`` (warm start), and raises
on invalid input instead of reading out of bounds.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math

import torch

from . import _lib
from ._lib import KmSolveInfo, ptr


@dataclasses.dataclass
class SolveInfo:
    iterations: int
    status: int
    rel_resid: float
    rel_resid_true: float
    bnorm: float
    refinements: int = 0
    fallback: int = 0

    @property
    def status_name(self) -> str:
        return _lib.STATUS_NAMES.get(self.status, str(self.status))

    @property
    def converged(self) -> bool:
        return self.status == _lib.KM_OK

    @classmethod
    def from_c(cls, s: KmSolveInfo) -> "SolveInfo":
        return cls(s.iterations, s.status, s.rel_resid, s.rel_resid_true, s.bnorm, s.refinements, s.fallback)


def as_soa(coef: torch.Tensor) -> torch.Tensor:
    """Return a contiguous (5, NX, NY) FP64 view or copy of a coefficient tensor."""
    if coef.dtype != torch.float64:
        raise TypeError("coefficients must be float64")
    if coef.dim() != 3:
        raise ValueError(f"coefficient tensor must be 3-D, got shape {tuple(coef.shape)}")
    if coef.shape[0] == 5 and coef.is_contiguous():
        return coef
    if coef.shape[2] != 5:
        raise ValueError(f"coefficient tensor must be (NX, NY, 5) or (5, NX, NY), got {tuple(coef.shape)}")
    soa = coef.permute(2, 0, 1)
    if soa.is_contiguous():
        return soa                      # SoA storage behind an (NX, NY, 5) view: zero copy
    return soa.contiguous()             # reference AoS layout: one layout conversion


def _check_vec(t: torch.Tensor, shape, name: str):
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if t.dtype != torch.float64 or tuple(t.shape) != tuple(shape) or not t.is_contiguous():
        raise ValueError(f"{name} must be a contiguous float64 tensor of shape {tuple(shape)}, "
                         f"got {t.dtype} {tuple(t.shape)}")
    if not t.is_cuda:
        raise ValueError(f"{name} must live on a ROCm device")


class CGSolver:
    """Conjugate gradient on the implicit 5-coefficient stencil (cgsolver.py:6-131)."""

    _kind = "cg"

    def __init__(self, coef, b, x):
        self.update_coef(coef, b, x)
        self.history = []

    def update_coef(self, coef, b, x):
        """Rebind coefficients, right-hand side and solution (cgsolver.py:41-47);
        zero-copy for SoA storage."""
        self._coef_in = coef
        self.coef = as_soa(coef)
        self.NX, self.NY = int(self.coef.shape[1]), int(self.coef.shape[2])
        _check_vec(b, (self.NX, self.NY), "b")
        _check_vec(x, (self.NX, self.NY), "x")
        if not self.coef.is_cuda or self.coef.device != x.device or b.device != x.device:
            raise ValueError("coef, b and x must be on the same ROCm device")
        self.b, self.x = b, x
        self.steps = self.NX * self.NY     # cgsolver.py:14
        self.ctx = _lib.context_for(x.device)

    def _refresh(self):
        # an AoS input may have been modified in place since construction
        if self._coef_in is not self.coef and not (self._coef_in.dim() == 3 and self._coef_in.shape[0] == 5):
            soa = self._coef_in.permute(2, 0, 1)
            if not soa.is_contiguous():
                self.coef = soa.contiguous()

    def reduce(self, p, q) -> float:
        """<p, q> over the whole padded field (cgsolver.py:49-54), FP64, deterministic."""
        out = ctypes.c_double()
        self.ctx.set_stream(torch.cuda.current_stream(self.x.device).cuda_stream)
        self.ctx.call("km_dot", ctypes.c_int64(p.numel()), _lib.KM_F64, ptr(p), ptr(q), ctypes.byref(out))
        return out.value

    def compute_Ax(self, out=None):
        """A x into ``out`` (cgsolver.py:62-66)."""
        out = torch.empty_like(self.x) if out is None else out
        self.ctx.set_stream(torch.cuda.current_stream(self.x.device).cuda_stream)
        self.ctx.call("km_stencil_apply", self.NX, self.NY, _lib.KM_F64, ptr(self.coef), ptr(self.x), ptr(out))
        return out

    def check_solution(self) -> float:
        """||b - A x|| recomputed from x (cgsolver.py:121-131 intent)."""
        n2 = ctypes.c_double()
        self.ctx.set_stream(torch.cuda.current_stream(self.x.device).cuda_stream)
        self.ctx.call("km_residual", self.NX, self.NY, ptr(self.coef), ptr(self.b), ptr(self.x), None,
                      ctypes.byref(n2))
        return math.sqrt(max(n2.value, 0.0))

    def _run(self, fn, eps, max_iter, flags) -> SolveInfo:
        self._refresh()
        info = KmSolveInfo()
        self.ctx.set_stream(torch.cuda.current_stream(self.x.device).cuda_stream)
        self.ctx.call(fn, self.NX, self.NY, ptr(self.coef), ptr(self.b), ptr(self.x), float(eps),
                      int(max_iter or 0), int(flags), ctypes.byref(info))
        return SolveInfo.from_c(info)

    def solve(self, eps=1e-8, quiet=True, max_iter=None, graph=True) -> SolveInfo:
        """CG from x0 = 0 until ||r|| < eps ||r0|| (cgsolver.py:83-114)."""
        flags = 0 if graph else _lib.KM_FLAG_NO_GRAPH
        info = self._run("km_cg", eps, max_iter, flags)
        self.history.append(f"{info.rel_resid * info.bnorm:e}\n")
        if not quiet:
            print(f">>> CG {info.status_name} after {info.iterations} iterations, "
                  f"relative residual {info.rel_resid_true:.3e}")
        return info

    def save_history(self, path="convergence.txt"):
        """cgsolver.py:116-119"""
        with open(path, "w") as f:
            f.writelines(self.history)


class BICGSolver(CGSolver):
    """Unpreconditioned BiCGSTAB (bicgsolver.py:6-136)."""

    _kind = "bicgstab"

    def solve(self, eps=1e-8, quiet=True, max_iter=None, mixed=False, x0=False, graph=True,
              precond=None, coarse=False, smooth=False, true_residual=False, gcr=False,
              symmetric=False) -> SolveInfo:
        """BiCGSTAB until sqrt(r.r / b.b) < eps (bicgsolver.py:130).

        ``mixed``: FP32 matvec/axpy inner solves, FP64 scalars, FP64 refinement
        judged by the true residual.  ``x0``: start from the current ``x``
        instead of zero (then the reference's r~ = b becomes r~ = b - A x0).
        ``precond``: a :class:`~karman_taichi_amd.learned.UNetPreconditioner`
        in the p^ / s^ hooks (bicgsolver.py:48-58): flexible BiCGSTAB with
        FP64 refinement (FP32 vectors with ``mixed``), judged by the true
        residual only, with plain BiCGSTAB as the fallback.  ``coarse``: the
        two-level preconditioner (aggregation coarse correction, then the
        network -- or red-black SOR sweeps without one).  ``smooth``: flexible
        BiCGSTAB with one or two red-black Gauss-Seidel sweeps as the preconditioner
        (the momentum systems), judged by the true residual like the others.
        ``gcr``: flexible GCR instead of flexible BiCGSTAB as the outer
        method of the preconditioned solve (the network's: a non-linear
        preconditioner cannot break it).
        ``true_residual`` (plain solve): after the reference's stop on the
        recursive residual, restart from x while the true residual is above
        eps (KM_FLAG_TRUE; the reference has no such check).
        ``symmetric`` (with ``coarse``): the caller asserts the system is
        symmetric bit for bit, as the assembled p' systems are
        (KM_FLAG_SYMMETRIC: three coefficient planes in the smoothing tiles,
        identical results).
        """
        flags = (_lib.KM_FLAG_MIXED if mixed else 0) | (_lib.KM_FLAG_X0 if x0 else 0)
        flags |= _lib.KM_FLAG_TRUE if true_residual else 0
        flags |= 0 if graph else _lib.KM_FLAG_NO_GRAPH
        if precond is not None or coarse or smooth or gcr:
            from .learned import fbicgstab
            self._refresh()
            info = fbicgstab(self.coef, self.b, self.x, precond, eps=eps, max_iter=max_iter, x0=x0, graph=graph,
                             mixed=mixed, coarse=coarse, smooth=smooth, gcr=gcr, symmetric=symmetric)
        else:
            info = self._run("km_bicgstab", eps, max_iter, flags)
        if not quiet:
            print(f">>> BICG {info.status_name} after {info.iterations} iterations, "
                  f"relative residual {info.rel_resid_true:.3e}")
        return info
