"""Batched SIMPLE: many channel cases of uneven size and Reynolds number per launch.

Each case is a full :class:`~karman_taichi_amd.simple.SIMPLESolver` (its own
fields on the device); :meth:`BatchedSIMPLE.step` advances every case by one
outer iteration with ``km_batch_step``: every stage is one launch over all
cases, the three BiCGSTAB solves run all cases' systems in the same fused
iteration kernels with per-case scalars and stopping tests, and the stage
sequences between solves replay as captured hipGraphs.  The reference runs
one case per process (simple.py:285-303); the per-case arithmetic is that of
``SIMPLESolver.step``.  :meth:`BatchedSIMPLE.step_cases` instead steps the
cases as independent ``SIMPLESolver.step`` calls, several at a time on their
own contexts and streams.
"""
from __future__ import annotations

import ctypes
from typing import List, Sequence

import torch

from . import _lib, cases as kcases
from ._lib import KmCase, KmSolveInfo
from .simple import SIMPLESolver
from .solvers import SolveInfo

_FIELDS = ("u", "v", "p", "u0", "v0", "u_mid", "v_mid", "pcor", "mdiv")


class BatchedSIMPLE:
    def __init__(self, case_list: Sequence[kcases.Case], device=None):
        if not case_list:
            raise ValueError("empty batch")
        self.cases = list(case_list)
        self.solvers: List[SIMPLESolver] = [SIMPLESolver.from_case(c, device=device) for c in self.cases]
        self.device = self.solvers[0].device
        self.ctx = self.solvers[0].ctx
        self.n = len(self.solvers)
        self._table = (KmCase * self.n)()
        self._info = (KmSolveInfo * (3 * self.n))()
        self._fill()

    def _fill(self):
        for k, s in enumerate(self.solvers):
            e = self._table[k]
            e.g = s._grid()
            for f in _FIELDS:
                setattr(e, f, getattr(s, f).data_ptr())
            e.coef_u, e.coef_v, e.coef_p = s._coef_u.data_ptr(), s._coef_v.data_ptr(), s._coef_p.data_ptr()
            e.b_u, e.b_v, e.b_p = s.b_u.data_ptr(), s.b_v.data_ptr(), s.b_p.data_ptr()
            e.mask = s.ct.data_ptr()

    def init(self):
        for s in self.solvers:
            s.init()

    def step(self, eps_mom: float = 1e-8, eps_p: float = 1e-8, max_iter: int = 0, graph: bool = True,
             coarse: bool = False, smooth: bool = False):
        """One outer iteration of every case in FP64 (``Case.mixed`` is honoured by
        :meth:`step_cases`); returns [(u, v, p') SolveInfo] per case.
        ``coarse``: p' by the two-level preconditioned flexible BiCGSTAB, case by
        case; ``smooth``: u and v by the smoothed flexible BiCGSTAB, case by case
        (km_fbicgstab2).  The stages, and the solves not selected, stay batched."""
        self.ctx.set_stream(torch.cuda.current_stream(self.device).cuda_stream)
        flags = (0 if graph else _lib.KM_FLAG_NO_GRAPH) | (_lib.KM_FLAG_COARSE if coarse else 0)
        flags |= _lib.KM_FLAG_SMOOTH if smooth else 0
        self.ctx.call("km_batch_step", self.n, self._table, float(eps_mom), float(eps_p), int(max_iter), flags,
                      self._info)
        return [tuple(SolveInfo.from_c(self._info[3 * k + f]) for f in range(3)) for k in range(self.n)]

    def step_cases(self, workers: int = 4, eps_mom: float = 1e-8, eps_p: float = 1e-8, group: int = 1):
        """One outer iteration of every case as independent ``SIMPLESolver.step``
        calls (``km_simple_step``: smoothed u/v pair, two-level p'), ``workers``
        cases at a time, each worker with its own context and stream, so that
        the latency-bound solves of different cases overlap on the GPU.  Each
        case runs in its own precision (``Case.mixed``: FP32 Krylov vectors
        with FP64 scalars and refinement).  Case k always runs on worker k mod
        ``workers`` (its coarse level stays there).  group > 1: a worker steps
        its cases ``group`` at a time with ``SIMPLESolver.step_group``
        (km_simple_step_n), whose FP64 p' solves share launches; each case's
        result is that of its own step.  Returns [(u, v, p') SolveInfo] per
        case."""
        import concurrent.futures
        main = torch.cuda.current_stream(self.device)
        if getattr(self, "_workers", None) is None or len(self._workers) != workers:
            idx = self.device.index if self.device.index is not None else torch.cuda.current_device()
            self._workers = []
            for _ in range(workers):
                st = torch.cuda.Stream(device=self.device)
                ctx = _lib.Context(idx, st.cuda_stream)
                # the workers' streams share the GPU: one momentum sweep (a
                # second saves an iteration but its work is not hidden here)
                ctx.call("km_set_momentum_sweeps", 1)
                self._workers.append((st, ctx))
            self._pool = concurrent.futures.ThreadPoolExecutor(max_workers=workers)
        for st, _ in self._workers:
            st.wait_stream(main)

        def run(w):
            st, ctx = self._workers[w]
            out = {}
            mine = list(range(w, self.n, workers))
            with torch.cuda.stream(st):
                for g0 in range(0, len(mine), max(1, group)):
                    ks = mine[g0:g0 + max(1, group)]
                    ss = [self.solvers[k] for k in ks]
                    saved = [(s.ctx, s.mom_smooth, s.p_coarse, s.eps_mom, s.eps_p, s.mixed) for s in ss]
                    try:
                        for k, s in zip(ks, ss):
                            s.ctx = ctx
                            s.mom_smooth = s.p_coarse = True
                            s.eps_mom, s.eps_p = eps_mom, eps_p
                            s.mixed = bool(self.cases[k].mixed)
                        lasts = SIMPLESolver.step_group(ss, ctx) if group > 1 else [ss[0].step()]
                    finally:
                        for s, sv in zip(ss, saved):
                            s.ctx, s.mom_smooth, s.p_coarse, s.eps_mom, s.eps_p, s.mixed = sv
                    for k, last in zip(ks, lasts):
                        out[k] = (last["u"], last["v"], last["p"])
            return out

        futures = [self._pool.submit(run, w) for w in range(workers)]
        concurrent.futures.wait(futures)
        for st, _ in self._workers:
            main.wait_stream(st)
        res = {}
        for f in futures:
            res.update(f.result())   # re-raises a worker's error once every worker has ended
        return [res[k] for k in range(self.n)]

    def time_forward(self):
        self.ctx.set_stream(torch.cuda.current_stream(self.device).cuda_stream)
        self.ctx.call("km_batch_time_forward", self.n, self._table)

    @property
    def tile(self):
        ti, tj = ctypes.c_int(), ctypes.c_int()
        self.ctx.call("km_batch_get_tile", ctypes.byref(ti), ctypes.byref(tj))
        return ti.value, tj.value

    def cells(self) -> int:
        """p' cells of the whole batch (padded)."""
        return sum(c.shape_p[0] * c.shape_p[1] for c in self.cases)
