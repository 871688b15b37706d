"""Drop-in ``SIMPLESolver`` (simple.py:9-285) on hand-written gfx950 kernels.

Fields are FP64 torch tensors on a ROCm device in the reference's padded
shapes (simple.py:27-47); ``coef_u``/``coef_v``/``coef_p`` are (NX, NY, 5)
views of SoA storage, so ``coef_u[i, j, 0]`` is aP exactly as in the
reference while the kernels stream five contiguous planes.  All arithmetic
runs in libkarman_hip.so; torch only allocates and moves memory.

Fused kernels change the intermediate state, not the result of the
reference's call sequence:

* ``compute_coef_u`` / ``compute_coef_v`` already apply the boundary folds of
  ``set_bc`` (simple.py:177-216) and the obstacle rows; ``set_bc`` then only
  writes the ghost velocities.  Calling them in the reference's order
  (compute_coef_u, compute_coef_v, set_bc) gives the reference's final state.
* ``compute_coef_p`` writes ``mdiv`` and ``b_p`` as simple.py:147-148 does.

The channel extensions (obstacle mask ``ct``, zero-gradient outlet with flux
fix, p' = 0 on the outlet column, time stepping) follow karman_taichi.py and
are selected by ``east`` and ``ct``.
"""
from __future__ import annotations

import ctypes
import math
import time
from typing import Optional

import numpy as np
import torch

from . import _lib, cases, post
from ._lib import KmCase, KmGrid, KmSolveInfo, ptr
from .solvers import BICGSolver, CGSolver, SolveInfo


class SIMPLESolver:
    def __init__(self, lx, ly, nx, ny, device=None, case: Optional[cases.Case] = None):
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device())
        self.device = torch.device(device)
        self.lx, self.ly, self.nx, self.ny = float(lx), float(ly), int(nx), int(ny)
        self.dx = self.lx / self.nx
        self.dy = self.ly / self.ny
        # reference defaults (simple.py:18-25); a Case overrides them
        self.rho = 1.00
        self.mu = 0.01
        self.dt = 10000
        self.alpha_p = 0.001
        self.alpha_u = 0.001
        self.alpha_m = 0.001
        self.bc = {'w': [0.0, 0.0], 'e': [0.0, 0.0], 'n': [0.0, 0.0], 's': [0.0, 0.0]}
        self.east = cases.EAST_WALL
        self.pin_pcor = False
        self.case = case
        if case is not None:
            self.rho, self.mu, self.dt = case.rho, case.mu, case.dt
            self.alpha_p, self.alpha_u, self.alpha_m = case.alpha_p, case.alpha_u, case.alpha_m
            self.bc = {k: list(v) for k, v in case.bc.items()}
            self.east, self.pin_pcor = case.east, case.pin_pcor

        f64 = dict(dtype=torch.float64, device=self.device)
        su, sv, sp = (nx + 3, ny + 2), (nx + 2, ny + 3), (nx + 2, ny + 2)
        self.u = torch.zeros(su, **f64)
        self.v = torch.zeros(sv, **f64)
        self.u_mid = torch.zeros(su, **f64)
        self.v_mid = torch.zeros(sv, **f64)
        self.u0 = torch.zeros(su, **f64)
        self.v0 = torch.zeros(sv, **f64)
        self.p = torch.zeros(sp, **f64)
        self.pcor = torch.zeros(sp, **f64)
        self.mdiv = torch.zeros(sp, **f64)
        # cell type: 1 = solid obstacle (karman_taichi.py:46-49); unused at simple.py:40
        self.ct = torch.zeros(sp, dtype=torch.uint8, device=self.device)
        self._coef_u = torch.zeros((5,) + su, **f64)
        self._coef_v = torch.zeros((5,) + sv, **f64)
        self._coef_p = torch.zeros((5,) + sp, **f64)
        self.coef_u = self._coef_u.permute(1, 2, 0)
        self.coef_v = self._coef_v.permute(1, 2, 0)
        self.coef_p = self._coef_p.permute(1, 2, 0)
        self.b_u = torch.zeros(su, **f64)
        self.b_v = torch.zeros(sv, **f64)
        self.b_p = torch.zeros(sp, **f64)
        # solvers are built once and rebound (bicgsimp.py:260-262) instead of
        # re-instantiated every outer step (simple.py:246-247, :254)
        self.u_momentum_solver = BICGSolver(self._coef_u, self.b_u, self.u_mid)
        self.v_momentum_solver = BICGSolver(self._coef_v, self.b_v, self.v_mid)
        self.p_correction_solver = BICGSolver(self._coef_p, self.b_p, self.pcor)
        self.p_correction_cg = CGSolver(self._coef_p, self.b_p, self.pcor)
        self.ctx = self.u_momentum_solver.ctx
        self._mdiv_stats = torch.zeros(2, dtype=torch.float64)
        self.eps_mom = 1e-8
        self.eps_p = 1e-8
        self.p_solver = "bicgstab"    # or "cg" (simple.py:254)
        self.mixed = False            # FP32 inner solves + FP64 refinement
        self.warm_start_p = False
        self.p_precond = None         # a learned.UNetPreconditioner: flexible BiCGSTAB for p'
        self.p_coarse = False         # two-level p' preconditioner (coarse correction + network or SOR sweeps)
        self.mom_smooth = False       # momentum: flexible BiCGSTAB with a red-black Gauss-Seidel sweep
        self.native_step = True       # step() as one km_simple_step call where the configuration allows
        self.true_residual = False    # plain solves: restart while the true residual is above eps (KM_FLAG_TRUE)
        self.p_gcr = False            # p' with a network: flexible GCR as the outer method (KM_FLAG_GCR)
        self._case = KmCase()
        self._step_info = (KmSolveInfo * 3)()
        self.last = {}

    @classmethod
    def from_case(cls, case: cases.Case, device=None) -> "SIMPLESolver":
        s = cls(case.lx, case.ly, case.nx, case.ny, device=device, case=case)
        s.eps_mom, s.eps_p = case.eps_mom, case.eps_p
        s.mixed = bool(case.mixed)
        s.ct.copy_(torch.from_numpy(cases.make_mask(case)))
        return s

    # ------------------------------------------------------------ plumbing
    def _grid(self) -> KmGrid:
        g = KmGrid()
        g.nx, g.ny = self.nx, self.ny
        g.east, g.pin_pcor = int(self.east), int(bool(self.pin_pcor))
        g.dx, g.dy, g.rho, g.mu, g.dt = self.dx, self.dy, self.rho, self.mu, float(self.dt)
        g.alpha_u, g.alpha_p, g.alpha_m = self.alpha_u, self.alpha_p, self.alpha_m
        for s, side in enumerate("wens"):
            g.bc[s][0] = float(self.bc[side][0])
            g.bc[s][1] = float(self.bc[side][1])
        return g

    def _call(self, name, *args):
        self.ctx.set_stream(torch.cuda.current_stream(self.device).cuda_stream)
        g = self._grid()
        return self.ctx.call(name, ctypes.byref(g), *args)

    # ------------------------------------------------------------ kernels
    def compute_coef_u(self):
        """simple.py:77-92 + the u half of set_bc (:177-207) + obstacle rows."""
        self._call("km_assemble_u", ptr(self.u), ptr(self.v), ptr(self.p), ptr(self.u0), ptr(self.ct),
                   ptr(self._coef_u), ptr(self.b_u))

    def compute_coef_v(self):
        """simple.py:95-110 + the v half of set_bc (:188-216) + obstacle rows."""
        self._call("km_assemble_v", ptr(self.u), ptr(self.v), ptr(self.p), ptr(self.v0), ptr(self.ct),
                   ptr(self._coef_v), ptr(self.b_v))

    def set_bc(self):
        """Ghost velocities of set_bc (simple.py:181, :186, :192, :196); the
        coefficient folds were applied by compute_coef_u/v."""
        self._call("km_apply_ghosts", ptr(self.u), ptr(self.v))

    def compute_mdiv(self):
        """simple.py:137-141."""
        self._call("km_mdiv", ptr(self.u), ptr(self.v), ptr(self.mdiv), None)

    def mdiv_stats(self):
        """(max|mdiv|, ||mdiv||_2) over interior cells; recomputes mdiv (bicgsimp.py:114-122)."""
        out = (ctypes.c_double * 2)()
        self._call("km_mdiv", ptr(self.u), ptr(self.v), ptr(self.mdiv), out)
        return out[0], math.sqrt(max(out[1], 0.0))

    def compute_coef_p(self):
        """simple.py:143-170 with the channel's p' Dirichlet cells."""
        self._call("km_assemble_p", ptr(self.u), ptr(self.v), ptr(self._coef_u), ptr(self._coef_v),
                   ptr(self.ct), ptr(self._coef_p), ptr(self.b_p), ptr(self.mdiv))

    def correct_pressure(self):
        """simple.py:218-222."""
        self._correct(1)

    def correct_velocity(self):
        """simple.py:224-230."""
        self._correct(2)

    def _correct(self, parts=3):
        """Both corrections in one launch when parts == 3."""
        self._call("km_correct", ptr(self.p), ptr(self.u), ptr(self.v), ptr(self.pcor), ptr(self._coef_u),
                   ptr(self._coef_v), ptr(self.ct), int(parts))

    def update_velocity(self):
        """simple.py:232-238."""
        self._call("km_blend", ptr(self.u), ptr(self.v), ptr(self.u_mid), ptr(self.v_mid))

    def correct_conserv(self):
        """Zero-gradient outlet + flux fix (karman_taichi.py:180-183, :586-614)."""
        if self.east == cases.EAST_OUTFLOW:
            self._call("km_outflow_fix", ptr(self.u))

    def time_forward(self):
        """karman_taichi.py:617-622."""
        self._call("km_time_forward", ptr(self.u), ptr(self.v), ptr(self.u0), ptr(self.v0))

    # ------------------------------------------------------------ solves
    def solve_momentum_eqn(self, n_iter=1):
        """simple.py:241-250."""
        iu = iv = None
        for _ in range(n_iter):
            self.compute_coef_u()
            self.compute_coef_v()
            self.set_bc()
            if self.mom_smooth and not self.mixed:
                # both systems' sweeps in the same launches (km_fbicgstab2)
                from .learned import fbicgstab_pair
                us, vs = self.u_momentum_solver, self.v_momentum_solver
                us._refresh()
                vs._refresh()
                iu, iv = fbicgstab_pair(((us.coef, us.b, us.x), (vs.coef, vs.b, vs.x)), eps=self.eps_mom)
            else:
                iu = self.u_momentum_solver.solve(eps=self.eps_mom, mixed=self.mixed, smooth=self.mom_smooth,
                                                  true_residual=self.true_residual)
                iv = self.v_momentum_solver.solve(eps=self.eps_mom, mixed=self.mixed, smooth=self.mom_smooth,
                                                  true_residual=self.true_residual)
            self.update_velocity()
        return iu, iv

    def solve_pcorrection_eqn(self):
        """simple.py:252-255 (BiCGSTAB by default, CG on request)."""
        self.compute_coef_p()
        if self.p_solver == "cg":
            return self.p_correction_cg.solve(eps=self.eps_p)
        if self.p_precond is not None or self.p_coarse:
            # the assembled p' system is symmetric bit for bit (compute_coef_p)
            return self.p_correction_solver.solve(eps=self.eps_p, x0=self.warm_start_p, precond=self.p_precond,
                                                 mixed=self.mixed, coarse=self.p_coarse, gcr=self.p_gcr,
                                                 symmetric=True)
        return self.p_correction_solver.solve(eps=self.eps_p, mixed=self.mixed, x0=self.warm_start_p,
                                             true_residual=self.true_residual)

    def init(self, seed_fields=True):
        """simple.py:257-264; with a Case, the seeded channel start state."""
        if self.case is not None and seed_fields:
            u, v, p, u0, v0 = cases.initial_fields(self.case, self.ct.cpu().numpy())
            for dst, src in ((self.u, u), (self.v, v), (self.p, p), (self.u0, u0), (self.v0, v0)):
                dst.copy_(torch.from_numpy(src))
        else:
            for t in (self.u, self.v, self.p, self.u0, self.v0):
                t.zero_()
        self.pcor.zero_()

    def _step_flags(self):
        """(momentum, p') flags of km_simple_step, or None when the step runs
        stage by stage (CG for p', a network preconditioner, or opted out)."""
        if not self.native_step or self.p_solver != "bicgstab" or self.p_precond is not None:
            return None
        mixed = _lib.KM_FLAG_MIXED if self.mixed else 0
        mixed |= _lib.KM_FLAG_TRUE if self.true_residual else 0
        mom = mixed | (_lib.KM_FLAG_SMOOTH if self.mom_smooth else 0)
        pf = mixed | (_lib.KM_FLAG_COARSE if self.p_coarse else 0) | (_lib.KM_FLAG_X0 if self.warm_start_p else 0)
        return mom, pf

    def _fill_case(self):
        """This solver's fields as a km_case (the layout contract of the C-ABI)."""
        c = self._case
        c.g = self._grid()
        for f in ("u", "v", "p", "u0", "v0", "u_mid", "v_mid", "pcor", "mdiv", "b_u", "b_v", "b_p"):
            setattr(c, f, getattr(self, f).data_ptr())
        c.coef_u, c.coef_v, c.coef_p = self._coef_u.data_ptr(), self._coef_v.data_ptr(), self._coef_p.data_ptr()
        c.mask = self.ct.data_ptr()
        return c

    @staticmethod
    def step_group(solvers, ctx=None):
        """One outer iteration of several solvers in one call
        (km_simple_step_n, on ``ctx`` or the first solver's context and the
        current stream): every case's stages, then the p' solves -- the
        two-level ones of equal flags with their first sweeps in shared
        launches -- then every case's corrections.  Each solver's result is
        that of its own ``step()``; returns each one's ``last``."""
        if not solvers:
            return []
        flags = [s._step_flags() for s in solvers]
        if any(f is None for f in flags):
            raise ValueError("step_group needs solvers whose step runs natively (km_simple_step)")
        n = len(solvers)
        ctx = ctx if ctx is not None else solvers[0].ctx
        arr = (KmCase * n)()
        for k, s in enumerate(solvers):
            arr[k] = s._fill_case()
        mom = (ctypes.c_int * n)(*[f[0] for f in flags])
        pf = (ctypes.c_int * n)(*[f[1] for f in flags])
        info = (KmSolveInfo * (3 * n))()
        eps_mom, eps_p = float(solvers[0].eps_mom), float(solvers[0].eps_p)
        if any(float(s.eps_mom) != eps_mom or float(s.eps_p) != eps_p for s in solvers):
            raise ValueError("step_group: the solvers differ in eps_mom / eps_p")
        ctx.set_stream(torch.cuda.current_stream(solvers[0].device).cuda_stream)
        ctx.call("km_simple_step_n", n, arr, eps_mom, eps_p, mom, pf, info)
        for k, s in enumerate(solvers):
            s.last = {key: SolveInfo.from_c(info[3 * k + i]) for i, key in enumerate("uvp")}
        return [s.last for s in solvers]

    def _native_step(self, mom_flags, p_flags):
        """The whole step in one call (km_simple_step): the same kernels in the
        same order as the stage-by-stage path below."""
        c = self._fill_case()
        self.ctx.set_stream(torch.cuda.current_stream(self.device).cuda_stream)
        info = self._step_info
        self.ctx.call("km_simple_step", ctypes.byref(c), float(self.eps_mom), float(self.eps_p), mom_flags, p_flags,
                      info)
        self.last = {k: SolveInfo.from_c(info[i]) for i, k in enumerate("uvp")}
        return self.last

    def step(self):
        """One SIMPLE outer iteration (simple.py:270-281 order; the outlet fix
        sits between the momentum blend and the p' assembly as in
        karman_taichi.py:645-646)."""
        flags = self._step_flags()
        if flags is not None:
            return self._native_step(*flags)
        iu, iv = self.solve_momentum_eqn(1)
        self.correct_conserv()
        ip = self.solve_pcorrection_eqn()
        self._correct()
        self.compute_mdiv()
        self.last = {"u": iu, "v": iv, "p": ip}
        return self.last

    def solve(self, n_steps=None, n_outer=None, callback=None):
        """Time loop: n_outer SIMPLE iterations per time step, then u0 <- u."""
        c = self.case
        n_steps = n_steps if n_steps is not None else (c.n_steps if c else 2000)
        n_outer = n_outer if n_outer is not None else (c.n_outer if c else 1)
        for ts in range(n_steps):
            for k in range(n_outer):
                self.step()
                if callback is not None:
                    callback(self, ts, k)
            self.time_forward()

    # ------------------------------------------------------------ outputs
    def fields_numpy(self):
        return {name: getattr(self, name).detach().cpu().numpy()
                for name in ("u", "v", "p", "pcor", "mdiv")}

    def dump_matrix(self, step, msg, folder="log"):
        """CSV dumps named as simple.py:51-53."""
        for name, val in self.fields_numpy().items():
            np.savetxt(f"{folder}/{step:06}-{name}-{msg}.csv", val, delimiter=",")

    def dump_coef(self, step, msg, folder="log"):
        """The coefficient planes and right-hand sides of the three systems as
        CSV, named as simple.py:55-75 (including the space in "bu -")."""
        for sys_, coef, b in (("u", self._coef_u, self.b_u), ("v", self._coef_v, self.b_v),
                              ("p", self._coef_p, self.b_p)):
            planes = coef.detach().cpu().numpy()
            for k, pre in enumerate(("ap", "aw", "ae", "an", "as")):
                np.savetxt(f"{folder}/{step:06}-{pre}{sys_}-{msg}.csv", planes[k], delimiter=",")
            np.savetxt(f"{folder}/{step:06}-b{sys_} -{msg}.csv", b.detach().cpu().numpy(), delimiter=",")

    def write_fields(self, time_step, folder="."):
        """The five field files the reference writes every 10th time step
        (karman_taichi.py:678-683): ``timestep{n:06}_{pcor,p,u,v,udiv}.csv``
        (udiv is mdiv), each as ``write_matrix`` writes it."""
        f = self.fields_numpy()
        src = {"pcor": f["pcor"], "p": f["p"], "u": f["u"], "v": f["v"], "udiv": f["mdiv"]}
        return [post.write_matrix(src[k], f"{folder}/{name}") for k, name in post.field_dump_names(time_step).items()]

    def forces(self):
        """Pressure + viscous force (drag, lift) per unit depth on the
        obstacle (``post.obstacle_forces``)."""
        return post.obstacle_forces(self.p.cpu().numpy(), self.u.cpu().numpy(), self.v.cpu().numpy(),
                                    self.ct.cpu().numpy(), self.dx, self.dy, self.mu)

    def drag_lift_coefficients(self, u_ref=1.0):
        d = self.case.obstacle_diameter() if self.case else 1.0
        fx, fy = self.forces()
        return post.force_coefficients(fx, fy, d, self.rho, u_ref)
