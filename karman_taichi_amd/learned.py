"""Learned p' preconditioner: parameters, harvesting and training (host side).

The network's forward and backward passes, the residual loss, Adam and the
flexible BiCGSTAB that uses the network run in libkarman_hip.so
(``csrc/km_unet.hip``, ``csrc/km_fbicg.hip``).  This module owns the FP32
parameter vector and the Adam state on the device, harvests training pairs
from plain GPU solves and runs the training loop.

The network fills the right-preconditioner hooks ``update_phat`` /
``update_shat`` of the reference's BiCGSTAB (bicgsolver.py:48-58, identity
there; karman_taichi.py:399,419 plugs Jacobi into the same hooks).

Architecture: a 4-level U-Net of 3x3 convolutions, 16 -> 32 -> 64 -> 128
channels (2x2 average pools down, 2x nearest upsampling and skip
concatenation up, ReLU), a 1x1 output layer; inputs (r / rs, aP / abar,
identity-row mask), output scaled by rs / abar.  Parameters in the order
e0 e1 e2 e3 d2 d1 d0 (W[cout][3][3][cin], b[cout]) then w_out[16], b_out.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
import time
from pathlib import Path
from typing import Iterable, Optional, Sequence

import numpy as np
import torch

from . import _lib
from ._lib import KmSolveInfo, ptr

# (name, input channels, output channels) in parameter order
LAYERS = (("e0", 3, 16), ("e1", 16, 32), ("e2", 32, 64), ("e3", 64, 128),
          ("d2", 192, 64), ("d1", 96, 32), ("d0", 48, 16))
OUT_IN = 16

WEIGHTS_DIR = Path(__file__).resolve().parent / "weights"


def param_count() -> int:
    return sum(co * 9 * ci + co for _, ci, co in LAYERS) + OUT_IN + 1


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal convolution weights (std sqrt(2 / (9 cin))), zero biases,
    N(0, 1/16) output weights."""
    rng = np.random.default_rng(seed)
    parts = []
    for _, ci, co in LAYERS:
        parts.append(rng.standard_normal(co * 9 * ci) * math.sqrt(2.0 / (9 * ci)))
        parts.append(np.zeros(co))
    parts.append(rng.standard_normal(OUT_IN) * 0.25)
    parts.append(np.zeros(1))
    return np.concatenate(parts).astype(np.float32)


class UNetPreconditioner:
    """FP32 parameters + Adam state of the network on one device."""

    def __init__(self, params=None, device=None, seed: int = 0):
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device())
        self.device = torch.device(device)
        n = _lib.unet_param_count()
        if n != param_count():
            raise _lib.KarmanError(f"library parameter count {n} != {param_count()}")
        src = init_params(seed) if params is None else np.asarray(params, dtype=np.float32)
        if src.shape != (n,):
            raise ValueError(f"expected {n} parameters, got {src.shape}")
        self.params = torch.tensor(src, dtype=torch.float32, device=self.device)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.grad = torch.zeros_like(self.params)
        self.step_count = 0
        self.ctx = _lib.context_for(self.device)

    # ---------------------------------------------------------------- io
    def save(self, path):
        np.savez(path, params=self.params.cpu().numpy(), m=self.m.cpu().numpy(), v=self.v.cpu().numpy(),
                 step=np.array([self.step_count]))

    @classmethod
    def load(cls, path, device=None) -> "UNetPreconditioner":
        with np.load(path, allow_pickle=False) as d:
            net = cls(params=d["params"], device=device)
            if "m" in d:
                net.m.copy_(torch.from_numpy(d["m"]))
                net.v.copy_(torch.from_numpy(d["v"]))
                net.step_count = int(d["step"][0])
        return net

    # ---------------------------------------------------------------- compute
    def _bind(self):
        self.ctx.set_stream(torch.cuda.current_stream(self.device).cuda_stream)

    def apply(self, coef: torch.Tensor, r: torch.Tensor, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        """z = M(r) for one system: coef (5, NX, NY) FP64 SoA, r (NX, NY) FP64."""
        NX, NY = r.shape
        z = torch.empty_like(r) if out is None else out
        self._bind()
        self.ctx.call("km_unet_apply", ptr(self.params), NX, NY, ptr(coef), ptr(r), ptr(z))
        return z

    def loss_grad(self, coefs: torch.Tensor, rs: torch.Tensor, corrs: Optional[torch.Tensor] = None):
        """Batch mean of the relative loss of z = M(r) and its gradient (into self.grad).

        Without ``corrs``: ||r - A z||^2 / ||r||^2.  With the converged
        corrections c (A c = r): the A-norm error (c - z)^T A (c - z) / c^T r.
        coefs: (B, 5, NX, NY) FP32, rs / corrs: (B, NX, NY) FP32."""
        B, _, NX, NY = coefs.shape
        if coefs.dtype != torch.float32 or rs.dtype != torch.float32:
            raise TypeError("training batches are float32")
        coefs, rs = coefs.contiguous(), rs.contiguous()
        corrs = None if corrs is None else corrs.contiguous()
        loss = ctypes.c_double()
        self._bind()
        self.ctx.call("km_unet_loss_grad", ptr(self.params), B, NX, NY, ptr(coefs), ptr(rs),
                      None if corrs is None else ptr(corrs), ptr(self.grad), ctypes.byref(loss))
        return loss.value, self.grad

    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, grad: Optional[torch.Tensor] = None):
        self.step_count += 1
        g = self.grad if grad is None else grad
        self._bind()
        self.ctx.call("km_adam_step", ctypes.c_int64(self.params.numel()), ptr(self.params), ptr(g), ptr(self.m),
                      ptr(self.v), self.step_count, lr, b1, b2, eps)

    def train_step(self, coefs, rs, lr=1e-3, corrs=None) -> float:
        loss, _ = self.loss_grad(coefs, rs, corrs)
        self.adam_step(lr)
        return loss


def fbicgstab(coef: torch.Tensor, b: torch.Tensor, x: torch.Tensor, net: Optional[UNetPreconditioner], eps=1e-8,
              max_iter=None, x0=False, graph=True, mixed=False, coarse=False, smooth=False, gcr=False,
              jacobi=False, mg=False, symmetric=False):
    """Flexible BiCGSTAB with the network as right preconditioner
    (km_fbicgstab); ``net=None`` uses the Jacobi preconditioner instead.
    ``mixed``: FP32 Krylov vectors (FP64 scalars and refinement).
    ``coarse``: two-level preconditioner (aggregation coarse correction, then
    the network, or red-black SOR sweeps without one).  ``smooth``: one
    red-black Gauss-Seidel sweep as the preconditioner (no network, no coarse
    level; the momentum systems).  ``gcr``: flexible GCR as the outer method
    (KM_FLAG_GCR), which a non-linear preconditioner cannot break;
    ``jacobi`` (with ``gcr`` and ``coarse``, no network): one Jacobi step
    after the coarse correction instead of the SOR sweeps.  ``mg`` (no
    network, FP64): one multigrid V-cycle as the preconditioner
    (KM_FLAG_MG).  ``symmetric`` (with ``coarse``): the caller asserts the
    system is symmetric bit for bit (KM_FLAG_SYMMETRIC)."""
    from .solvers import SolveInfo
    NX, NY = b.shape
    info = KmSolveInfo()
    flags = (_lib.KM_FLAG_X0 if x0 else 0) | (0 if graph else _lib.KM_FLAG_NO_GRAPH)
    flags |= (_lib.KM_FLAG_MIXED if mixed else 0) | (_lib.KM_FLAG_COARSE if coarse else 0)
    flags |= _lib.KM_FLAG_SMOOTH if smooth else 0
    flags |= (_lib.KM_FLAG_GCR if gcr else 0) | (_lib.KM_FLAG_JACOBI if jacobi else 0)
    flags |= _lib.KM_FLAG_MG if mg else 0
    flags |= _lib.KM_FLAG_SYMMETRIC if symmetric else 0
    ctx = _lib.context_for(b.device) if net is None else net.ctx
    ctx.set_stream(torch.cuda.current_stream(b.device).cuda_stream)
    ctx.call("km_fbicgstab", NX, NY, ptr(coef), ptr(b), ptr(x), float(eps), int(max_iter or 0), flags,
             None if net is None else ptr(net.params), ctypes.byref(info))
    return SolveInfo.from_c(info)



def fbicgstab_pair(systems, eps=1e-8, max_iter=None, graph=True, smooth=True):
    """Two independent flexible BiCGSTAB solves ``systems = ((coef, b, x), (coef, b, x))``
    with the same settings (km_fbicgstab2): with ``smooth`` their first sweeps
    iterate in the same launches; results as two :func:`fbicgstab` calls."""
    from .solvers import SolveInfo
    (c0, b0, x0), (c1, b1, x1) = systems
    infos = (KmSolveInfo(), KmSolveInfo())
    flags = (0 if graph else _lib.KM_FLAG_NO_GRAPH) | (_lib.KM_FLAG_SMOOTH if smooth else 0)
    ctx = _lib.context_for(b0.device)
    ctx.set_stream(torch.cuda.current_stream(b0.device).cuda_stream)
    ctx.call("km_fbicgstab2", b0.shape[0], b0.shape[1], ptr(c0), ptr(b0), ptr(x0), b1.shape[0], b1.shape[1],
             ptr(c1), ptr(b1), ptr(x1), float(eps), int(max_iter or 0), flags, ctypes.byref(infos[0]),
             ctypes.byref(infos[1]))
    return SolveInfo.from_c(infos[0]), SolveInfo.from_c(infos[1])


# -------------------------------------------------------------------- data

@dataclasses.dataclass
class PairSet:
    """Harvested (p' residual, converged correction) pairs on the device.

    coefs: (S, 5, NX, NY) FP32 p' matrices, one per harvested system;
    rs: (P, NX, NY) FP32 residuals r_k = b - A x_k; sys: (P,) system index;
    corr: (P, NX, NY) FP32 corrections x* - x_k (A corr = r_k), optional."""
    coefs: torch.Tensor
    rs: torch.Tensor
    sys: torch.Tensor
    corr: Optional[torch.Tensor] = None
    meta: dict = dataclasses.field(default_factory=dict)

    def __len__(self):
        return self.rs.shape[0]

    def batch(self, idx: torch.Tensor):
        """(coefs, rs, corrs or None) of the pairs ``idx``, contiguous."""
        corr = None if self.corr is None else self.corr.index_select(0, idx).contiguous()
        return (self.coefs.index_select(0, self.sys.index_select(0, idx)).contiguous(),
                self.rs.index_select(0, idx).contiguous(), corr)


def harvest(case_list: Sequence, outer_steps: int, snapshots: Sequence[int] = (0, 8, 32, 128, 512),
            device=None, corrections: bool = False, progress=None, coarse: bool = False,
            noise_per_system: int = 0, seed: int = 0) -> PairSet:
    """Run the plain GPU SIMPLE loop on each case and record, at every outer
    step's p' system, the residuals r_k = b - A x_k of the plain BiCGSTAB
    iterates at the iteration counts ``snapshots`` (0: r = b) with the
    converged solve's correction x* - x_k.  The iterates come from plain
    BiCGSTAB restarted from x_prev for the next increment of iterations.

    ``noise_per_system`` adds that many white-noise residuals per system (free
    cells; their correction A^-1 r is not known, so the residual loss only).
    ``coarse``: store what the network sees inside the two-level
    preconditioner -- r1 = r - A z1 and c - z1 with z1 the coarse correction
    of r (km_coarse_correct)."""
    from .simple import SIMPLESolver
    from .solvers import BICGSolver
    if device is None:
        device = torch.device("cuda", torch.cuda.current_device())
    snaps = sorted(set(int(s) for s in snapshots))
    shapes = {c.shape_p for c in case_list}
    if len(shapes) != 1:
        raise ValueError("all harvested cases must share one grid")
    NX, NY = shapes.pop()
    S = len(case_list) * outer_steps
    per = len(snaps) + noise_per_system
    P = S * per
    coefs = torch.empty((S, 5, NX, NY), dtype=torch.float32, device=device)
    rs = torch.empty((P, NX, NY), dtype=torch.float32, device=device)
    corr = torch.empty((P, NX, NY), dtype=torch.float32, device=device) if corrections else None
    sysidx = torch.empty(P, dtype=torch.int64)
    xk = torch.zeros((NX, NY), dtype=torch.float64, device=device)
    rk = torch.empty_like(xk)
    z1 = torch.empty_like(xk)
    r1 = torch.empty_like(xk)
    gen = torch.Generator(device="cpu").manual_seed(seed)

    def put(ctx, coef, r, corr_k, slot, sys_i):
        if coarse:
            ctx.call("km_coarse_correct", NX, NY, ptr(coef), ptr(r), ptr(z1), ptr(r1))
            rs[slot].copy_(r1)
            if corr is not None and corr_k is not None:
                corr[slot].copy_(corr_k - z1)
        else:
            rs[slot].copy_(r)
            if corr is not None and corr_k is not None:
                corr[slot].copy_(corr_k)
        sysidx[slot] = sys_i
    si = pi = 0
    t0 = time.perf_counter()
    for case in case_list:
        s = SIMPLESolver.from_case(case, device=device)
        s.init()
        trunc = BICGSolver(s._coef_p, s.b_p, xk)
        for k in range(outer_steps):
            s.solve_momentum_eqn(1)
            s.correct_conserv()
            s.compute_coef_p()
            s.p_correction_solver.solve(eps=case.eps_p)
            coefs[si].copy_(s._coef_p)
            xk.zero_()
            done = 0
            for kk in snaps:
                if kk > done:
                    trunc.solve(eps=0.0, max_iter=kk - done, x0=True)
                    done = kk
                s.ctx.call("km_residual", NX, NY, ptr(s._coef_p), ptr(s.b_p), ptr(xk), ptr(rk), None)
                put(s.ctx, s._coef_p, rk, s.pcor - xk if corr is not None else None, pi, si)
                pi += 1
            if noise_per_system:
                free = (s._coef_p[1:] != 0).any(dim=0)
                for _ in range(noise_per_system):
                    noise = torch.randn((NX, NY), generator=gen, dtype=torch.float64).to(device) * free
                    put(s.ctx, s._coef_p, noise.contiguous(), None, pi, si)
                    if corr is not None:
                        corr[pi].zero_()
                    pi += 1
            si += 1
            s._correct()
            s.compute_mdiv()
            if (k + 1) % case.n_outer == 0:
                s.time_forward()
            if progress is not None and si % 25 == 0:
                progress(f"harvest: {si}/{S} systems, {pi} pairs, {time.perf_counter() - t0:.1f} s")
    torch.cuda.synchronize(device)
    return PairSet(coefs, rs, sysidx.to(device), corr,
                   meta={"cases": [c.name for c in case_list], "outer_steps": outer_steps, "snapshots": snaps,
                         "noise_per_system": noise_per_system, "coarse": coarse,
                         "seconds": time.perf_counter() - t0})


def noise_residuals(coefs: torch.Tensor, generator: torch.Generator) -> torch.Tensor:
    """White-noise residuals on the free interior cells (a link is non-zero) of
    each system: equal weight on every eigenmode, the input spectrum a Krylov
    preconditioner has to handle."""
    free = (coefs[:, 1:] != 0).any(dim=1)
    noise = torch.randn(free.shape, generator=generator, dtype=torch.float32).to(coefs.device)
    return (noise * free).contiguous()


def train(net: UNetPreconditioner, pairs: PairSet, steps: int, batch: int = 8, lr: float = 1e-3, seed: int = 0,
          log_every: int = 50, progress=None, loss: str = "anorm", lr_final: Optional[float] = None,
          noise_frac: float = 0.0):
    """Adam on random batches of harvested pairs; returns [(step, mean loss over the log window)].

    ``loss``: "anorm" (needs the harvested corrections) or "residual".
    ``lr_final``: cosine decay of the learning rate from ``lr`` to it.
    ``noise_frac`` (residual loss): that fraction of each batch's residuals is
    replaced by white noise on the same systems."""
    if loss == "anorm" and pairs.corr is None:
        raise ValueError("the A-norm loss needs harvested corrections")
    if noise_frac and loss != "residual":
        raise ValueError("noise residuals have no known correction: use the residual loss")
    g = torch.Generator(device="cpu").manual_seed(seed)
    hist, window = [], []
    t0 = time.perf_counter()
    n_noise = int(round(noise_frac * batch))
    for it in range(1, steps + 1):
        idx = torch.randint(len(pairs), (batch,), generator=g).to(pairs.rs.device)
        coefs, rs, corrs = pairs.batch(idx)
        if n_noise:
            rs[:n_noise] = noise_residuals(coefs[:n_noise], g)
        step_lr = lr if lr_final is None else lr_final + 0.5 * (lr - lr_final) * (1 + math.cos(math.pi * it / steps))
        window.append(net.train_step(coefs, rs, step_lr, corrs if loss == "anorm" else None))
        if it % log_every == 0 or it == steps:
            m = float(np.mean(window))
            hist.append((it, m))
            window = []
            if progress is not None:
                progress(f"train: step {it}/{steps}, loss {m:.4f}, {time.perf_counter() - t0:.1f} s")
    return hist


# -------------------------------------------------------------------- cost

def inference_flops(NX: int, NY: int) -> int:
    """Multiply-add FLOPs (2 per MAC) of one inference on an NX x NY padded
    system: the network runs on the interior rounded up to a multiple of 8,
    H x W, with levels at H, H/2, H/4, H/8 (3x3 convolutions) and the 1x1
    output layer at full resolution."""
    H, W = (NX - 2 + 7) // 8 * 8, (NY - 2 + 7) // 8 * 8
    level = {"e0": 0, "e1": 1, "e2": 2, "e3": 3, "d2": 2, "d1": 1, "d0": 0}
    f = 0
    for name, ci, co in LAYERS:
        s = 1 << level[name]
        f += 2 * 9 * ci * co * (H // s) * (W // s)
    return f + 2 * OUT_IN * H * W


def main(argv=None):
    """Apply the network ``--reps`` times to the p' system of the first
    outer step of the 1280 x 256 channel (a kernel-profiling workload) and
    print the mean wall time per inference and the FLOP rate."""
    import argparse
    import json
    import sys
    from . import cases
    from .simple import SIMPLESolver
    ap = argparse.ArgumentParser()
    ap.add_argument("--weights", default=str(WEIGHTS_DIR / "unet_c2.npz"))
    ap.add_argument("--reps", type=int, default=20)
    args = ap.parse_args(argv)
    dev = torch.device("cuda", 0)
    s = SIMPLESolver.from_case(cases.c2(), device=dev)
    s.init()
    s.solve_momentum_eqn(1)
    s.correct_conserv()
    s.compute_coef_p()
    net = UNetPreconditioner.load(args.weights, device=dev)
    r = s.b_p.clone()
    z = torch.empty_like(r)
    for _ in range(3):
        net.apply(s._coef_p, r, z)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    for _ in range(args.reps):
        net.apply(s._coef_p, r, z)
    torch.cuda.synchronize(dev)
    ms = (time.perf_counter() - t0) * 1e3 / args.reps
    NX, NY = r.shape
    fl = inference_flops(NX, NY)
    out = {"system": [NX, NY], "reps": args.reps, "ms_per_inference": round(ms, 4), "gflop": round(fl / 1e9, 3),
           "tflops_wall": round(fl / (ms * 1e-3) / 1e12, 2), "z_finite": bool(torch.isfinite(z).all())}
    print(json.dumps(out), flush=True)
    print(f"network inference on {NX}x{NY}: {ms:.3f} ms, {fl / 1e9:.2f} GFLOP", file=sys.stderr)


if __name__ == "__main__":
    main()
