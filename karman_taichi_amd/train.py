"""Train the learned p' preconditioner on one GPU and measure what it gains.

    python -m karman_taichi_amd.train --out karman_taichi_amd/weights/unet_c2.npz

1. harvest: plain GPU SIMPLE runs of the 1280x256 channel at several Reynolds
   numbers in [100, 250] (150 itself is held out for evaluation) record
   (p' residual, converged correction) pairs at several plain-BiCGSTAB
   iteration counts of every outer step's p' system;
2. train: Adam on random batches, loss = the relative A-norm error
   (c - M(r))^T A (c - M(r)) / c^T r with c the converged correction (or the
   residual loss ||r - A M(r)||^2 / ||r||^2), forward and backward passes in
   the library's HIP kernels;
3. evaluate: on the p' systems of the Re = 150 case from its seeded start,
   plain FP64 BiCGSTAB against the learned flexible solve (FP32 vectors, FP64
   scalars and true residual), iterations and wall time per solve, and the
   same flexible solve with an untrained network and with Jacobi (ablations).

Prints one JSON line with the harvest, training and evaluation summaries.
"""
from __future__ import annotations

import argparse
import json
import statistics
import sys
import time

import torch

from . import cases, learned
from .simple import SIMPLESolver


def c2_case(re: float, seed: int = 0) -> cases.Case:
    return cases.channel(f"c2_1280x256_re{re:g}", 1280, 256, 32, float(re), 0.0025, 500, seed=seed)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def evaluate(case, outer_steps: int, variants: dict, eps: float = 1e-8, device=None):
    """Per outer step of ``case``: plain FP64 BiCGSTAB and every variant's
    flexible solve on the same p' system (the SIMPLE run continues with the
    plain solution).  variants: label -> keyword arguments of
    ``learned.fbicgstab`` (``net``: the network or None).  Returns
    {label: {"it": [...], "ms": [...], "fallback": n, "true": max}}."""
    s = SIMPLESolver.from_case(case, device=device)
    s.init()
    res = {k: {"it": [], "ms": [], "fallback": 0, "true": 0.0} for k in ["plain", *variants]}
    x = torch.zeros_like(s.b_p)
    for k in range(outer_steps):
        s.solve_momentum_eqn(1)
        s.correct_conserv()
        s.compute_coef_p()
        for label, kw in variants.items():
            kw = dict(kw)
            net = kw.pop("net", None)
            x.zero_()
            torch.cuda.synchronize(device)
            t0 = time.perf_counter()
            info = learned.fbicgstab(s._coef_p, s.b_p, x, net, eps=eps, **kw)
            torch.cuda.synchronize(device)
            r = res[label]
            r["ms"].append((time.perf_counter() - t0) * 1e3)
            r["it"].append(info.iterations)
            r["fallback"] += info.fallback
            r["true"] = max(r["true"], info.rel_resid_true)
        torch.cuda.synchronize(device)
        t0 = time.perf_counter()
        info = s.p_correction_solver.solve(eps=eps)
        torch.cuda.synchronize(device)
        r = res["plain"]
        r["ms"].append((time.perf_counter() - t0) * 1e3)
        r["it"].append(info.iterations)
        r["true"] = max(r["true"], info.rel_resid_true)
        s._correct()
        s.compute_mdiv()
        if (k + 1) % case.n_outer == 0:
            s.time_forward()
    return res


def variants_for(net, untrained, coarse: bool) -> dict:
    """The learned solve and its ablations (keyword arguments of
    learned.fbicgstab): the network as the right preconditioner of flexible
    BiCGSTAB and of flexible GCR (``gcr``, which a non-linear preconditioner
    cannot break), in FP32 directions; the same with an untrained network
    (what training gains); with no network -- a Jacobi step after the coarse
    correction in GCR (the cheapest smoother in the network's place), and the
    default two-level solve (red-black SOR sweeps, flexible BiCGSTAB)."""
    v = {"learned_fp32": dict(net=net, coarse=coarse, mixed=True),
         "learned_gcr_fp32": dict(net=net, coarse=coarse, mixed=True, gcr=True),
         "learned_gcr_jacobi_fp32": dict(net=net, coarse=coarse, mixed=True, gcr=True, jacobi=True),
         "untrained_gcr_fp32": dict(net=untrained, coarse=coarse, mixed=True, gcr=True),
         "untrained_gcr_jacobi_fp32": dict(net=untrained, coarse=coarse, mixed=True, gcr=True, jacobi=True)}
    if coarse:
        v["coarse_jacobi_gcr_fp32"] = dict(net=None, coarse=True, mixed=True, gcr=True, jacobi=True)
        v["twolevel_sor"] = dict(net=None, coarse=True)
        v["twolevel_sor_fp32"] = dict(net=None, coarse=True, mixed=True)
    else:
        v["jacobi_gcr_fp32"] = dict(net=None, mixed=True, gcr=True)
    return v


def summarize(res):
    out = {}
    for k, r in res.items():
        out[k] = {"iters_median": statistics.median(r["it"]), "ms_median": round(statistics.median(r["ms"]), 3),
                  "fallbacks": r["fallback"], "true_resid_max": r["true"], "iters": r["it"]}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--re", default="100,125,175,200,225,250")
    ap.add_argument("--outer-steps", type=int, default=80, help="harvested outer steps per Reynolds number")
    ap.add_argument("--snapshots", default="0,4,16,64,256,1024")
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-final", type=float, default=None, help="cosine decay to this learning rate")
    ap.add_argument("--loss", default="anorm", choices=["anorm", "residual"])
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--init", default=None, help="start from these weights (.npz)")
    ap.add_argument("--out", default=str(learned.WEIGHTS_DIR / "unet_c2.npz"))
    ap.add_argument("--eval-steps", type=int, default=8)
    ap.add_argument("--eps", type=float, default=1e-8)
    ap.add_argument("--noise-frac", type=float, default=0.0, help="residual loss: fraction of white-noise residuals")
    ap.add_argument("--eval-only", action="store_true", help="evaluate --init (and the baselines) without training")
    ap.add_argument("--coarse", action="store_true",
                    help="train the network as the smoother of the two-level preconditioner (residuals after the "
                         "coarse correction) and evaluate the two-level solve")
    ap.add_argument("--noise-per-system", type=int, default=0, help="white-noise residuals harvested per system")
    args = ap.parse_args(argv)
    if args.noise_per_system and args.loss == "anorm":
        ap.error("noise residuals have no known correction: use --loss residual")
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    res = [float(v) for v in args.re.split(",") if v]
    snaps = [int(v) for v in args.snapshots.split(",") if v]
    case_list = [c2_case(re, seed=args.seed + i) for i, re in enumerate(res)]
    if args.eval_only:
        net = learned.UNetPreconditioner.load(args.init, device=dev) if args.init else \
            learned.UNetPreconditioner(device=dev, seed=args.seed)
        untrained = learned.UNetPreconditioner(device=dev, seed=args.seed)
        ev = evaluate(cases.c2(), args.eval_steps, variants_for(net, untrained, args.coarse), args.eps, dev)
        out = {"eval": summarize(ev), "weights": args.init, "coarse": args.coarse}
        log(json.dumps({k: {kk: vv for kk, vv in v.items() if kk != "iters"} for k, v in out["eval"].items()}))
        print(json.dumps(out), flush=True)
        return
    log(f"harvest: Re {res}, {args.outer_steps} outer steps each, snapshots {snaps}")
    pairs = learned.harvest(case_list, args.outer_steps, snaps, device=dev, corrections=args.loss == "anorm",
                            progress=log, coarse=args.coarse, noise_per_system=args.noise_per_system,
                            seed=args.seed)
    log(f"harvest: {len(pairs)} pairs from {pairs.coefs.shape[0]} systems in {pairs.meta['seconds']:.1f} s")
    net = learned.UNetPreconditioner.load(args.init, device=dev) if args.init else \
        learned.UNetPreconditioner(device=dev, seed=args.seed)
    t0 = time.perf_counter()
    hist = learned.train(net, pairs, args.steps, batch=args.batch, lr=args.lr, seed=args.seed,
                         log_every=max(1, args.steps // 40), progress=log, loss=args.loss, lr_final=args.lr_final,
                         noise_frac=args.noise_frac)
    train_s = time.perf_counter() - t0
    out = {"harvest": {"pairs": len(pairs), "systems": int(pairs.coefs.shape[0]), "re": res,
                       "outer_steps": args.outer_steps, "snapshots": snaps, "seconds": round(pairs.meta["seconds"], 1)},
           "train": {"steps": args.steps, "batch": args.batch, "lr": args.lr, "lr_final": args.lr_final,
                     "loss": args.loss, "noise_frac": args.noise_frac, "seconds": round(train_s, 1),
                     "loss_first": hist[0][1], "loss_last": hist[-1][1], "history": hist}}
    del pairs
    torch.cuda.empty_cache()
    if args.out:
        net.save(args.out)
        log(f"saved {args.out}")
    if args.eval_steps:
        untrained = learned.UNetPreconditioner(device=dev, seed=args.seed)
        ev = evaluate(cases.c2(), args.eval_steps, variants_for(net, untrained, args.coarse), args.eps, dev)
        out["eval"] = summarize(ev)
        log(json.dumps({k: {kk: vv for kk, vv in v.items() if kk != "iters"} for k, v in out["eval"].items()}))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
