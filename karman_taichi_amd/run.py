"""Time-accurate runs of one channel case with drag, lift and Strouhal number.

    python -m karman_taichi_amd.run --case c5 --time-steps 300 \\
        --paths default,mixed,tight --reference tight --out run_c5.json

Each path advances the same seeded start state by ``--time-steps`` time
steps of ``n_outer`` SIMPLE outer iterations (the time loop of
karman_taichi.py:628-659, with u0 <- u after each time step as in
:617-622).  After every time step the drag and lift coefficients on the
obstacle are recorded (``post.obstacle_forces``).  At the end the fields of
every path are compared with the reference path's (relative L2 error of u,
v and p), and Cd (mean over the last half of the run), Cl amplitude and the
Strouhal number of the lift signal are reported.

Paths:
  default  the package's default step: smoothed u/v pair, two-level p'
           (FP64 vectors), inner eps 1e-8
  mixed    the same with FP32 Krylov vectors (FP64 scalars, true residuals
           and refinement), inner eps 1e-8
  tight    the default step with inner eps 1e-10 (the FP64 comparator: its
           p' solves reach 1e-10 in the true residual)
  plain    plain BiCGSTAB for u, v and p' (bicgsolver.py:88-136), eps 1e-8
  learned  the two-level p' preconditioner with the U-Net as its smoother in
           FP32 vectors (``--weights``), smoothed u/v pair, eps 1e-8

Prints progress to stderr and one JSON document to stdout (and ``--out``).
"""
from __future__ import annotations

import argparse
import json
import math
import statistics
import sys
import time

import numpy as np
import torch

from . import cases, post
from .simple import SIMPLESolver


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_solver(case, path, dev, weights=None):
    s = SIMPLESolver.from_case(case, device=dev)
    s.init()
    s.mixed = False
    s.eps_mom = s.eps_p = 1e-8
    if path in ("default", "mixed", "tight", "learned"):
        s.mom_smooth = s.p_coarse = True
    if path == "mixed":
        s.mixed = True
    if path == "tight":
        s.eps_mom = s.eps_p = 1e-10
    if path == "learned":
        from . import learned
        if not weights:
            raise SystemExit("the learned path needs --weights")
        s.p_precond = learned.UNetPreconditioner.load(weights, device=dev)
        s.mixed = True
    return s


def run_path(case, path, steps, dev, weights=None, progress_every=50):
    s = make_solver(case, path, dev, weights)
    d = case.obstacle_diameter()
    cd, cl, p_it, p_true, mdiv = [], [], [], [], []
    t0 = time.perf_counter()
    for ts in range(steps):
        for k in range(case.n_outer):
            info = s.step()
            p_it.append(info["p"].iterations)
            p_true.append(info["p"].rel_resid_true)
            for f in "uvp":
                if info[f].status != 0:
                    raise RuntimeError(f"{path}: time step {ts} outer {k}: {f} solve status {info[f].status}")
        s.time_forward()
        fx, fy = s.forces()
        c_d, c_l = post.force_coefficients(fx, fy, d, s.rho)
        cd.append(c_d)
        cl.append(c_l)
        if (ts + 1) % 10 == 0 or ts == steps - 1:
            mdiv.append(s.mdiv_stats()[1])
        if progress_every and (ts + 1) % progress_every == 0:
            log(f"{case.name} {path}: time step {ts + 1}/{steps}, Cd {c_d:.4f}, Cl {c_l:+.4f}, "
                f"p' iterations {p_it[-1]}, {time.perf_counter() - t0:.1f} s")
    torch.cuda.synchronize(dev)
    wall = time.perf_counter() - t0
    half = cd[len(cd) // 2:]
    cl_half = cl[len(cl) // 2:]
    res = {
        "wall_s": round(wall, 2),
        "outer_steps_per_s": round(steps * case.n_outer / wall, 2),
        "cd_last": cd[-1], "cd_mean_last_half": statistics.fmean(half),
        "cl_last": cl[-1], "cl_amplitude_last_half": 0.5 * (max(cl_half) - min(cl_half)),
        "strouhal": post.strouhal(cl, case.dt, d),
        "p_iters_median": statistics.median(p_it), "p_iters_max": max(p_it),
        "p_true_resid_max": max(p_true),
        "mdiv_l2_last": mdiv[-1],
        "cd": cd, "cl": cl,
    }
    fields = {f: getattr(s, f).detach().cpu().numpy() for f in ("u", "v", "p")}
    return res, fields


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--case", default="c5", choices=sorted(cases.CASES))
    ap.add_argument("--time-steps", type=int, default=None, help="default: the case's own count")
    ap.add_argument("--paths", default="default,mixed,tight")
    ap.add_argument("--reference", default="tight", help="path the others' fields are compared with")
    ap.add_argument("--weights", default=None, help="network parameters (.npz) for the learned path")
    ap.add_argument("--progress-every", type=int, default=50)
    ap.add_argument("--out", default=None)
    ap.add_argument("--signals", action="store_true", help="keep the per-time-step Cd / Cl lists in the output")
    args = ap.parse_args(argv)
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    case = cases.CASES[args.case]()
    steps = args.time_steps or case.n_steps
    paths = [p for p in args.paths.split(",") if p]
    out = {"case": case.name, "grid": [case.nx, case.ny], "re": case.re, "dt": case.dt, "n_outer": case.n_outer,
           "time_steps": steps, "convective_time": steps * case.dt, "obstacle_d": case.obstacle_diameter(),
           "paths": {}}
    fields = {}
    for path in paths:
        log(f"{case.name}: path {path}, {steps} time steps x {case.n_outer} outer steps")
        res, fields[path] = run_path(case, path, steps, dev, args.weights, args.progress_every)
        if not args.signals:
            res.pop("cd")
            res.pop("cl")
        out["paths"][path] = res
        log(f"{case.name} {path}: " + json.dumps({k: v for k, v in res.items() if k not in ("cd", "cl")}))
    ref = args.reference if args.reference in fields else None
    if ref:
        out["reference"] = ref
        for path in paths:
            if path == ref:
                continue
            err = {f: float(np.linalg.norm(fields[path][f] - fields[ref][f]) / np.linalg.norm(fields[ref][f]))
                   for f in ("u", "v", "p")}
            r, q = out["paths"][path], out["paths"][ref]
            err["cd"] = abs(r["cd_mean_last_half"] - q["cd_mean_last_half"]) / abs(q["cd_mean_last_half"])
            st_r, st_q = r["strouhal"], q["strouhal"]
            err["strouhal"] = abs(st_r - st_q) / st_q if math.isfinite(st_r) and math.isfinite(st_q) else None
            out["paths"][path]["error_vs_" + ref] = err
    text = json.dumps(out)
    if args.out:
        with open(args.out, "w") as f:
            f.write(text + "\n")
    print(text, flush=True)


if __name__ == "__main__":
    main()
