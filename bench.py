#!/usr/bin/env python3
"""Benchmark: SIMPLE outer steps/s and p' solve time on one MI355X.

Workload: the 1280x256 square-cylinder channel (Re = 150, dt = 0.0025,
32x32-cell obstacle, seeded +-1e-3 v perturbation).  A "step" is one SIMPLE
outer iteration: u/v assembly with boundary folds, two BiCGSTAB momentum
solves to 1e-8, blend, outlet fix, p' assembly, BiCGSTAB p' solve to 1e-8,
corrections and the continuity residual.  Every ``n_outer`` steps the time
level advances (u0 <- u), as in a time-accurate run.

Prints ONE JSON line (rank 0).  Extra keys:
  roofline      achieved HBM GB/s of the dominant solver kernel vs 8 TB/s,
                algorithmic bytes / kernel time from HIP events on the
                solver's stream; traffic from rocprofv3 counters (profiles/)
  cpu_baseline  the C/OpenMP FP64 restatement with the reference's kernel
                decomposition (oracle/cpu_simple.c) on this host's cores,
                timed on a bounded sample of the same workload
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import math
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per cell of the three BiCGSTAB kernels (FP64):
#   bicg_p reads r, p, v, r~ and 5 coefficient planes, writes p', v'
#   bicg_s reads r, v' and 5 coefficient planes, writes s, t
#   bicg_r reads s, t, p', x, r~, writes x, r
BYTES_PER_CELL = {"bicg_p": 11 * 8, "bicg_s": 9 * 8, "bicg_r": 7 * 8,
                  # the fused two-level iteration (FP64 vectors, 1-byte free-cell mask):
                  #   tl_pv reads p, 5 coefficient planes, r~, mask; writes p^, v
                  #   tl_st reads r, v, 5 coefficient planes, r~, mask; writes s, s^, t
                  #   tl_ct (flat workgroups) reads s, t, p, v, x, p^, s^; writes x, r, p
                  "tl_pv": 9 * 8 + 1, "tl_st": 11 * 8 + 1, "tl_ct": 10 * 8,
                  # the same for a symmetric p' system (KM_FLAG_SYMMETRIC, as the
                  # step's p' solve runs them): aP, aE, aN planes only
                  "tl_pv+sym": 7 * 8 + 1, "tl_st+sym": 9 * 8 + 1}
# the coarse rows of tl_cv / tl_ct read A_c^-1 (m x m) and a few m-vectors
COARSE_VECTORS = {"tl_cv": 4, "tl_ct": 6}


def mall_note(cells, m):
    """Working set of one p' solve: 5 coefficient planes and ~10 vectors in
    FP64, and A_c^-1, against the 256 MB MALL (Infinity Cache)."""
    ws = cells * 8 * 15 + 8 * m * m
    return (f"working set of one p' system ~{ws / 1e6:.0f} MB: "
            f"{'fits' if ws < 256e6 else 'exceeds'} the 256 MB MALL")


def alg_bytes(name, cells, m, sym=False):
    """Algorithmic bytes of one launch of kernel `name` (sym: the three-plane
    smoothing kernels of a symmetric system)."""
    coarse = 8 * m * m + 8 * COARSE_VECTORS[name] * m if name in COARSE_VECTORS else 0
    key = name + "+sym" if sym and name + "+sym" in BYTES_PER_CELL else name
    return coarse + BYTES_PER_CELL.get(key, 0) * cells


# p' solver of each mode: (solver, FP32 vectors, two-level, network)
#   plain          BiCGSTAB in FP64 (bicgsolver.py)
#   mixed          FP32 BiCGSTAB inner solves + FP64 refinement (all three solves)
#   cg             CG for p' (cgsolver.py)
#   twolevel       flexible BiCGSTAB, aggregation coarse correction + damped Jacobi, FP64 vectors
#   twolevel_fp32  the same with FP32 vectors (FP64 scalars and refinement)
#   learned        flexible BiCGSTAB, coarse correction + the U-Net on the remaining residual, FP32 vectors
#   learned_fp64   the same with FP64 vectors
MODES = {"plain": ("bicgstab", False, False, False), "mixed": ("bicgstab", True, False, False),
         "cg": ("cg", False, False, False), "twolevel": ("fbicgstab", False, True, False),
         "twolevel_fp32": ("fbicgstab", True, True, False), "learned": ("fbicgstab", True, True, True),
         "learned_fp64": ("fbicgstab", False, True, True)}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--case", default="c2")
    ap.add_argument("--mode", default="twolevel", choices=list(MODES))
    ap.add_argument("--weights", default=os.path.join(ROOT, "karman_taichi_amd", "weights", "unet_c2.npz"),
                    help="network parameters for the learned modes")
    ap.add_argument("--coarse-block", type=int, default=0,
                    help="aggregate edge of the two-level coarse level (0: automatic)")
    ap.add_argument("--totals", type=int, default=-1, choices=[-1, 0, 1],
                    help="two-level partial sums: -1 automatic, 0 every workgroup sums, 1 one workgroup first")
    ap.add_argument("--plain-totals", type=int, default=-1, choices=[-1, 0, 1],
                    help="plain BiCGSTAB partial sums: -1 automatic, 0 every workgroup sums, 1 one workgroup first")
    ap.add_argument("--momentum", default="auto", choices=["auto", "plain", "rbgs"],
                    help="u/v solver: plain BiCGSTAB (bicgsolver.py) or flexible BiCGSTAB with one red-black "
                         "Gauss-Seidel sweep; auto: rbgs with the FP64 two-level modes, plain otherwise")
    ap.add_argument("--coarse-update", type=int, default=32,
                    help="two-level: update the coarse inverse every k-th p' solve (1 each, 0 factor each; default 32)")
    ap.add_argument("--chunks", default="",
                    help="two-level launch schedule big,small,lead (iterations per graph far from / near the expected "
                         "end of a solve, and where near begins); default 8,2,2")
    ap.add_argument("--case-workers", type=int, default=3,
                    help="c3 with the two-level mode: step the cases as independent SIMPLE steps, this many at a "
                         "time on their own contexts and streams (0: km_batch_step)")
    ap.add_argument("--case-group", type=int, default=1,
                    help="c3 with case workers: a worker steps this many cases per call (km_simple_step_n), their "
                         "FP64 p' solves in shared launches")
    ap.add_argument("--case-precision", choices=["per-case", "f64", "f32"], default="per-case",
                    help="c3 with case workers: each case's own precision (half the cases FP32 Krylov vectors), or "
                         "every case FP64 / FP32 vectors")
    ap.add_argument("--spinup-time-steps", type=int, default=100,
                    help="time steps the flow is advanced from the seeded start before the warm-up (untimed): the "
                         "timed window starts at this time step")
    ap.add_argument("--learned-reps", type=int, default=3,
                    help="repetitions of the learned+FP32 vs plain p' comparison on the final system (0: skip)")
    ap.add_argument("--smoothing-tile", type=int, default=-1,
                    help="two-level modes: pin the smoothing-tile shape (an index of the library's candidates) after "
                         "the first outer step instead of the autotuner's choice (-1)")
    ap.add_argument("--staged", action="store_true",
                    help="time the step stage by stage from Python instead of one km_simple_step call")
    ap.add_argument("--stage-steps", type=int, default=8,
                    help="extra stage-by-stage steps after the timed region (per-solve times)")
    ap.add_argument("--compare-reps", type=int, default=5,
                    help="repetitions of the p' solve on the final system, plain BiCGSTAB vs the mode's solver")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=24, help="outer steps in the CPU baseline sample")
    ap.add_argument("--traffic-json", default=None,
                    help="HBM bytes per launch from rocprofv3 counters (default: the mode's file under profiles/)")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="also run a device-to-device copy of this many MiB (counter calibration for rocprofv3 --pmc)")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); this build has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from karman_taichi_amd import _lib, cases
    from karman_taichi_amd.simple import SIMPLESolver

    if args.case == "c3":
        return bench_batch(args)
    case = cases.CASES[args.case]()
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    s = SIMPLESolver.from_case(case, device=dev)
    s.init()
    p_solver, fp32, twolevel, learned = MODES[args.mode]
    s.mixed = fp32
    s.p_coarse = twolevel
    momentum = args.momentum if args.momentum != "auto" else ("rbgs" if twolevel and not fp32 else "plain")
    s.mom_smooth = momentum == "rbgs"
    s.ctx.call("km_set_coarse", int(args.coarse_block))
    s.ctx.call("km_set_coarse_update", int(args.coarse_update))
    s.ctx.call("km_set_plain_totals", int(args.plain_totals))
    if p_solver == "cg":
        s.p_solver = "cg"
    if learned:
        from karman_taichi_amd.learned import UNetPreconditioner
        if not os.path.exists(args.weights):
            raise SystemExit(f"mode {args.mode} needs trained network parameters; {args.weights} does not exist")
        s.p_precond = UNetPreconditioner.load(args.weights, device=dev)
    n_outer = case.n_outer
    cells_p = case.shape_p[0] * case.shape_p[1]

    step_count = [0]
    native = not args.staged and s._step_flags() is not None

    def one_step(rec=None, staged=not native):
        if not staged:
            # the whole step in one call (km_simple_step)
            last = s.step()
            step_count[0] += 1
            if step_count[0] % n_outer == 0:
                s.time_forward()
            if rec is not None:
                rec.append({"p_it": last["p"].iterations, "u_it": last["u"].iterations,
                            "v_it": last["v"].iterations, "p_true": last["p"].rel_resid_true,
                            "p_status": last["p"].status})
            return
        t0 = time.perf_counter()
        iu, iv = s.solve_momentum_eqn(1)
        s.correct_conserv()
        t1 = time.perf_counter()
        ip = s.solve_pcorrection_eqn()      # synchronous: returns after the solve
        t2 = time.perf_counter()
        s._correct()
        s.compute_mdiv()
        step_count[0] += 1
        if step_count[0] % n_outer == 0:
            s.time_forward()
        if rec is not None:
            rec.append({"mom_ms": (t1 - t0) * 1e3, "p_ms": (t2 - t1) * 1e3, "p_it": ip.iterations,
                        "u_it": iu.iterations, "v_it": iv.iterations, "p_true": ip.rel_resid_true, "p_status": ip.status})

    log(f"bench: case {case.name} mode {args.mode}, spin-up {args.spinup_time_steps} time steps, "
        f"warmup {args.warmup}, steps {args.steps}")
    t_spin = time.perf_counter()
    for _ in range(args.spinup_time_steps * n_outer):
        one_step()
    torch.cuda.synchronize()
    if args.spinup_time_steps:
        log(f"bench: spin-up to time step {args.spinup_time_steps} in {time.perf_counter() - t_spin:.1f} s")
    for w in range(args.warmup):
        one_step()
        if w == 0 and twolevel and args.smoothing_tile >= 0:
            s.ctx.call("km_twolevel_set_tile", int(args.smoothing_tile))
    if twolevel and args.totals != -1:   # the coarse level exists after the first solve
        s.ctx.call("km_set_coarse_totals", int(args.totals))
    if twolevel and args.chunks:
        s.ctx.call("km_set_coarse_chunks", *[int(v) for v in args.chunks.split(",")])
    torch.cuda.synchronize()
    rec = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        one_step(rec)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    steps_per_s = args.steps / wall
    p_it = [r["p_it"] for r in rec]
    # per-solve host times from stage-by-stage steps after the timed region
    srec = rec if not native else []
    for _ in range(args.stage_steps if native else 0):
        one_step(srec, staged=True)
    p_ms = [r["p_ms"] for r in srec]
    p_ms_it = [r["p_it"] for r in srec]
    log(f"bench: {steps_per_s:.3f} steps/s ({'one call per step' if native else 'stage by stage'}), "
        f"p' iters median {statistics.median(p_it)}"
        + (f", staged p' median {statistics.median(p_ms):.2f} ms" if p_ms else ""))
    compare = p_compare(s, args.compare_reps) if args.compare_reps > 0 else None
    lcompare = learned_compare(s, args.weights, args.learned_reps) if args.learned_reps > 0 else None

    # ---- roofline: dominant kernel of the mode's p' iteration, timed with HIP
    # events on the solver stream (one launch of each kernel per iteration)
    import ctypes
    ctx = s.ctx
    ctx.set_stream(torch.cuda.current_stream(dev).cuda_stream)
    NX, NY = case.shape_p
    # the step's p' solve (km_simple_step) runs the smoothing tiles of a
    # symmetric system (three coefficient planes); the stage-by-stage path
    # the five-plane ones
    symk = bool(twolevel and not learned and native)
    if twolevel and not learned:
        kinds = (("tl_pv", _lib.KM_KERNEL_TL_PV), ("tl_cv", _lib.KM_KERNEL_TL_CV), ("tl_st", _lib.KM_KERNEL_TL_ST),
                 ("tl_ct", _lib.KM_KERNEL_TL_CT))
        if symk:
            kinds = tuple((nm, kid | _lib.KM_KERNEL_SYMMETRIC) for nm, kid in kinds)
        reps = 100       # well inside the stall window of the flexible iteration
    else:
        kinds = (("bicg_p", _lib.KM_KERNEL_BICG_A), ("bicg_s", _lib.KM_KERNEL_BICG_B),
                 ("bicg_r", _lib.KM_KERNEL_BICG_R))
        reps = 200
    kern, kern_ev = {}, {}
    for name, kid in kinds:
        us = ctypes.c_double()
        ctx.call("km_time_kernel", kid | _lib.KM_KERNEL_BACK_TO_BACK, NX, NY, _lib.ptr(s._coef_p), _lib.ptr(s.b_p),
                 reps, ctypes.byref(us))
        kern[name] = us.value
        ctx.call("km_time_kernel", kid, NX, NY, _lib.ptr(s._coef_p), _lib.ptr(s.b_p), reps, ctypes.byref(us))
        kern_ev[name] = us.value
    # one launch of each per iteration: the longest dominates.  The achieved rate
    # uses the back-to-back timing (events around `reps` identical launches),
    # which agrees with rocprofv3's per-kernel averages; events around every
    # dispatch of a running iteration add ~3 us of event overhead per launch.
    dom = max(kern, key=kern.get)
    tile = coarse = None
    if twolevel and not learned:
        ti, tj = ctypes.c_int(), ctypes.c_int()
        ctx.call("km_twolevel_tile", ctypes.byref(ti), ctypes.byref(tj))
        nt = ctypes.c_int()
        ctx.call("km_twolevel_tile_threads", ctypes.byref(nt))
        tile = [ti.value, tj.value, nt.value]
        blk, mc = ctypes.c_int(), ctypes.c_int()
        ctx.call("km_coarse_info", ctypes.byref(blk), ctypes.byref(mc))
        coarse = {"aggregate": blk.value, "unknowns": mc.value}
    launch_bytes = alg_bytes(dom, cells_p, coarse["unknowns"] if coarse else 0, sym=symk)
    achieved = launch_bytes / (kern[dom] * 1e-6) / 1e9
    traffic = None
    tjson = args.traffic_json or os.path.join(
        ROOT, "profiles", "k_tl_traffic.json" if twolevel and not learned else "k_bicg_traffic.json")
    try:
        with open(tjson) as f:
            tj = json.load(f)
        per_case = tj.get(args.case, {})
        # the smoothing kernels' traffic depends on the tile shape the run uses
        sfx = "+sym" if symk else ""
        traffic = per_case.get(f"{dom}{sfx}@{tile[0]}x{tile[1]}x{tile[2]}" if tile else dom + sfx)
    except (OSError, ValueError):
        pass
    log(f"bench: kernel us {kern} (dispatch events {kern_ev}), dominant {dom}: {achieved:.0f} GB/s")

    if args.calib_copy_mb:
        # a plain copy of known size: its FETCH_SIZE / WRITE_SIZE readings
        # calibrate the counters of the solver kernels in the same run
        nb = args.calib_copy_mb << 20
        src = torch.ones(nb // 8, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(4):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"bench: calibration copy of {nb} bytes x 4")

    # ---- CPU baseline: C/OpenMP restatement with the reference decomposition
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(case, args.cpu_steps)

    total_p_it = sum(p_ms_it)
    total_p_s = sum(p_ms) / 1e3
    out = {
        "metric": "SIMPLE steps/s + p' solve time to 1e-8 rel resid: learned+FP32 vs plain, 1 GPU",
        "value": round(steps_per_s, 4),
        "unit": "SIMPLE steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if fp32 else "f64",
        "data": f"synthetic (seeded channel start state advanced {args.spinup_time_steps} time steps before the "
                f"warm-up)",
        "config": {"workload": f"{case.name}: {case.nx}x{case.ny} channel, Re={case.re:g}, dt={case.dt}, "
                               f"{case.obstacle_cells}x{case.obstacle_cells} cylinder, "
                               f"{n_outer} outer steps per time step, inner eps 1e-8",
                   "mode": args.mode, "p_solver": p_solver, "p_twolevel": twolevel, "p_network": learned,
                   "p_vectors": "f32" if fp32 else "f64", "smoothing_tile": tile, "coarse": coarse,
                   "momentum": "fbicgstab+rbgs" if momentum == "rbgs" else "bicgstab",
                   "step": "km_simple_step" if native else "stage by stage"},
        "p_solve_ms_median": round(statistics.median(p_ms), 3) if p_ms else None,
        "p_iters_median": statistics.median(p_it),
        "p_iters": p_it,
        "p_ms": [round(v, 3) for v in p_ms],
        "momentum_ms_median": round(statistics.median(r["mom_ms"] for r in srec), 3) if srec else None,
        "solve_times_from": (f"{len(srec)} stage-by-stage steps after the timed region" if native
                             else "the timed steps"),
        "uv_iters_median": [statistics.median(r["u_it"] for r in rec), statistics.median(r["v_it"] for r in rec)],
        "p_rel_resid_true_max": max(r["p_true"] for r in rec),
        "p_status_all_ok": all(r["p_status"] == _lib.KM_OK for r in rec),
        "p_compare": compare,
        "learned_compare": lcompare,
        "p_cells_iters_per_s": round(cells_p * total_p_it / total_p_s, 1) if p_ms else None,
        "us_per_p_iter": round(total_p_s / total_p_it * 1e6, 3) if p_ms else None,
        "kernel_us_back_to_back": {k: round(v, 3) for k, v in kern.items()},
        "kernel_us_dispatch_events": {k: round(v, 3) for k, v in kern_ev.items()},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "alg_bytes_per_launch": launch_bytes,
                     **({"coefficient_planes": 3 if symk else 5} if twolevel and not learned else {}),
                     "note": mall_note(cells_p, coarse["unknowns"] if coarse else 0)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def p_compare(s, reps):
    """The last outer step's p' system solved from zero, ``reps`` times each,
    by plain FP64 BiCGSTAB and by the mode's solver: median wall ms (the solves
    return after a stream synchronisation), iterations, true residuals."""
    import torch
    sol = s.p_correction_solver
    runs = {"plain": dict(), "mode": dict(precond=s.p_precond, coarse=s.p_coarse, mixed=s.mixed)}
    if s.p_solver == "cg":
        return None
    out = {}
    for name, kw in runs.items():
        fn = (lambda: sol.solve(eps=s.eps_p, **kw)) if (kw.get("precond") is not None or kw.get("coarse")) else \
            (lambda: sol.solve(eps=s.eps_p, mixed=kw.get("mixed", False)))
        fn()                                  # warm: graphs, coarse set-up
        torch.cuda.synchronize()
        ms, last = [], None
        for _ in range(reps):
            t0 = time.perf_counter()
            last = fn()
            ms.append((time.perf_counter() - t0) * 1e3)
        out[name] = {"ms_median": round(statistics.median(ms), 3), "iterations": last.iterations,
                     "rel_resid_true": last.rel_resid_true, "status": last.status_name}
    out["speedup"] = round(out["plain"]["ms_median"] / out["mode"]["ms_median"], 3)
    log(f"bench: p' compare {out}")
    return out


def learned_compare(s, weights, reps):
    """The metric's second leg on the last outer step's p' system, each solve
    from zero to 1e-8 in the true residual, ``reps`` times: plain FP64
    BiCGSTAB against the learned solve -- flexible GCR whose every
    preconditioner application gives two directions, the two-level M(r)
    (coarse correction + red-black SOR sweeps) and the coarse correction
    plus the trained U-Net (BF16 on the matrix cores), FP32 residual-space
    vectors, FP64 solution-space vectors, scalars and true residual, plain
    BiCGSTAB as the fallback when a restart cycle stalls -- beside the same
    solve without the network (the two-level direction alone), the default
    two-level flexible BiCGSTAB and flexible BiCGSTAB / GCR with one
    multigrid V-cycle as the preconditioner.  Median wall ms (each solve returns after a
    stream synchronisation), iterations, true residuals, fallbacks."""
    import torch
    from karman_taichi_amd import learned as kl
    if not os.path.exists(weights):
        return {"status": f"unmeasured: no trained network parameters at {os.path.relpath(weights, ROOT)}"}
    net = kl.UNetPreconditioner.load(weights, device=s.device)
    x = torch.zeros_like(s.b_p)
    runs = {
        "plain": lambda: s.p_correction_solver.solve(eps=1e-8),
        "learned_fp32": lambda: kl.fbicgstab(s._coef_p, s.b_p, x, net, eps=1e-8, mixed=True, coarse=True, gcr=True),
        "no_network_fp32": lambda: kl.fbicgstab(s._coef_p, s.b_p, x, None, eps=1e-8, mixed=True, coarse=True,
                                                gcr=True),
        "twolevel_sor": lambda: kl.fbicgstab(s._coef_p, s.b_p, x, None, eps=1e-8, coarse=True),
        "multigrid_fbicgstab": lambda: kl.fbicgstab(s._coef_p, s.b_p, x, None, eps=1e-8, mg=True),
        "multigrid_gcr": lambda: kl.fbicgstab(s._coef_p, s.b_p, x, None, eps=1e-8, mg=True, gcr=True),
    }
    out = {"weights": os.path.relpath(weights, ROOT)}
    for name, fn in runs.items():
        fn()
        torch.cuda.synchronize()
        ms, last = [], None
        for _ in range(reps):
            t0 = time.perf_counter()
            last = fn()
            torch.cuda.synchronize()
            ms.append((time.perf_counter() - t0) * 1e3)
        out[name] = {"ms_median": round(statistics.median(ms), 3), "iterations": last.iterations,
                     "rel_resid_true": last.rel_resid_true, "status": last.status_name,
                     "fallback": bool(getattr(last, "fallback", 0))}
    out["speedup_vs_plain"] = round(out["plain"]["ms_median"] / out["learned_fp32"]["ms_median"], 3)
    out["speedup_vs_no_network"] = round(out["no_network_fp32"]["ms_median"] / out["learned_fp32"]["ms_median"], 3)
    log(f"bench: learned compare {out}")
    return out


def coarse_unknowns(nx, ny):
    """Coarse unknowns of the two-level level of an nx x ny interior: the
    smallest aggregate of 8, 12, 16, 24, ... 256 that keeps them <= 2048, or
    on larger grids within 4 m^2 <= 21 nx ny (kmcoarse::choose_block)."""
    for B in (8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256):
        m = -(-nx // B) * -(-ny // B)
        if -(-ny // B) > 63:
            continue
        if m <= 2048 or (m <= 8192 and 4 * m * m <= 21 * nx * ny):
            return m
    return -(-nx // 512) * -(-ny // 512)


def bench_batch(args):
    """The 96 uneven channel cases (nx 256..2048, Re 80..300) advanced together
    with km_batch_step: one launch per SIMPLE stage over all cases; momentum
    and p' batched (plain modes) or case by case with the smoothed momentum
    and two-level p' solvers (two-level modes).  A step = one outer iteration
    of every case."""
    import torch
    from karman_taichi_amd import cases
    from karman_taichi_amd.batch import BatchedSIMPLE
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    case_list = cases.batched_cases(96, seed=0)
    if args.case_precision != "per-case":
        case_list = [dataclasses.replace(c, mixed=args.case_precision == "f32") for c in case_list]
    b = BatchedSIMPLE(case_list, device=dev)
    b.init()
    coarse = MODES[args.mode][2]
    fp32 = MODES[args.mode][1]
    # momentum as the single-case bench: the smoothed solve with the two-level p' solve
    momentum = args.momentum if args.momentum != "auto" else ("rbgs" if coarse and not fp32 else "plain")
    smooth = momentum == "rbgs"
    log(f"bench c3: {b.n} cases, {b.cells()} p' cells, warmup {args.warmup}, steps {args.steps}, mode {args.mode}")
    workers = args.case_workers if coarse and smooth else 0

    def batch_step():
        return b.step_cases(workers, group=args.case_group) if workers > 0 else b.step(coarse=coarse, smooth=smooth)

    for _ in range(args.warmup):
        batch_step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    recs = [batch_step() for _ in range(args.steps)]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    p_it = [inf[2].iterations for rec in recs for inf in rec]
    n_fp32 = sum(bool(c.mixed) for c in case_list) if workers else 0
    cell_its = sum(c.shape_p[0] * c.shape_p[1] * inf[2].iterations for rec in recs for c, inf in zip(case_list, rec))
    # algorithmic bytes of the p' iterations alone (the two-level path's
    # K_PV, K_CV, K_ST, K_CT: per cell 57 + 73 + 80 B with FP64 vectors, 41 +
    # 49 + 40 B with FP32 vectors and FP64 coefficient planes, plus two reads
    # of the m x m FP64 coarse inverse and the coarse vectors); the momentum
    # solves, stages and refinement residuals are not counted
    p_bytes = 0.0
    if coarse:
        for rec in recs:
            for c, inf in zip(case_list, rec):
                m = coarse_unknowns(c.nx, c.ny)
                per_cell = (41 + 49 + 40) if (workers and c.mixed) else (57 + 73 + 80)
                cells = c.shape_p[0] * c.shape_p[1]
                p_bytes += inf[2].iterations * (per_cell * cells + 16.0 * m * m + 80.0 * m)
    out = {"metric": "batched SIMPLE outer steps/s (every case advanced one outer iteration)",
           "value": round(args.steps / wall, 4), "unit": "batch steps/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64" if n_fp32 else "f64",
           "data": "synthetic (seeded channel start states)",
           "config": {"workload": f"96 uneven channels nx 256..2048, ny nx/5, D ny/8, Re 80..300, "
                                  f"{b.cells()} p' cells in total", "tile": None if workers else list(b.tile),
                      "p_solver": "two-level flexible BiCGSTAB per case" if coarse else "batched BiCGSTAB",
                      "momentum": "fbicgstab+rbgs per case (u, v together)" if smooth else "batched BiCGSTAB",
                      "step": (f"km_simple_step_n, {args.case_group} cases per call, {workers} workers"
                               if workers and args.case_group > 1 else
                               f"km_simple_step per case, {workers} at a time" if workers else "km_batch_step"),
                      "fp32_vector_cases": n_fp32},
           "p_iters_min_median_max": [min(p_it), sorted(p_it)[len(p_it) // 2], max(p_it)],
           "p_cells_iters_per_s": round(cell_its / wall, 1),
           "p_iter_bytes_per_s": round(p_bytes / wall, 1) if coarse else None,
           "p_iter_hbm_fraction": round(p_bytes / wall / (HBM_PEAK_GBS * 1e9), 4) if coarse else None,
           "case_steps_per_s": round(args.steps * b.n / wall, 3)}
    print(json.dumps(out), flush=True)


def cpu_baseline(case, n_steps):
    """Time the C/OpenMP oracle on the same seeded start state: the first
    ``n_steps`` outer steps (all cores of this process's affinity mask)."""
    import numpy as np
    from karman_taichi_amd import cases as kc
    cores = len(os.sched_getaffinity(0))
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    from oracle import cpu as ocpu
    if not ocpu.LIB.exists():
        ocpu.build()
    m = kc.make_mask(case)
    st = ocpu.State(case, *kc.initial_fields(case, m), m)
    threads = ocpu.threads()
    its = []
    t0 = time.perf_counter()
    for k in range(n_steps):
        _, it = st.step()
        its.append(it)
        if (k + 1) % case.n_outer == 0:
            st.time_forward()
    dt = time.perf_counter() - t0
    cells = case.shape_p[0] * case.shape_p[1]
    p_it = sum(i[2] for i in its)
    log(f"cpu baseline: {n_steps} steps in {dt:.1f} s on {threads} threads")
    p_its = [i[2] for i in its]
    return {"value": round(n_steps / dt, 5), "unit": "SIMPLE steps/s", "cores": threads, "kind": "port",
            "sample": f"first {n_steps} outer steps ({n_steps // case.n_outer} time steps) of {case.name} from the "
                      f"seeded start, {dt:.1f} s, p' BiCGSTAB iterations {min(p_its)}..{max(p_its)}; "
                      f"C/OpenMP FP64 with the reference's kernel decomposition",
            "p_cells_iters_per_s": round(cells * p_it / dt, 1)}


if __name__ == "__main__":
    main()
