#!/usr/bin/env python3
"""Throughput of the per-time-step hot path on one MI355X, FP64 and mixed precision.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cfg 2] [--no-cpu]

A step is one RK2 time step (two projection stages: Smagorinsky eddy
viscosity, advection/diffusion fluxes, predictor, PPE source, BiCGSTAB
pressure solve, corrector, face-velocity correction) of the channel workload
on resident device state.  Two legs run the same steps from the same initial
state on one GPU:
  fp64   the reference's arithmetic (BiCGSTAB in FP64);
  mixed  the pressure solve as FP32 Jacobi-preconditioned conjugate gradients
         on the correction, with reliable updates (the true FP64 residual
         replaces the recursive one when it has fallen tenfold) and the
         reference's stopping rule on the true residual; --legs also takes
         mixed_cascade (an FP32 BiCGSTAB pass to 1e-5, then the FP64
         BiCGSTAB), mixed_restart and mixed_reliable (FP32 BiCGSTAB forms).
Each leg: W warm-up steps, K timed steps (wall clock, no events), then one
step with HIP events around every solver kernel for the per-kernel roofline.
Prints one JSON line:
  value        prism-cell-steps/s of the mixed leg = interior prisms x K / wall
  fp64         the same for the FP64 leg; mixed_over_fp64 their ratio
  rel_err      relative L2 error of u, v, w, p (interior) of the mixed leg
               against the FP64 leg after the same steps, and both legs' max
               |cell divergence| and CFL number (glob_cfl on the device)
  roofline     the dominant kernel of the mixed leg: algorithmic bytes per
               launch / mean launch time (HIP events on the library's stream)
  cpu_baseline the oracle (CPU restatement of the reference, 1 core) on a
               bounded sample of the same workload, extrapolated per step;
               cpu_baseline_all_cores the same with the oracle's solver on
               OMP_NUM_THREADS threads
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "3d-channel-flow-code_amd"))

METRIC = "prism-cell-steps/s: mixed-prec vs FP64 GPU at ≤1e-8 rel err; % HBM peak"
HBM_PEAK_GBS = 8000.0
# algorithmic HBM bytes per interior prism of each BiCGSTAB kernel, in units of
# the vector width (8 B FP64, 4 B FP32): A reads rt,q,rhat, writes vt; B
# reads rt,vt, writes t; C reads x,rt,q,vt,t,rhat, writes x,rt,q
PHASE_UNITS = {"a": 4, "b": 3, "c": 9}
KERNEL_NAMES = {"a": "k_bicg_a", "b": "k_bicg_b", "c": "k_bicg_c"}
# workload tag used in file names of the profiles: quads in x, quads in y, layers
TAG = {0: "32x16x32", 1: "128x64x64", 2: "256x128x96", 3: "64x32x48", 4: "512x256x128"}
WORKLOAD = {
    0: "32x16 quads x 32 layers = 32,768 prisms, Re_tau 180, dt 1e-3",
    1: "128x64 quads x 64 layers = 1,048,576 prisms, Re_tau 180, dt 5e-4, jittered mesh",
    2: "256x128 quads x 96 layers = 6,291,456 prisms, Re_tau 395, dt 2.5e-4, jittered mesh",
    3: "64x32 quads x 48 layers = 196,608 prisms, Re_tau 180, dt 5e-4, jittered mesh",
    4: "512x256 quads x 128 layers = 33,554,432 prisms, Re_tau 590, dt 1e-4, jittered mesh",
}


def cpu_baseline(case, mean_iters: float, cap: int = 20, parallel: bool = False) -> dict:
    """One projection stage of the oracle on the full mesh with BiCGSTAB capped
    at `cap` iterations; the step time is extrapolated with the GPU's mean
    iteration count per solve (FP64 leg).  parallel: the oracle built with
    OpenMP (the solver's loops over OMP_NUM_THREADS threads; the rest of the
    stage, 3 % of the serial step, stays on one thread)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle

    O = oracle.Oracle(case, parallel=parallel)
    cores = O.threads()
    O.set_max_iters(cap)
    for n in ("uf", "vf", "wf", "pf"):
        O.state(n)[:] = O.state(n[0] + "fn")
    t0 = time.perf_counter()
    O.hydro()
    wall = time.perf_counter() - t0
    t = O.timers()
    its = max(O.solver_record()["iter"], 1)
    t_iter = t["solve"] / its
    t_rest = t["hydro"] - t["solve"]
    step_s = case.params.FinRK * (t_rest + mean_iters * t_iter)
    what = "1 thread" if not parallel else f"OpenMP, {cores} threads on the solver"
    return {"value": case.n_interior / step_s, "unit": "prism-cell-steps/s", "cores": cores, "kind": "port",
            "sample": (f"oracle (C restatement of the reference, {what}) on the full mesh: one projection stage "
                       f"with BiCGSTAB capped at {its} iterations ({wall:.1f} s); step time = "
                       f"{case.params.FinRK} x (non-solver {t_rest:.2f} s + {mean_iters:.0f} iterations x "
                       f"{t_iter * 1e3:.1f} ms) with the FP64 GPU leg's mean iteration count"),
            "s_per_step": step_s}


# "mixed" is the library's default form: single-reduction FP32 CG with the
# 4-layer block-Jacobi preconditioner and reliable updates; "mixed_line" the
# same with the line preconditioner (whole columns), "mixed_point" with point
# Jacobi; "mixed_cg2" the point-Jacobi CG in two kernels per iteration
MIXED_MODES = {"mixed": (5, 0.1), "mixed_line": (6, 0.1), "mixed_point": (4, 0.1), "mixed_cg2": (3, 0.1), "mixed_cascade": (2, 1e-5),
               "mixed_restart": (1, 1e-5), "mixed_reliable": (0, 0.1)}
# the FP32 CG kernels: W reads rt, p and writes w; X reads e, rt, p, w and
# writes e, rt, p (units of 4 B per interior prism; the few X launches that
# form a true FP64 residual move more)
CG_UNITS = {"w": 3, "x": 7}


def leg(case, name: str, steps: int, warmup: int, max_iters=None) -> dict:
    from nsx.device import Context

    precision = 0 if name == "fp64" else 1
    ctx = Context(case)
    if max_iters:
        ctx.set_max_iters(max_iters)
    if precision:
        ctx.set_precision(1, *MIXED_MODES[name])
    ctx.init_state()
    if warmup:
        ctx.step(warmup, precision)
    info0 = ctx.solve_info()
    t0 = time.perf_counter()
    info = ctx.step(steps, precision)
    wall = time.perf_counter() - t0
    solves = info["n_solves"] - info0["n_solves"]
    iters = info["iters_total"] - info0["iters_total"]
    # per-kernel times: one more step with HIP events around each solver kernel
    ctx.timing(True)
    ctx.timing_reset()
    ctx.step(1, precision)
    ctx.timing(False)
    n = case.n_interior
    kernels = {}
    for width, prefix in ((8, "bicg_"), (4, "bicg32_")):
        for ph, units in PHASE_UNITS.items():
            ms, launches = ctx.timer(prefix + ph)
            if launches:
                avg = ms / launches
                b = units * width * n
                kernels[prefix + ph] = {"kernel": KERNEL_NAMES[ph] + ("<float>" if width == 4 else "<double>"),
                                        "launches": launches, "avg_us": avg * 1e3, "bytes_per_launch": b,
                                        "GBps": b / (avg * 1e-3) / 1e9}
    for ph, units in CG_UNITS.items():
        ms, launches = ctx.timer("cg32_" + ph)
        if launches:
            avg = ms / launches
            b = units * 4 * n
            kernels["cg32_" + ph] = {"kernel": "k_cg_" + ph, "launches": launches, "avg_us": avg * 1e3,
                                     "bytes_per_launch": b, "GBps": b / (avg * 1e-3) / 1e9}
    # the single-reduction CG: one kernel per iteration reading and writing
    # rt, om, sg (6 units), and every other iteration also reading the last
    # u, p and e and writing p and e (11 units): 8.5 units a launch on
    # average; its rare true-residual and product launches are counted with it
    ms, launches = ctx.timer("cg32_sr")
    if launches:
        avg = ms / launches
        b = 34 * n
        kname = "k_cgs<%d>" % {4: 0, 5: 1, 6: 2}[MIXED_MODES[name][0]]
        kernels["cg32_sr"] = {"kernel": kname, "launches": launches, "avg_us": avg * 1e3, "bytes_per_launch": b,
                              "GBps": b / (avg * 1e-3) / 1e9}
    stage_ms, stage_n = ctx.timer("stage")
    fields = {f: np.asarray(ctx.get(f))[:case.N_CELL0, 1:-1].copy() for f in ("UFN", "VFN", "WFN", "PFN")}
    # glob_cfl on the device: the largest net cell flux (cfl.F90:131-167) and the CFL number
    cfl = ctx.glob_cfl()
    fields["div"] = abs(cfl["flux_max"])
    fields["cfl"] = cfl["cfl"]
    ctx.close()
    return {"value": n * steps / wall, "ms_per_step": wall / steps * 1e3, "wall": wall,
            "mean_iters_per_solve": iters / max(solves, 1), "solves": solves, "exit": info["exit"],
            "last_iters_fp32": info.get("iters_fp32", 0),
            "last_e1": info["e1"], "last_e2": info["e2"], "outer_passes_last": info.get("outer_passes", 0),
            "kernels": kernels, "non_solver_ms_per_stage": stage_ms / max(stage_n, 1) * 2 if stage_n else None,
            "fields": fields}


def traffic_of(cfg: int, kernel):
    """HBM bytes per launch of `kernel` (fetch + write) from the committed
    rocprofv3 --pmc summary of this workload, or None"""
    if not kernel:
        return None
    from nsx.profsum import read_traffic

    t = read_traffic(os.path.join(REPO, "profiles", f"bicgstab_traffic_channel_{TAG[cfg]}.txt"))
    if kernel not in t:
        return None
    return t[kernel][0] + t[kernel][1]


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cfg", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--legs", default="fp64,mixed",
                    help="comma list of fp64, mixed, mixed_cascade, mixed_restart, mixed_reliable")
    ap.add_argument("--max-iters", type=int, default=None,
                    help="cap BiCGSTAB iterations (profiling runs only: the line is then marked capped)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first copy one cell field with k_copy (the known-size copy for HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding of the path is a design note, not code)", file=sys.stderr)
        return 2

    from nsx.case import config_case

    cfg = args.cfg
    steps = args.steps if args.steps is not None else {0: 10, 1: 20, 2: 10, 3: 20, 4: 3}[cfg]
    case = config_case(cfg, seed=args.seed)
    legs = [x.strip() for x in args.legs.split(",") if x.strip()]
    if args.calibrate_copy:
        from nsx.device import Context
        c0 = Context(case)
        c0.copy(64, 0)  # CELL_SCRATCH <- UFN: N_CELL0 x NS doubles, NS = NZ rounded up to a multiple of 4
        c0.close()
        ns = (case.NZ + 3) // 4 * 4
        print(f"# k_copy calibration: {case.N_CELL0 * ns} doubles", file=sys.stderr, flush=True)
    res = {}
    for name in legs:
        res[name] = leg(case, name, steps, args.warmup, args.max_iters)
        print(f"# leg {name}: {res[name]['value']:.4g} prism-cell-steps/s, {res[name]['ms_per_step']:.1f} ms/step, "
              f"{res[name]['mean_iters_per_solve']:.1f} iterations/solve", file=sys.stderr, flush=True)
    head = res.get("mixed") or res.get("fp64") or next(iter(res.values()))
    kern = head["kernels"]
    dom_name = max(kern, key=lambda k: kern[k]["avg_us"] * kern[k]["launches"]) if kern else None
    dom = kern.get(dom_name, {"GBps": 0.0, "bytes_per_launch": 0, "avg_us": 0.0, "kernel": None})
    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "prism-cell-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f32+f64 (block-Jacobi single-reduction FP32 CG in FP64 reliable-update refinement)" if "mixed" in res
                  else ("f64" if "fp64" in res else "f32+f64")),
        "data": "synthetic",
        "config": {"workload": f"channel {WORKLOAD[cfg]}", "prisms": case.n_interior,
                   "time_scheme": "RK2 (FinRK=2)", "les": "Smagorinsky Cs=0.1", "ppe_tolerance": case.params.eps,
                   "reliable_update_delta": MIXED_MODES["mixed"][1], "seed": args.seed},
        "roofline": {"bound": "hbm", "kernel": dom["kernel"], "achieved": dom["GBps"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": dom["GBps"] / HBM_PEAK_GBS, "traffic": traffic_of(cfg, dom["kernel"]),
                     "avg_launch_us": dom["avg_us"], "bytes_per_launch": dom["bytes_per_launch"]},
    }
    for name, r in res.items():
        out[name] = {k: v for k, v in r.items() if k != "fields"}
    if args.max_iters:
        out["capped_max_iters"] = args.max_iters
    for name in res:
        if "fp64" not in res or not name.startswith("mixed"):
            continue
        sfx = "" if name == "mixed" else name[5:]
        out["mixed_over_fp64" + sfx] = res[name]["value"] / res["fp64"]["value"]
        fa, fb = res["fp64"]["fields"], res[name]["fields"]
        err = {}
        for f, nm in (("UFN", "u"), ("VFN", "v"), ("WFN", "w"), ("PFN", "p")):
            err[nm] = float(np.linalg.norm(fb[f] - fa[f]) / max(np.linalg.norm(fa[f]), 1e-300))
        err["max_div_fp64"] = fa["div"]
        err["max_div_mixed"] = fb["div"]
        err["cfl_fp64"] = fa["cfl"]
        err["cfl_mixed"] = fb["cfl"]
        out["rel_err" + sfx] = err
    if not args.no_cpu:
        base = res.get("fp64", head)
        cb = cpu_baseline(case, base["mean_iters_per_solve"])
        out["cpu_baseline"] = {k: v for k, v in cb.items() if k != "s_per_step"}
        out["gpu_over_cpu"] = out["value"] / cb["value"]
        cbp = cpu_baseline(case, base["mean_iters_per_solve"], parallel=True)
        out["cpu_baseline_all_cores"] = {k: v for k, v in cbp.items() if k != "s_per_step"}
        out["gpu_over_cpu_all_cores"] = out["value"] / cbp["value"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
