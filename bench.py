#!/usr/bin/env python3
"""Throughput of the D3Q19 per-step path on one MI355X.

Workload (the default): 256x128x128 channel, D3Q19 TRT (Lambda = 3/16), FP32
populations; velocity inlet u = 0.05 at x = 0, zero-gradient outlet, no-slip
y/z walls; one sphere of radius 12 translating at (0.01, 0, 0) with moving
bounce-back, momentum correction and equilibrium refill of uncovered nodes;
momentum-exchange force and torque every step.  tau+ = 0.56 (Re_D = 60 on the
inlet speed).  A step is one full time step of that case; the timed region
starts with the populations resident in HBM.

Prints one JSON line.  MLUPS counts every cell, solid ones included
(src/Utilities.jl:913-916).  ``roofline`` is the fused stream-collide kernel's
algorithmic bytes per launch of the per-step kernel k_step (2 x 19 x 4 B per
cell, all cells: the launch covers bulk rows, faces and body regions) over its
mean launch time from HIP events recorded on the solver's stream around every
16th launch in the timed region.  ``cpu_baseline`` times the oracle (a plain C restatement,
OpenMP) on the same workload for a bounded number of steps.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "channel256": dict(nx=256, ny=128, nz=128, storage="fp32", dtype="f32"),
    "channel256_fp16s": dict(nx=256, ny=128, nz=128, storage="fp16s", dtype="f16 storage, f32 compute"),
    "channel256_bf16s": dict(nx=256, ny=128, nz=128, storage="bf16s", dtype="bf16 storage, f32 compute"),
}


def make_case(w):
    from lbm3d_bb_amd.bodies import ConstantMotion
    from lbm3d_bb_amd.solver import Config, no_slip, velocity, zero_gradient

    tau_p = 0.56
    tau_m = 0.5 + (3.0 / 16.0) / (tau_p - 0.5)
    cfg = Config(w["nx"], w["ny"], w["nz"], collision="TRT", storage=w["storage"], omega_plus=1.0 / tau_p,
                 omega_minus=1.0 / tau_m,
                 faces=[velocity((0.05, 0, 0)), zero_gradient(), no_slip(), no_slip(), no_slip(), no_slip()],
                 u0=(0.05, 0.0, 0.0), refill=True)
    law = ConstantMotion("sphere", (48.0, 64.0, 64.0), (12.0,), vel=(0.01, 0.0, 0.0))
    return cfg, law


def bytes_per_cell(storage: str) -> int:
    return 2 * 19 * {"fp64": 8, "fp32": 4, "fp16s": 2, "bf16s": 2}[storage]


def cpu_baseline(w, budget_s: float) -> dict:
    """The oracle on the same workload, timed for a bounded number of steps."""
    from oracle.oracle import OracleSolver, load_oracle

    cfg, law = make_case(w)
    prec = "fp64" if w["storage"] == "fp64" else "fp32"
    o = OracleSolver(cfg, prec)
    o.set_motion([law])
    o.init_equilibrium()
    o.step(1)  # warm-up (first geometry build)
    n = 0
    t0 = time.perf_counter()
    while True:
        o.step(1)
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s or n >= 500:
            break
    o.close()
    mlups = cfg.ncells * n / el / 1e6
    return {"value": round(mlups, 3), "unit": "MLUPS", "cores": int(load_oracle().orc_num_threads()),
            "kind": "port",
            "sample": f"{n} steps of the same {cfg.nx}x{cfg.ny}x{cfg.nz} case after 1 warm-up step, "
                      f"{prec} oracle, {el:.1f} s"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=500)
    ap.add_argument("--workload", default="channel256", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of oracle time for cpu_baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true", help="skip the per-launch HIP events (no roofline)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (asked for {args.gpus}): this build has no multi-GPU path; "
              "the sharding design is a slab decomposition along x, described in DESIGN.md.", file=sys.stderr)
        return 2

    from lbm3d_bb_amd.solver import Solver

    w = WORKLOADS[args.workload]
    cfg, law = make_case(w)
    s = Solver(cfg)
    s.set_motion([law])
    s.init_equilibrium()
    if args.warmup:
        s.step(args.warmup)
    s.sync()
    s.profile(not args.no_profile)
    t0 = time.perf_counter()
    forces = s.step(args.steps)
    s.sync()
    el = time.perf_counter() - t0
    prof = s.profile_read()
    s.profile(False)
    info = s.info()
    rho, u = s.macros()
    s.close()

    N = cfg.ncells
    mlups = N * args.steps / el / 1e6
    bpc = bytes_per_cell(w["storage"])
    # k_step is the whole time step (bulk rows, faces, body regions) in one launch
    step_avg_s = prof["bulk_ms"] / max(prof["bulk_launches"], 1) / 1e3
    achieved = bpc * N / step_avg_s / 1e9 if step_avg_s > 0 else None
    traffic = None
    tfile = os.path.join(REPO, "profiles", f"traffic_{args.workload}.json")
    if os.path.exists(tfile):
        with open(tfile, encoding="utf-8") as fh:
            traffic = json.load(fh).get("bytes_per_launch")
    out = {
        "metric": "MLUPS",
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": w["dtype"],
        "data": "synthetic (uniform start, prescribed body motion; no dataset)",
        "config": {
            "workload": f"{args.workload}: {cfg.nx}x{cfg.ny}x{cfg.nz} D3Q19 TRT channel, {w['storage']} populations, "
                        "translating sphere R=12 with moving bounce-back, refill, per-step MEA force",
            "cells": N,
            "tau_plus": 0.56,
            "lambda": 0.1875,
            "step_hbm_frac_of_peak": round(bpc * N * args.steps / el / 1e9 / HBM_PEAK_GBS, 4),
            "step_kernel_ms": round(step_avg_s * 1e3, 5),
            "step_kernel_samples": prof["bulk_launches"],
            "boundary_nodes_last_static_build": info["nbb_nodes"],
            "drag_last_step": float(forces[-1, 0, 0]) if forces is not None else None,
            "rho_min": float(rho.min()),
            "u_max": float(np.abs(u).max()),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1) if achieved else None,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
            "traffic": traffic,
        },
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_budget)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
