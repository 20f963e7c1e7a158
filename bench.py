#!/usr/bin/env python3
"""Throughput of the per-step path (D3Q19, or D3Q27) on one MI355X.

Workload (the default): 256x128x128 channel, D3Q19 TRT (Lambda = 3/16), FP32
populations; velocity inlet u = 0.05 at x = 0, zero-gradient outlet, no-slip
y/z walls; one sphere of radius 12 translating at (0.01, 0, 0) with moving
bounce-back, momentum correction and equilibrium refill of uncovered nodes;
momentum-exchange force and torque every step.  tau+ = 0.56 (Re_D = 60 on the
inlet speed).  A step is one full time step of that case; the timed region
starts with the populations resident in HBM.

Prints one JSON line.  MLUPS counts every cell, solid ones included
(src/Utilities.jl:913-916).  ``roofline`` is the per-step kernel k_step (one
launch per time step: bulk rows, faces and body regions; k27_step for the
D3Q27 workloads): its algorithmic bytes per launch (2 x q x s bytes per cell,
q = 19 or 27 populations, s the storage size, all cells)
over its mean launch time from HIP events recorded on the solver's stream
around every 4th launch in the timed region (``--sample-stride``); ``traffic`` is the HBM bytes per
launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE, when
profiles/traffic_<workload>.json holds it.  ``cpu_baseline`` times the oracle (a
plain C restatement, OpenMP) on the same workload for a bounded number of
steps.  ``config.drag_err_vs_fp64`` compares each body's drag, averaged over
the last 100 steps of the run, with an untimed FP64-storage run of the same
workload, start and step count.  ``companion`` (default workload only) times
the same case in shifted-FP16 storage right after, with the same warm-up and
steps, and reports its MLUPS, its ratio to the FP32 line, its roofline and its
drag error against the same FP64 run.

--copy-probe MiB runs, after the timed region, ten device-to-device copies of
that many MiB (a plain copy kernel of known size, the calibration of the
rocprofv3 byte counters).
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

_DT = {"fp64": "f64", "fp32": "f32", "fp16s": "f16 storage, f32 compute", "bf16s": "bf16 storage, f32 compute"}

# channel256*: the translating-sphere channel above (the default, the metric's
# configuration); box512_spheres64*: 512x256x256 periodic box driven by a body
# force g = 1e-6 along x (the reference-runnable 64^3 case's value; at 1e-7 the
# per-step forcing increment 3 w_d g of a population is within a few half-ulps
# of w_d in FP32) with 64 fixed spheres (radius uniform in [6, 10], centres by rejection
# sampling from seed 0); channel1024_ellipsoids8: 1024x384x384 channel in
# shifted FP16 with 8 ellipsoids (16, 8, 8) at random orientations, each
# rotating at 1e-3 rad/step about a random axis; refcase_*: the reference's
# own input file (input/params.txt: 1201x1201x3, D3Q19 BGK at Re 185, velocity
# inlet and y walls at U = 0.075, anti-bounce-back outlet, z periodic, a
# cylinder R = 16.16 heaving with the reference's law) with its interpolated
# bounce-back (IBB), the simple one (SBB), or the NACA 0012 airfoil of chord 120
# (IBB); channel256_q27*: the translating-sphere channel on the D3Q27 lattice;
# refcase_cyl_ibb_q27: the reference's case with q = 27 in its input file.
WORKLOADS = {
    "channel256": dict(case="channel", nx=256, ny=128, nz=128, storage="fp32"),
    "channel256_fp16s": dict(case="channel", nx=256, ny=128, nz=128, storage="fp16s"),
    "channel256_bf16s": dict(case="channel", nx=256, ny=128, nz=128, storage="bf16s"),
    "channel256_q27": dict(case="channel", nx=256, ny=128, nz=128, storage="fp32", q=27),
    "channel256_q27_fp64": dict(case="channel", nx=256, ny=128, nz=128, storage="fp64", q=27),
    "channel256_q27_fp16s": dict(case="channel", nx=256, ny=128, nz=128, storage="fp16s", q=27),
    "channel256_q27_bf16s": dict(case="channel", nx=256, ny=128, nz=128, storage="bf16s", q=27),
    "box512_spheres64": dict(case="box", nx=512, ny=256, nz=256, storage="fp32"),
    "box512_spheres64_fp16s": dict(case="box", nx=512, ny=256, nz=256, storage="fp16s"),
    "box512_spheres64_fp64": dict(case="box", nx=512, ny=256, nz=256, storage="fp64"),
    "channel1024_ellipsoids8": dict(case="ellipsoids", nx=1024, ny=384, nz=384, storage="fp16s"),
    "refcase_cyl_ibb": dict(case="refparams", shape="CYL", bc="IBB", nx=1201, ny=1201, nz=3, storage="fp32"),
    "refcase_cyl_sbb": dict(case="refparams", shape="CYL", bc="SBB", nx=1201, ny=1201, nz=3, storage="fp32"),
    "refcase_naca_ibb": dict(case="refparams", shape="AIR", bc="IBB", nx=1201, ny=1201, nz=3, storage="fp32"),
    "refcase_cyl_ibb_q27": dict(case="refparams", shape="CYL", bc="IBB", nx=1201, ny=1201, nz=3, storage="fp32", q=27),
}
for _w in WORKLOADS.values():
    _w["dtype"] = _DT[_w["storage"]]
    _w.setdefault("q", 19)


def _trt(tau_p: float) -> tuple[float, float]:
    return 1.0 / tau_p, 1.0 / (0.5 + (3.0 / 16.0) / (tau_p - 0.5))


def make_case(w):
    """(Config, motion laws or None, fixed bodies or None, description) of a workload."""
    from lbm3d_bb_amd.bodies import ConstantMotion, rotation_matrix, sphere
    from lbm3d_bb_amd.solver import Config, no_slip, periodic, velocity, zero_gradient

    walls = [velocity((0.05, 0, 0)), zero_gradient(), no_slip(), no_slip(), no_slip(), no_slip()]
    if w["case"] == "refparams":
        from lbm3d_bb_amd.params import LBMParams, config_from_params

        p = LBMParams(Nx=w["nx"], Ny=w["ny"], Nz=w["nz"], h=1.0, R=16.16, x_c=600.5, y_c=600.5, theta0=0.0, q=w["q"],
                      Re=185.0, U_inf=0.075, V_inf=0.0, W_inf=0.0, rho_inf=1.0, dt=1.0, Lambda=0.1875,
                      n_max=40000, n_save=10000, bc=w["bc"], coll="BGK", shape=w["shape"], naca_code="0012",
                      chord=120.0)
        cfg, laws = config_from_params(p, storage=w["storage"])
        body = "cylinder R=16.16" if w["shape"] == "CYL" else "NACA 0012 airfoil, chord 120"
        desc = (f"{cfg.nx}x{cfg.ny}x{cfg.nz} D3Q{w['q']} BGK, the reference's input/params.txt case: {body} heaving "
                f"with the reference's law, {w['bc']} links, {w['storage']} populations, per-step MEA force")
        return cfg, laws, None, desc
    if w["case"] == "channel":
        wp, wm = _trt(0.56)
        cfg = Config(w["nx"], w["ny"], w["nz"], collision="TRT", storage=w["storage"], omega_plus=wp, omega_minus=wm,
                     faces=walls, u0=(0.05, 0.0, 0.0), refill=True, q=w["q"])
        law = ConstantMotion("sphere", (48.0, 64.0, 64.0), (12.0,), vel=(0.01, 0.0, 0.0))
        desc = (f"{cfg.nx}x{cfg.ny}x{cfg.nz} D3Q{w['q']} TRT channel, {w['storage']} populations, translating sphere "
                "R=12 with moving bounce-back, refill, per-step MEA force")
        return cfg, [law], None, desc
    rng = np.random.default_rng(0)
    L = np.array([w["nx"], w["ny"], w["nz"]], dtype=float)
    if w["case"] == "box":
        wp, wm = _trt(0.6)
        cfg = Config(w["nx"], w["ny"], w["nz"], collision="TRT", storage=w["storage"], omega_plus=wp, omega_minus=wm,
                     faces=[periodic()] * 6, g=(1e-6, 0.0, 0.0))
        placed = []
        while len(placed) < 64:
            r = rng.uniform(6.0, 10.0)
            c = rng.uniform(0.0, 1.0, size=3) * L
            ok = all(np.linalg.norm(np.minimum(np.abs(c - c2), L - np.abs(c - c2))) >= r + r2 + 2.0
                     for c2, r2 in placed)
            if ok:
                placed.append((c, r))
        bodies = [sphere(tuple(c), r) for c, r in placed]
        desc = (f"{cfg.nx}x{cfg.ny}x{cfg.nz} periodic D3Q19 TRT box, body force, {w['storage']} populations, "
                "64 fixed spheres r in [6,10] (seed 0), per-step MEA force")
        return cfg, None, bodies, desc
    wp, wm = _trt(0.56)
    cfg = Config(w["nx"], w["ny"], w["nz"], collision="TRT", storage=w["storage"], omega_plus=wp, omega_minus=wm,
                 faces=walls, u0=(0.05, 0.0, 0.0), refill=True)
    laws, placed = [], []
    while len(laws) < 8:
        c = np.array([rng.uniform(64.0, L[0] - 64.0), rng.uniform(24.0, L[1] - 24.0), rng.uniform(24.0, L[2] - 24.0)])
        if any(np.linalg.norm(c - c2) < 36.0 for c2 in placed):
            continue
        placed.append(c)
        axis0, axis1 = rng.standard_normal(3), rng.standard_normal(3)
        rot0 = rotation_matrix(axis0, rng.uniform(0.0, 2.0 * np.pi))
        om = 1e-3 * axis1 / np.linalg.norm(axis1)
        laws.append(ConstantMotion("ellipsoid", tuple(c), (16.0, 8.0, 8.0), omega=tuple(om), rot0=rot0))
    desc = (f"{cfg.nx}x{cfg.ny}x{cfg.nz} D3Q19 TRT channel, {w['storage']} populations, 8 ellipsoids (16,8,8) "
            "rotating at 1e-3 rad/step, moving bounce-back, refill, per-step MEA force and torque")
    return cfg, laws, None, desc


def setup(solver, laws, bodies) -> None:
    if laws is not None:
        solver.set_motion(laws)
    else:
        solver.set_bodies(bodies)
    solver.init_equilibrium()


def bytes_per_cell(storage: str, q: int = 19) -> int:
    return 2 * q * {"fp64": 8, "fp32": 4, "fp16s": 2, "bf16s": 2}[storage]


def cpu_baseline(w, budget_s: float) -> dict:
    """The oracle on the same workload, timed for a bounded number of steps."""
    from oracle.oracle import OracleSolver, load_oracle

    cfg, laws, bodies, _ = make_case(w)
    prec = "fp64" if w["storage"] == "fp64" else "fp32"
    o = OracleSolver(cfg, prec)
    setup(o, laws, bodies)
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


# The 16-bit storage of the default workload, timed in the same run and
# reported beside it (the FP16S throughput claim and its drag error).
COMPANIONS = {"channel256": "channel256_fp16s"}


def timed_run(w, args) -> dict:
    """Set-up, warm-up and the timed region of one workload; the per-launch
    HIP-event samples of k_step come back in `prof`."""
    from lbm3d_bb_amd.solver import Solver

    cfg, laws, bodies, desc = make_case(w)
    s = Solver(cfg)
    setup(s, laws, bodies)
    start = {"start": "uniform", "reason": "uniform start asked"}
    if args.start == "net":
        start = net_start(s, cfg, laws, bodies, args.params)
    warm_forces = s.step(args.warmup) if args.warmup else None
    s.sync()
    s.profile(0 if args.no_profile else max(1, args.sample_stride))
    t0 = time.perf_counter()
    forces = s.step(args.steps)
    s.sync()
    el = time.perf_counter() - t0
    prof = s.profile_read()
    s.profile(0)
    info = s.info()
    rho, u = s.macros()
    s.close()
    return {"case": (cfg, laws, bodies, desc), "forces": forces, "warm_forces": warm_forces, "seconds": el,
            "prof": prof, "info": info, "start": start, "rho": rho, "u": u}


def roofline(w, name, ncells, prof) -> tuple[float, dict]:
    """Mean k_step launch time (s) and the roofline object: k_step is the whole
    time step (bulk rows, faces and body regions) in one launch."""
    bpc = bytes_per_cell(w["storage"], w["q"])
    step_avg_s = prof["bulk_ms"] / max(prof["bulk_launches"], 1) / 1e3
    achieved = bpc * ncells / step_avg_s / 1e9 if step_avg_s > 0 else None
    traffic = None
    tfile = os.path.join(REPO, "profiles", f"traffic_{name}.json")
    if os.path.exists(tfile):
        with open(tfile, encoding="utf-8") as fh:
            traffic = json.load(fh).get("bytes_per_launch")
    return step_avg_s, {
        "bound": "hbm",
        "achieved": round(achieved, 1) if achieved else None,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
        "traffic": traffic,
    }


def _case_key(w) -> dict:
    return {k: v for k, v in w.items() if k not in ("storage", "dtype")}


def companion(name, args, primary, primary_mlups, F64) -> dict:
    """Another workload with the same start, warm-up and steps: its MLUPS,
    ratio to the primary line's, roofline and drag error against FP64 (the
    primary's FP64 run when the companion is the same case in another
    storage, else an FP64 run of its own)."""
    if name not in WORKLOADS:
        raise SystemExit(f"--companion: unknown workload {name!r} (one of {', '.join(sorted(WORKLOADS))}, or none)")
    w = WORKLOADS[name]
    run = timed_run(w, args)
    cfg = run["case"][0]
    mlups = cfg.ncells * args.steps / run["seconds"] / 1e6
    step_avg_s, roof = roofline(w, name, cfg.ncells, run["prof"])
    out = {"workload": name, "value": round(mlups, 1), "unit": "MLUPS", "dtype": w["dtype"],
           "ms_per_step": round(run["seconds"] / args.steps * 1e3, 5), "ratio_to_primary": round(mlups /
           primary_mlups, 3), "step_kernel_ms": round(step_avg_s * 1e3, 5),
           "step_kernel_samples": run["prof"]["bulk_launches"], "roofline": roof}
    if w["storage"] in ("fp16s", "bf16s") and w["q"] == 19:
        # the D3Q19 16-bit target: 1.7 x the MLUPS of FP32 storage (2 x 19 x 4 bytes per cell) at 60 % of
        # the HBM roofline
        target = 1.7 * 0.6 * HBM_PEAK_GBS * 1e9 / bytes_per_cell("fp32", 19) / 1e6
        out["target"] = {"value": round(target, 1), "met": bool(mlups >= target),
                         "rule": "1.7 x the FP32-storage MLUPS at 60 % of the 8 TB/s roofline"}
    if not args.no_fp64_ref:
        F = run["forces"] if run["warm_forces"] is None else np.concatenate([run["warm_forces"], run["forces"]])
        if F64 is None or _case_key(w) != _case_key(primary):
            F64 = fp64_forces(w, F.shape[0], args, run["case"][1], run["case"][2])
        out["drag_err_vs_fp64"] = drag_err(F, F64)
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=500)
    ap.add_argument("--workload", default="channel256", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of oracle time for cpu_baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true", help="skip the per-launch HIP events (no roofline)")
    ap.add_argument("--sample-stride", type=int, default=4,
                    help="time every n-th step launch with HIP events (an event pair costs ~3 us of the step)")
    ap.add_argument("--start", default="uniform", choices=["uniform", "net"],
                    help="net: start from the 128^3 U-Net applied tile-wise (untimed, before the warm-up)")
    ap.add_argument("--params", default=None, help="U-Net parameters (.npy) for --start net; random init if absent")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MiB",
                    help="after the run, ten device-to-device copies of this many MiB (counter calibration)")
    ap.add_argument("--companion", default="auto",
                    help="another workload timed after the primary and reported beside it ('auto': the 16-bit "
                         "storage of the default workload; 'none' for no companion)")
    ap.add_argument("--no-fp64-ref", action="store_true",
                    help="skip the FP64 run of the same steps that drag_err_vs_fp64 compares against")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (asked for {args.gpus}): there is no multi-GPU path; the path "
              "would shard as x slabs, one per GPU, exchanging the five x-moving populations of each face plane "
              "every step.", file=sys.stderr)
        return 2

    if args.copy_probe > 0:
        import torch  # initialised before the library so the probe's copies share its HIP runtime

        torch.cuda.init()
    w = WORKLOADS[args.workload]
    run = timed_run(w, args)
    cfg, laws, bodies, desc = run["case"]
    forces, warm_forces, el, prof, info, start = (run[k] for k in ("forces", "warm_forces", "seconds", "prof",
                                                                   "info", "start"))
    rho, u = run["rho"], run["u"]

    N = cfg.ncells
    mlups = N * args.steps / el / 1e6
    bpc = bytes_per_cell(w["storage"], w["q"])
    step_avg_s, roof = roofline(w, args.workload, N, prof)
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
        "data": f"synthetic ({start['start']} start, prescribed body motion; no dataset)",
        "config": {
            "workload": f"{args.workload}: {desc}",
            "cells": N,
            "tau_plus": round(1.0 / cfg.omega_plus, 6),
            "lambda": 0.1875 if cfg.collision == "TRT" else None,
            "solid_bc": cfg.solid_bc,
            "lattice": f"D3Q{w['q']}",
            "step_hbm_frac_of_peak": round(bpc * N * args.steps / el / 1e9 / HBM_PEAK_GBS, 4),
            "step_kernel_ms": round(step_avg_s * 1e3, 5),
            "step_kernel_samples": prof["bulk_launches"],
            "boundary_nodes_last_static_build": info["nbb_nodes"],
            "drag_last_step": float(forces[-1, 0, 0]) if forces is not None else None,
            "rho_min": float(rho.min()),
            "u_max": float(np.abs(u).max()),
        },
        "roofline": roof,
    }
    if args.start == "net":
        out["start"] = start
    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
    F64 = None
    if not args.no_fp64_ref and forces is not None:
        F = forces if warm_forces is None else np.concatenate([warm_forces, forces])
        F64 = fp64_forces(w, F.shape[0], args, laws, bodies)
        out["config"]["drag_err_vs_fp64"] = drag_err(F, F64)
    comp = COMPANIONS.get(args.workload) if args.companion == "auto" else (args.companion or None)
    if comp and comp != "none":
        out["companion"] = companion(comp, args, w, mlups, F64)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_budget)
    print(json.dumps(out), flush=True)
    return 0


def fp64_forces(w, n, args, laws, bodies) -> np.ndarray | None:
    """Per-step forces of the same workload, start and step count in FP64
    storage (untimed, after the timed run); None when the run is FP64 itself."""
    from lbm3d_bb_amd.solver import Solver

    if w["storage"] == "fp64":
        return None
    cfg64 = make_case(dict(w, storage="fp64"))[0]
    r = Solver(cfg64)
    setup(r, laws, bodies)
    if args.start == "net":
        net_start(r, cfg64, laws, bodies, args.params)
    F64 = r.step(n)
    r.close()
    return F64


def drag_err(F, F64) -> dict:
    """Relative error of each body's drag (x-force) averaged over the last
    min(100, steps) steps, against the FP64 run's."""
    n = F.shape[0]
    win = min(100, n)
    if F64 is None:
        return {"max": 0.0, "total": 0.0, "steps": n, "window": win, "note": "this run is FP64"}
    d = F[-win:, :, 0].mean(axis=0)
    d64 = F64[-win:, :, 0].mean(axis=0)
    e = np.abs(d - d64) / np.abs(d64)
    return {"max": float(e.max()), "median": float(np.median(e)), "total": float(abs(d.sum() - d64.sum()) /
            abs(d64.sum())), "steps": n, "window": win}


def net_start(s, cfg, laws, bodies, params) -> dict:
    """The pack network (4 levels, 16/32/64/128 channels, warmstart.PACK_NET) on 128^3 tiles (halo 16) from
    the solver's solid flags; Re = |u0| D / nu with D the smallest diameter of
    the first body.  Returns the start taken and the inference time."""
    from lbm3d_bb_amd.warmstart import pack_net, warm_start_tiled

    u_ref = float(np.linalg.norm(cfg.u0))
    if u_ref == 0.0:
        raise SystemExit("--start net needs a workload with an inflow velocity")
    semi = (laws[0] if laws is not None else bodies[0]).semi
    nu = (1.0 / cfg.omega_plus - 0.5) / 3.0
    re = u_ref * 2.0 * min(semi) / nu
    tile = tuple(min(128, n) for n in (cfg.nz, cfg.ny, cfg.nx))
    net = pack_net(tile, batch=2, seed=0)
    try:
        if params:
            net.load(params)
        info = warm_start_tiled(s, net, re, u_ref, halo=16)
    finally:
        net.close()
    info.update(re=round(re, 3), tile=list(tile), halo=16, params=params or "random init")
    info["seconds"] = round(info["seconds"], 3)
    return info


def copy_probe(mib: int) -> None:
    """Ten device-to-device copies of `mib` MiB: the plain copy kernel whose
    rocprofv3 FETCH_SIZE / WRITE_SIZE calibrate the byte counters."""
    import torch

    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(10):
        b.copy_(a)
    torch.cuda.synchronize()
    del a, b


if __name__ == "__main__":
    sys.exit(main())
