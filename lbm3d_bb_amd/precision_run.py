"""Per-body drag and torque of one benchmark configuration in several storage
formats, against the FP64 run of the same configuration.

  python -m lbm3d_bb_amd.precision_run --workload box512_spheres64 --steps 50000 \\
      --storages fp64,fp32,fp16s,bf16s --out profiles/precision_box512_spheres64.json
  python -m lbm3d_bb_amd.precision_run combine uniform=A.json net=B.json fp64=C.json --out SUMMARY.json

Each storage runs the configuration from the same start for the same steps
(the FP64 run is the reference: its step is pinned to the oracle at the FP64
bar).  `--start net` starts every run from the tile-wise prediction of the
trained pack network instead of the uniform state (bench.net_start); the
output records the start taken.  The force of a body is the momentum-exchange sum of src/MEA.jl:76-116
(FP64 accumulation, with torque).  Reported per body: the mean x-force (drag)
and the mean torque vector over the last `--window` steps, and for every
non-FP64 storage the relative drag error |D - D64| / |D64| and the torque error
|T - T64| / max_bodies |T64|; and the history: per-body force and torque
averaged over every `--chunk` steps.  Prints a progress line every `--chunk`
steps and rewrites `--out` with the history so far (key "partial"), so a run
cut off by a time limit keeps what it did.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_storage(w: dict, storage: str, steps: int, window: int, chunk: int, log, start=None, partial=None) -> dict:
    """Steps the workload in one storage format; returns the window means.
    partial(record) is called after every chunk with the history so far."""
    sys.path.insert(0, REPO)
    import bench

    from .solver import Solver

    wk = dict(w, storage=storage)
    cfg, laws, bodies, desc = bench.make_case(wk)
    s = Solver(cfg)
    if laws is not None:
        s.set_motion(laws)
    else:
        s.set_bodies(bodies)
    info = {"start": "uniform"}
    if start is not None:
        info = start(s, cfg, laws, bodies)
    else:
        s.init_equilibrium()
    nb = len(laws) if laws is not None else len(bodies)
    tail = np.zeros((0, nb, 6))
    drag_hist = []
    hist = []
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        k = min(chunk, steps - done)
        F = s.step(k)
        done += k
        if not np.isfinite(F).all():
            raise FloatingPointError(f"{storage}: non-finite force by step {done}")
        tail = np.concatenate([tail, F])[-window:]
        drag_hist.append([done, F[-1, :, 0].sum()])
        hist.append([done, F.mean(axis=0).tolist()])
        el = time.perf_counter() - t0
        log(f"{storage}: step {done}/{steps}, {cfg.ncells * done / el / 1e6:.0f} MLUPS, "
            f"total drag {F[-1, :, 0].sum():.6e}")
        if partial is not None:
            partial({"storage": storage, "steps_done": done, "drag_total_history": drag_hist,
                     "chunk_mean_history": hist, "start": info, "workload": desc})
    el = time.perf_counter() - t0
    rho, _ = s.macros()
    s.close()
    return {"storage": storage, "steps": steps, "seconds": el, "mlups": cfg.ncells * steps / el / 1e6,
            "drag": tail[:, :, 0].mean(axis=0).tolist(), "force": tail[:, :, :3].mean(axis=0).tolist(),
            "torque": tail[:, :, 3:].mean(axis=0).tolist(), "rho_min": float(rho.min()),
            "rho_max": float(rho.max()), "drag_total_history": drag_hist, "chunk_mean_history": hist,
            "start": info, "workload": desc}


def compare(ref: dict, run: dict) -> dict:
    d64 = np.asarray(ref["drag"])
    d = np.asarray(run["drag"])
    e = np.abs(d - d64) / np.abs(d64)
    t64 = np.asarray(ref["torque"])
    t = np.asarray(run["torque"])
    et = np.linalg.norm(t - t64, axis=1) / max(float(np.linalg.norm(t64, axis=1).max()), 1e-300)
    return {"drag_err_per_body": e.tolist(), "drag_err_max": float(e.max()), "drag_err_median": float(np.median(e)),
            "drag_err_total": float(abs(d.sum() - d64.sum()) / abs(d64.sum())),
            "torque_err_per_body": et.tolist(), "torque_err_max": float(et.max())}


def combine(named: list[str], out: str) -> dict:
    """One summary of runs of the same workload written by separate calls
    (NAME=PATH, one storage per file): each run's throughput and window means,
    the errors of every non-FP64 run against the FP64 one (same start), and
    for every pair of runs of the same storage but different starts the
    history of the largest per-body difference of the chunk-mean drag,
    relative to the largest body drag of the chunk."""
    runs = {}
    for item in named:
        name, path = item.split("=", 1)
        with open(path, encoding="utf-8") as fh:
            d = json.load(fh)
        (st, r), = d["runs"].items()
        runs[name] = dict(r, storage=st, file=os.path.basename(path))
    summary = {"runs": {k: {kk: v for kk, v in r.items() if kk not in ("chunk_mean_history",
                                                                       "drag_total_history")}
                        for k, r in runs.items()}}
    ref = [k for k, r in runs.items() if r["storage"] == "fp64"]
    if ref:
        r64 = runs[ref[0]]
        summary["errors_vs_fp64"] = {k: compare(r64, r) for k, r in runs.items()
                                     if r["storage"] != "fp64" and r["start"]["start"] == r64["start"]["start"]}
    pairs = {}
    names = list(runs)
    for i, a in enumerate(names):
        for b in names[i + 1:]:
            ra, rb = runs[a], runs[b]
            if ra["storage"] != rb["storage"] or ra["start"]["start"] == rb["start"]["start"]:
                continue
            ha = np.array([h[1] for h in ra["chunk_mean_history"]])[:, :, 0]
            hb = np.array([h[1] for h in rb["chunk_mean_history"]])[:, :, 0]
            steps = [h[0] for h in ra["chunk_mean_history"]]
            rel = np.abs(ha - hb).max(axis=1) / np.abs(hb).max(axis=1)
            first = {}
            for tol in (1e-1, 1e-2, 3e-3, 2e-3):
                ok = np.nonzero(rel < tol)[0]
                # the first chunk from which every later chunk stays below tol
                later = [j for j in ok if (rel[j:] < tol).all()]
                first[f"{tol:g}"] = steps[later[0]] if later else None
            pairs[f"{a}_vs_{b}"] = {"chunk_steps": steps, "max_body_drag_rel_diff": rel.tolist(),
                                    "steps_until_within": first, "final_rel_diff": float(rel[-1])}
    summary["start_pairs"] = pairs
    with open(out, "w", encoding="utf-8") as fh:
        json.dump(summary, fh, indent=1)
    return summary


def main(argv=None) -> int:
    sys.path.insert(0, REPO)
    import bench

    if argv is None:
        argv = sys.argv[1:]
    if argv and argv[0] == "combine":
        # precision_run combine NAME=PATH ... --out SUMMARY.json
        k = argv.index("--out")
        s = combine(argv[1:k], argv[k + 1])
        print(json.dumps({"errors_vs_fp64": s.get("errors_vs_fp64"),
                          "start_pairs": {p: v["steps_until_within"] for p, v in s["start_pairs"].items()}}))
        return 0

    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="box512_spheres64", choices=sorted(bench.WORKLOADS))
    ap.add_argument("--steps", type=int, default=50000)
    ap.add_argument("--window", type=int, default=1000, help="steps averaged at the end")
    ap.add_argument("--chunk", type=int, default=2000, help="steps between progress lines")
    ap.add_argument("--storages", default="fp64,fp32,fp16s,bf16s")
    ap.add_argument("--ref", default=None, help="JSON of an earlier FP64 run of the same workload and steps")
    ap.add_argument("--start", default="uniform", choices=["uniform", "net"],
                    help="net: the trained pack network's tile-wise start (lbm3d_bb_amd/data/unet_pack128.npy)")
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)

    def log(m):
        print(m, flush=True)

    w = bench.WORKLOADS[a.workload]
    runs = {}
    if a.ref:
        with open(a.ref, encoding="utf-8") as fh:
            prev = json.load(fh)
        runs["fp64"] = prev["runs"]["fp64"]
        if runs["fp64"]["steps"] != a.steps:
            raise SystemExit("the reference run has a different step count")
    start = None
    if a.start == "net":
        from .warmstart import PACK_PARAMS

        def start(s, cfg, laws, bodies):
            return bench.net_start(s, cfg, laws, bodies, PACK_PARAMS)
    for st in [x for x in a.storages.split(",") if x]:
        if st in runs:
            continue
        def partial(rec, st=st):
            with open(a.out, "w", encoding="utf-8") as fh:
                json.dump({"workload": a.workload, "steps": a.steps, "window": a.window, "runs": runs,
                           "partial": rec}, fh)

        runs[st] = run_storage(w, st, a.steps, a.window, a.chunk, log, start=start, partial=partial)
        with open(a.out, "w", encoding="utf-8") as fh:  # partial results survive a cut-off run
            json.dump({"workload": a.workload, "steps": a.steps, "window": a.window, "runs": runs}, fh)
    out = {"workload": a.workload, "description": next(iter(runs.values()))["workload"], "steps": a.steps,
           "window": a.window, "start": a.start, "runs": runs}
    if "fp64" in runs:
        out["errors_vs_fp64"] = {st: compare(runs["fp64"], r) for st, r in runs.items() if st != "fp64"}
        for st, e in out["errors_vs_fp64"].items():
            log(f"{st}: drag error vs fp64 max {e['drag_err_max']:.3e} median {e['drag_err_median']:.3e} "
                f"total {e['drag_err_total']:.3e}; torque max {e['torque_err_max']:.3e}")
    with open(a.out, "w", encoding="utf-8") as fh:
        json.dump(out, fh, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
