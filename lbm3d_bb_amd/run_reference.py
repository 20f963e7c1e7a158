"""A params.txt run the way the reference's time loop drives it, on one GPU.

  python -m lbm3d_bb_amd.run_reference --params input/params.txt --out output

main_simulation_HPC (src/MainCore.jl:57-379) with one worker: the set-up from
the input file (config_from_params: faces, collision, BC, body, heave law),
n_max time steps, the aerodynamic coefficients of every step
(compute_aero!, src/MEA.jl:57-59) into output/aero_coeffs.h5
(save_coefficients!, src/I_O.jl:587-662), a snapshot of rho, u, v, w every
n_save steps into output/final_fields.h5 (write_fields_h5_time!,
src/I_O.jl:296-405), and the MLUPS summary of measure_Mlups
(src/Utilities.jl:903-926).  The steps between two outputs run as one lbm_step
call (the body's law is evaluated on the device).  The start is the reference's
(initialize_f, src/Initialization.jl:57-83: the equilibrium of a start velocity
perturbed by up to 20 % per node, seeded) or, with --start uniform, the
unperturbed equilibrium.
"""
from __future__ import annotations

import argparse
import math
import os
import sys
import time

from .h5io import AeroCoefficients, FieldSnapshots
from .params import aero_coefficients, config_from_params, perturbed_start, read_params_input


def run(params_path: str, out_dir: str = "output", storage: str = "fp32", n_max: int | None = None,
        n_save: int | None = None, wall_point: str = "physical", log=print, params=None, make_solver=None,
        start: str = "perturbed", seed: int = 0) -> dict:
    """params (an LBMParams) replaces reading params_path; make_solver(config)
    builds the solver (default: the HIP Solver); start "perturbed" (the
    reference's, seeded) or "uniform"."""
    if start not in ("perturbed", "uniform"):
        raise ValueError(f"unknown start {start!r}")
    from .solver import Solver

    p = params if params is not None else read_params_input(params_path)
    if n_max is not None:
        p.n_max = int(n_max)
    if n_save is not None:
        p.n_save = int(n_save)
    cfg, laws = config_from_params(p, storage=storage, wall_point=wall_point)
    num_snaps = math.ceil(p.n_max / p.n_save)
    os.makedirs(out_dir, exist_ok=True)
    t_ini = time.perf_counter()
    fields = FieldSnapshots(os.path.join(out_dir, "final_fields.h5"), p.Nx, p.Ny, p.Nz, num_snaps)
    coeffs = AeroCoefficients(os.path.join(out_dir, "aero_coeffs.h5"), p.n_max)
    s = make_solver(cfg) if make_solver is not None else Solver(cfg)
    s.set_motion(laws)
    if start == "perturbed":
        s.init_equilibrium(*perturbed_start(cfg, seed))
    else:
        s.init_equilibrium()
    s.sync()
    t0 = time.perf_counter()
    n, t_steps = 0, 0.0
    while n < p.n_max:
        k = min(p.n_save - n % p.n_save, p.n_max - n)
        ts = time.perf_counter()
        F = s.step(k)
        s.sync()
        t_steps += time.perf_counter() - ts
        cl, cd, cz = aero_coefficients(F[:, 0, :], p)
        coeffs.write(n + 1, cl, cd, cz)
        n += k
        if n % p.n_save == 0:
            rho, u = s.macros()
            fields.write(n // p.n_save, n, rho, u)
            log(f"Saving snapshot for n = {n}.")
    t_end = time.perf_counter()
    s.close()
    N = p.Nx * p.Ny * p.Nz
    res = {"n_max": p.n_max, "cells": N, "seconds_total": t_end - t_ini, "seconds_loop": t_end - t0,
           "mlups_loop": N * p.n_max / ((t_end - t0) * 1e6), "mlups_steps": N * p.n_max / (t_steps * 1e6)}
    log("--------------------------------------------------------")
    log("  Simulation finished.")
    log(f"  Wall-clock (incl. I/O)         : {res['seconds_total']:.4f} s")
    log(f"  MLUPS (loop incl. output)      : {res['mlups_loop']:.2f} MLUPS")
    log(f"  MLUPS (steps only)             : {res['mlups_steps']:.2f} MLUPS")
    log("--------------------------------------------------------")
    return res


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--params", required=True)
    ap.add_argument("--out", default="output")
    ap.add_argument("--storage", default="fp32", choices=["fp64", "fp32", "fp16s", "bf16s"])
    ap.add_argument("--n-max", type=int, default=None, help="override n_max of the input file")
    ap.add_argument("--n-save", type=int, default=None, help="override n_save of the input file")
    ap.add_argument("--wall-point", default="physical", choices=["physical", "reference"])
    ap.add_argument("--start", default="perturbed", choices=["perturbed", "uniform"],
                    help="perturbed: the reference's start, u = U (1 + 0.2 r) per node; uniform: u = U")
    ap.add_argument("--seed", type=int, default=0, help="seed of the perturbed start")
    a = ap.parse_args(argv)
    run(a.params, a.out, a.storage, a.n_max, a.n_save, a.wall_point,
        log=lambda m: print(m, flush=True), start=a.start, seed=a.seed)
    return 0


if __name__ == "__main__":
    sys.exit(main())
