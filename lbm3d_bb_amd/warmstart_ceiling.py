"""How much any start state can shorten the run: steps to a converged drag of
held-out sphere packs from starts of known accuracy.

  python -m lbm3d_bb_amd.warmstart_ceiling --size 128 --seeds 1000000,1000001,... \\
      --out profiles/warmstart_ceiling_128.json

For each pack the uniform start runs `--settle` steps; the state it ends in is
the reference steady state (its drag residual is recorded).  Then the pack runs
again from starts built from that state:

  populations   the steady populations themselves (the best any start can be)
  fields_ce     the steady (rho, u) with the first-order non-equilibrium part
                (what a perfect network output gives through
                populations_from_fields)
  fields_eq     the equilibrium of the steady (rho, u)
  noise_<e>     the steady fields with smooth errors of rms e u_in on u and of
                relative rms e on rho - 1 (a network of that accuracy), with
                the non-equilibrium part

  rho1          the steady u with rho = 1 everywhere (a network that predicts
                the velocity only)
  rho_mean1     the steady fields with rho shifted to a fluid mean of 1
                (the pressure field without its level)
  net, net_rho1, net_u_exact_rho, net_flux, net_flux_eq
                with --params: the network's (rho, u); its u with rho = 1;
                its u with the steady rho; its fields with every x-slice's
                flux set to the inlet's (flux_correct), with and without the
                non-equilibrium part

and the uniform start itself.  Every run records the drag (x-force summed over
the bodies) of every step, so the steps to convergence of any residual rule
(window length, tolerance, patience) are evaluated afterwards on the same
histories.  The result holds the histories' rule evaluations, not the
histories.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from .packs import make_pack
from .warmstart import flux_correct, populations_from_fields


def steps_to_converge(drag: np.ndarray, every: int, tol: float = 1e-4, patience: int = 3) -> int | None:
    """The DragMonitor rule (lbm3d_bb_amd.convergence) on a per-step drag
    history: window means of `every` steps, relative change below tol for
    `patience` checks in a row.  None if the history never converges."""
    n = len(drag) // every
    if n < 2:
        return None
    m = drag[: n * every].reshape(n, every).mean(axis=1)
    ok = 0
    for i in range(1, n):
        r = abs(m[i] - m[i - 1]) / max(abs(m[i]), 1e-300)
        ok = ok + 1 if r < tol else 0
        if ok >= patience:
            return (i + 1) * every
    return None


def smooth_noise(shape, rng, scale_cells: float) -> np.ndarray:
    """A zero-mean, unit-rms random field on a periodic (nz, ny, nx) grid,
    smoothed with a Gaussian of `scale_cells` (in Fourier space)."""
    a = rng.standard_normal(shape)
    k = [np.fft.fftfreq(n) for n in shape]
    K2 = k[0][:, None, None] ** 2 + k[1][None, :, None] ** 2 + k[2][None, None, :] ** 2
    a = np.real(np.fft.ifftn(np.fft.fftn(a) * np.exp(-2.0 * (np.pi * scale_cells) ** 2 * K2)))
    a -= a.mean()
    return a / a.std()


def drag_history(solver, nsteps: int, chunk: int = 2000) -> np.ndarray:
    out = []
    done = 0
    while done < nsteps:
        k = min(chunk, nsteps - done)
        F = solver.step(k)
        out.append(F[:, :, 0].sum(axis=1))
        done += k
    return np.concatenate(out)


RULES = [(100, 1e-4), (200, 1e-4), (500, 1e-4), (1000, 1e-4)]


def evaluate_rules(drag: np.ndarray) -> dict:
    return {f"every{e}_tol{t:g}": steps_to_converge(drag, e, t) for e, t in RULES}


def run_case(seed: int, L: int, settle: int, steps: int, noise: list[float], log, net=None) -> dict:
    from .solver import Solver
    from .warmstart import predict

    case = make_pack(seed, L)
    shape = (L, L, L)
    s = Solver(case.config("fp32"))
    s.set_bodies(case.bodies)
    s.init_equilibrium()
    t0 = time.perf_counter()
    d_uni = drag_history(s, settle)
    f_ss = s.populations()
    rho, u = s.macros()
    log(f"seed {seed} Re {case.re:.1f} spheres {len(case.bodies)}: uniform {settle} steps in "
        f"{time.perf_counter() - t0:.1f} s, last-window residual {abs(d_uni[-200:-100].mean() / d_uni[-100:].mean() - 1):.2e}")
    solid = (s.flags() != 0)
    res = {"seed": seed, "re": case.re, "spheres": len(case.bodies), "tau_plus": case.tau_plus,
           "drag_steady": float(d_uni[-1000:].mean()),
           "uniform": evaluate_rules(d_uni)}
    fluid = ~solid
    starts = {"populations": ("pop", None), "fields_ce": ("ce", 0.0), "fields_eq": ("eq", 0.0),
              "rho1": ("ce", "rho1"), "rho_mean1": ("ce", "rho_mean1")}
    for e in noise:
        starts[f"noise_{e:g}"] = ("ce", e)
    if net is not None:
        rn, un = predict(net, case)
        starts.update({"net": ("ce", "net"), "net_rho1": ("ce", "net_rho1"), "net_u_exact_rho": ("ce", "net_u"),
                       "net_flux": ("ce", "net_flux"), "net_flux_eq": ("eq", "net_flux")})
        du = (un - u)[:, fluid] / case.u_in
        dr = (rn - rho)[fluid]
        res["net_pred"] = {"u_rms": float(np.sqrt((du * du).mean())),
                           "rho_rel_rms": float(np.sqrt((dr * dr).mean()) / np.sqrt(((rho[fluid] - 1.0) ** 2).mean())),
                           "rho_mean_pred": float(rn[fluid].mean()), "rho_mean_steady": float(rho[fluid].mean())}
    rng = np.random.default_rng(seed)
    for name, (kind, e) in starts.items():
        if kind == "pop":
            s.set_populations(f_ss)
        else:
            r2, u2 = rho.copy(), u.copy()
            if e == "rho1":
                r2[:] = 1.0
            elif e == "rho_mean1":
                r2 = rho - rho[fluid].mean() + 1.0
            elif e == "net":
                r2, u2 = rn.copy(), un.copy()
            elif e == "net_rho1":
                r2, u2 = np.ones_like(rho), un.copy()
            elif e == "net_u":
                u2 = un.copy()
            elif e == "net_flux":
                r2, u2 = rn.copy(), flux_correct(rn, un, shape, case.u_in, solid)
            elif e:
                # errors smooth on the scale of a sphere: rms e u_in on each velocity
                # component, rms e on the relative deviation of rho from the rest state
                r2 = 1.0 + (rho - 1.0) * (1.0 + e * smooth_noise(shape, rng, 4.0).reshape(-1))
                for a in range(3):
                    u2[a] = u[a] + e * case.u_in * smooth_noise(shape, rng, 4.0).reshape(-1)
                u2[:, solid] = 0.0
            if kind == "eq":
                s.init_equilibrium(r2, u2)
            else:
                s.set_populations(populations_from_fields(r2, u2, shape, case.tau_plus, periodic=(False, True, True)))
        d = drag_history(s, steps)
        res[name] = evaluate_rules(d)
        res[name]["drag_end_rel"] = float(abs(d[-1000:].mean() / res["drag_steady"] - 1.0))
        log(f"  {name}: " + ", ".join(f"{k} {v}" for k, v in res[name].items() if k.startswith("every")))
    s.close()
    return res


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--seeds", default="1000000,1000001,1000002,1000003")
    ap.add_argument("--settle", type=int, default=60000, help="uniform-start steps that define the steady state")
    ap.add_argument("--steps", type=int, default=30000, help="steps of every other start")
    ap.add_argument("--noise", default="0.01,0.03,0.1")
    ap.add_argument("--params", default=None, help="U-Net parameters (.npy): add the network's starts")
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    net = None
    if a.params:
        from .warmstart import pack_net

        net = pack_net((a.size,) * 3, batch=1)
        net.load(a.params)

    def log(m):
        print(m, flush=True)

    rows = []
    for sd in [int(x) for x in a.seeds.split(",") if x]:
        rows.append(run_case(sd, a.size, a.settle, a.steps, [float(x) for x in a.noise.split(",") if x], log, net))
        with open(a.out, "w", encoding="utf-8") as fh:
            json.dump({"size": a.size, "settle": a.settle, "steps": a.steps, "rules": RULES, "rows": rows}, fh,
                      indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
