"""Shared small cases: the same script runs on the HIP path and on the oracle."""
from __future__ import annotations

import json
import os

import numpy as np

from lbm3d_bb_amd.bodies import ConstantMotion, ellipsoid, rotation_matrix, sphere
from lbm3d_bb_amd.solver import Config, no_slip, periodic, pressure, velocity, zero_gradient


def trt_omegas(tau_p: float, lam: float = 3.0 / 16.0) -> tuple[float, float]:
    tau_m = 0.5 + lam / (tau_p - 0.5)
    return 1.0 / tau_p, 1.0 / tau_m


def periodic_box(n=24, storage="fp64", collision="BGK", tau=0.8, g=(1e-5, 0.0, 0.0)):
    """Periodic box, BGK tau=0.8 (or TRT), body force along x."""
    wp, wm = trt_omegas(tau)
    return Config(n, n, n, collision=collision, storage=storage, omega_plus=1.0 / tau if collision == "BGK" else wp,
                  omega_minus=wm, faces=[periodic()] * 6, g=g)


def channel(nx=40, ny=20, nz=20, storage="fp32", u_in=0.05, outlet="zero_gradient", tau=0.56, refill=True):
    """Channel: inlet velocity at x=0, outlet at x=Nx-1, no-slip y/z walls, TRT 3/16."""
    wp, wm = trt_omegas(tau)
    out = zero_gradient() if outlet == "zero_gradient" else pressure()
    return Config(nx, ny, nz, collision="TRT", storage=storage, omega_plus=wp, omega_minus=wm,
                  faces=[velocity((u_in, 0, 0)), out, no_slip(), no_slip(), no_slip(), no_slip()],
                  u0=(u_in, 0.0, 0.0), refill=refill)


def reference_channel(nx=40, ny=24, nz=6, storage="fp64", U=0.075, tau=0.6):
    """The reference's own face set (src/Utilities.jl:795-855): velocity inlet and
    y-walls moving at U_inf, anti-bounce-back outlet, z periodic; BGK."""
    return Config(nx, ny, nz, collision="BGK", storage=storage, omega_plus=1.0 / tau, omega_minus=1.0 / tau,
                  faces=[velocity((U, 0, 0)), pressure(), velocity((U, 0, 0)), velocity((U, 0, 0)), periodic(),
                         periodic()], u0=(U, 0.0, 0.0), wall_point="reference")


def sphere_pack(n_bodies, L, rmin, rmax, seed):
    """Non-overlapping spheres by rejection sampling in a periodic box of side L."""
    rng = np.random.default_rng(seed)
    out = []
    tries = 0
    while len(out) < n_bodies:
        tries += 1
        if tries > 100000:
            raise RuntimeError("packing failed")
        r = rng.uniform(rmin, rmax)
        c = rng.uniform(0, L, size=3)
        ok = True
        for c2, r2 in out:
            d = np.abs(c - c2)
            d = np.minimum(d, L - d)
            if np.linalg.norm(d) < r + r2 + 2.0:
                ok = False
                break
        if ok:
            out.append((c, r))
    return [sphere(tuple(c), r) for c, r in out]


def force_scale(oracle_solver, rho0: float = 1.0) -> float:
    """Momentum-exchange scale of the force sums: S = sum over links of 2 w_d rho0,
    the summed magnitude of the link terms the per-body force cancels down from."""
    from lbm3d_bb_amd.lattice import W, W27

    _, dirs = oracle_solver.links()
    w = W27 if oracle_solver.config.q == 27 else W
    return float(2.0 * rho0 * w[dirs.astype(np.int64)].sum())


def force_err(Fg, Fo, scale: float) -> float:
    return float(np.abs(np.asarray(Fg) - np.asarray(Fo)).max() / scale)


def rel_err(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))


# ------------------------------------------------------------ parity bars ----
# Populations and density: normwise relative error max|gpu - oracle| /
# max|oracle| over the whole array below TOL (1e-12 for FP64 storage against the
# FP64 oracle, 1e-5 for FP32 against the FP32 oracle).
#
# Velocity: below TOL * max|u_oracle| plus what the population differences and
# the moment sum's rounding can move it by.  u_a = sum_d c_da f_d / rho (+ g/2)
# is a difference of populations ~w_d rho, so f differing by df moves u_a by up
# to (n_a + q |u_a|) max|df| / rho (n_a directions with c_da != 0, q of them in
# all), and each side rounds the sum by up to gamma_q sum_d |c_da| f_d / rho <=
# gamma_q, whatever |u|.  With the populations at their bar this is the bound
# the velocity can be held to; the log records max|du| / max|u| beside it.
#
# Per-body force and torque: max|F_gpu - F_oracle| (every step, body and
# component) below TOL * max|F_oracle| plus the rounding floor of the sum the
# force is.  F = sum over n links of c_d t_link with |t_link| ~ 2 w_d rho: the
# n terms cancel by orders of magnitude (a sphere's drag is ~1e-5 of
# A = sum |t_link|), and each side's summation in its own order may round by
# up to gamma_n A, gamma_n = n eps / (1 - n eps) (Higham, Accuracy and
# Stability of Numerical Algorithms, 2nd ed., eq. 4.4), so the two sides can
# differ by 2 gamma_n A whatever the populations.  The torque's terms carry an
# arm |r| <= the body's extent, so its floor is 2 gamma_n A R.  The log records
# the error relative to max|F_oracle| next to the floor.
TOL = {"fp64": 1e-12, "fp32": 1e-5}
EPS = {"fp64": 2.0 ** -53, "fp32": 2.0 ** -24}


def _log(kind: str, what: str, rec: dict) -> None:
    path = os.environ.get("PARITY_LOG")
    if not path:
        return
    rec = {"test": os.environ.get("PYTEST_CURRENT_TEST", "").split(" ")[0], "kind": kind, "what": what, **rec}
    with open(path, "a") as fh:
        fh.write(json.dumps(rec) + "\n")


def gamma(n: int, prec: str) -> float:
    """Higham's gamma_n = n eps / (1 - n eps), eps the unit roundoff of prec."""
    return n * EPS[prec] / (1.0 - n * EPS[prec])


def u_bar(uo, ro, df_max: float, q: int, prec: str, tol: float) -> float:
    """The velocity bar above, from the oracle's u, rho and the largest population
    difference."""
    na = 18 if q == 27 else 10
    umax = float(np.abs(uo).max())
    return tol * umax + (na + q * umax) * df_max / float(np.abs(ro).min()) + 2.0 * gamma(q, prec)


def check_state(g, o, prec: str, what: str = "", tol: float | None = None) -> dict:
    """Populations, rho and u of the HIP solver against the oracle's (the bars above)."""
    tol = TOL[prec] if tol is None else tol
    rg, ug = g.macros()
    ro, uo = o.macros()
    fg, fo = g.populations(), o.populations()
    assert fg.shape == fo.shape
    df = float(np.abs(fg - fo).max())
    du = float(np.abs(ug - uo).max())
    m = {"f": rel_err(fg, fo), "rho": rel_err(rg, ro), "u": rel_err(ug, uo), "u_max": float(np.abs(uo).max()),
         "du": du, "u_bar": u_bar(uo, ro, df, fg.shape[0], prec, tol), "tol": tol}
    _log("state", what, m)
    for k in ("f", "rho"):
        assert m[k] < tol, f"{what}: {k} {m[k]:.3e} >= {tol:.0e} ({m})"
    assert du <= m["u_bar"], f"{what}: u {m}"
    return m


def arm_bound(cfg, shapes) -> float:
    """Upper bound of the torque arm |x_w - X| over a body's links: its largest
    semi-axis (chord of an airfoil) plus 2 cells; z-extruded shapes add half the
    z extent."""
    arm = 0.0
    for s in shapes:
        ext = float(max(s.semi)) + 2.0
        if s.shape in ("cylinder", "naca4"):
            ext = float(np.hypot(ext, 0.5 * cfg.nz))
        arm = max(arm, ext)
    return arm


def check_forces(Fg, Fo, o, prec: str, what: str = "", shapes=None, tol: float | None = None) -> dict:
    """Per-body force and torque history of the HIP solver against the oracle's
    (the bars above).  ``shapes`` (bodies or motion laws) bound the torque arm;
    without them only the force is compared."""
    from lbm3d_bb_amd.lattice import W, W27

    tol = TOL[prec] if tol is None else tol
    Fg = np.asarray(Fg, dtype=np.float64)
    Fo = np.asarray(Fo, dtype=np.float64)
    assert Fg.shape == Fo.shape
    cells, dirs = o.links()
    rho, _ = o.macros()
    w = W27 if o.config.q == 27 else W
    n = max(len(dirs), 1)
    A = float((2.0 * w[dirs.astype(np.int64)] * np.abs(rho[cells.astype(np.int64)])).sum())
    # the sums run in FP64 on both sides; with FP32 populations each link term
    # t = f*_d + f_bb is also formed in FP32 (<= 3 roundings of |t| per side)
    floor = 2.0 * gamma(n, "fp64") * A + (6.0 * EPS["fp32"] * A if prec == "fp32" else 0.0)
    dF = float(np.abs(Fg[..., :3] - Fo[..., :3]).max())
    Fmax = float(np.abs(Fo[..., :3]).max())
    m = {"dF": dF, "F_max": Fmax, "F_rel": dF / Fmax if Fmax > 0 else 0.0, "A": A, "nlinks": int(len(dirs)),
         "floor_F": floor, "tol": tol}
    m["bar_F"] = tol * Fmax + floor
    if shapes is not None:
        arm = arm_bound(o.config, shapes)
        dT = float(np.abs(Fg[..., 3:] - Fo[..., 3:]).max())
        Tmax = float(np.abs(Fo[..., 3:]).max())
        m.update({"dT": dT, "T_max": Tmax, "T_rel": dT / Tmax if Tmax > 0 else 0.0, "arm": arm,
                  "bar_T": tol * Tmax + floor * arm})
    _log("force", what, m)
    assert dF <= m["bar_F"], f"{what}: force {m}"
    if shapes is not None:
        assert m["dT"] <= m["bar_T"], f"{what}: torque {m}"
    return m


__all__ = ["ConstantMotion", "ellipsoid", "rotation_matrix", "sphere", "trt_omegas", "periodic_box", "channel",
           "reference_channel", "sphere_pack", "rel_err", "force_scale", "force_err", "TOL", "EPS",
           "check_state", "check_forces", "arm_bound", "gamma", "u_bar"]
