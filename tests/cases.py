"""Shared small cases: the same script runs on the HIP path and on the oracle."""
from __future__ import annotations

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
    the magnitude of the summands the per-body force cancels down from.  FP64
    force parity is judged relative to S (a 10^3-term sum cancelling to 1e-2
    carries ~1e-12 relative rounding against |F| alone)."""
    from lbm3d_bb_amd.lattice import W

    _, dirs = oracle_solver.links()
    return float(2.0 * rho0 * W[dirs.astype(np.int64)].sum())


def force_err(Fg, Fo, scale: float) -> float:
    return float(np.abs(np.asarray(Fg) - np.asarray(Fo)).max() / scale)


def rel_err(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))


__all__ = ["ConstantMotion", "ellipsoid", "rotation_matrix", "sphere", "trt_omegas", "periodic_box", "channel",
           "reference_channel", "sphere_pack", "rel_err", "force_scale", "force_err"]
