"""LBMParams and the params.txt reader (src/Parameters.jl:8-164).

The derivation of the relaxation rates follows src/Parameters.jl:60-72 in FP32,
as the reference computes it:
  L = 2R (cylinder) or chord, in cells;  nu = U L / Re;
  tau+ = 1/2 + 3 nu;  tau- = 1/2 + Lambda / (tau+ - 1/2);  omega+- = 1/tau+-.
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass

import numpy as np

_NUM = re.compile(r"[+-]?\d+(?:\.\d*)?(?:[eE][+-]?\d+)?")


def trt_rates(nu: float, lam: float, dtype=np.float32) -> tuple[float, float]:
    """(omega+, omega-) from the lattice viscosity and the magic parameter Lambda."""
    f = dtype
    tau_p = f(0.5) + f(3.0) * f(nu)
    tau_m = f(0.5) + f(lam) / (tau_p - f(0.5))
    return float(f(1.0) / tau_p), float(f(1.0) / tau_m)


@dataclass
class LBMParams:
    Nx: int
    Ny: int
    Nz: int
    h: float
    R: float
    x_c: float
    y_c: float
    theta0: float
    q: int
    Re: float
    U_inf: float
    V_inf: float
    W_inf: float
    rho_inf: float
    dt: float
    Lambda: float
    n_max: int
    n_save: int
    bc: str
    coll: str
    shape: str  # "CYL" or "AIR"
    naca_code: str | None
    chord: float

    @property
    def nu(self) -> float:
        f = np.float32
        L = (f(2.0) * f(self.R) if self.shape == "CYL" else f(self.chord)) / f(self.h)
        return float(f(self.U_inf) * L / f(self.Re))

    @property
    def omega_p(self) -> float:
        return trt_rates(self.nu, self.Lambda)[0]

    @property
    def omega_m(self) -> float:
        return trt_rates(self.nu, self.Lambda)[1]


def read_params_input(path: str) -> LBMParams:
    """Positional numeric values plus keyed strings, as src/Parameters.jl:87-164 parses them.

    Numeric order (19 values): Nx Ny Nz h R x_c y_c theta0 Chord q Re U_inf V_inf
    W_inf rho_inf dt Lambda n_max n_save.  Keys: BC, Coll (default BGK), Shape
    (CYL/AIR), NACA (4 digits; default 0012 for an airfoil).
    """
    values: list[float] = []
    bc = coll = naca = None
    shape = "CYL"
    with open(path, encoding="utf-8") as fh:
        for raw in fh:
            line = raw.strip()
            if not line or line.startswith("#"):
                continue
            if line.startswith("BC:"):
                m = re.match(r"^BC:\s*\|?\s*([A-Za-z0-9_+\-]+)", line)
                if not m:
                    raise ValueError(f"Could not parse BC from: {raw!r}")
                bc = m.group(1)
                continue
            if line.startswith("Coll:"):
                m = re.match(r"^Coll:\s*\|?\s*([A-Za-z0-9_+\-]+)", line)
                if not m:
                    raise ValueError(f"Could not parse Coll from: {raw!r}")
                coll = m.group(1)
                continue
            if line.startswith("Shape:"):
                m = re.match(r"^Shape:\s*\|?\s*([A-Za-z]+)", line)
                if not m:
                    raise ValueError(f"Could not parse Shape from: {raw!r}")
                sh = m.group(1).upper()
                if sh not in ("CYL", "AIR"):
                    raise ValueError(f"Shape must be CYL or AIR, got {sh}")
                shape = sh
                continue
            if line.startswith("NACA:"):
                m = re.match(r"^NACA:\s*\|?\s*([0-9]{4})", line)
                if not m:
                    raise ValueError(f"Could not parse NACA code (expect 4 digits) from: {raw!r}")
                naca = m.group(1)
                continue
            m = _NUM.search(line)
            if not m:
                raise ValueError(f"Could not find a number in: {raw!r}")
            values.append(float(np.float32(m.group(0))))
    if len(values) != 19:
        raise ValueError(f"Expected 19 numeric parameters, got {len(values)}.")
    if bc is None:
        raise ValueError("Missing BC line in input file!")
    if coll is None:
        coll = "BGK"
    if shape == "AIR" and naca is None:
        naca = "0012"
    v = values
    return LBMParams(
        Nx=int(v[0]), Ny=int(v[1]), Nz=int(v[2]), h=v[3], R=v[4], x_c=v[5], y_c=v[6],
        theta0=float(np.float32(math.radians(v[7]))), q=int(v[9]), Re=v[10], U_inf=v[11], V_inf=v[12],
        W_inf=v[13], rho_inf=v[14], dt=v[15], Lambda=v[16], n_max=int(v[17]), n_save=int(v[18]),
        bc=bc, coll=coll, shape=shape, naca_code=naca, chord=v[8])


def ref_area_lat(p: LBMParams) -> float:
    """Reference area in lattice units (ref_area_lat, src/Utilities.jl:931-938):
    diameter (cylinder) or chord (airfoil) in cells times the span Nz."""
    f = np.float32
    L = f(2.0) * f(p.R) if p.shape == "CYL" else f(p.chord)
    return float((L / f(p.h)) * f(p.Nz))


def aero_coefficients(forces: np.ndarray, p: LBMParams) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(CL, CD, CZ) from per-step body forces (..., 6) as compute_aero! scales them
    (src/MEA.jl:57-59): fc = 1 / (0.5 rho_inf U_inf^2 A_ref)."""
    fc = 1.0 / (0.5 * p.rho_inf * p.U_inf * p.U_inf * ref_area_lat(p))
    F = np.asarray(forces)
    return F[..., 1] * fc, F[..., 0] * fc, F[..., 2] * fc


def config_from_params(p: LBMParams, storage: str = "fp32", wall_point: str = "physical",
                       refill: bool = False):
    """The reference's run set-up for one LBMParams, in lattice units
    (main_simulation_HPC, src/MainCore.jl:57-80 and initialize_distributed):

    * faces as compile_face_BC builds them (src/Utilities.jl:795-855): velocity
      (U_inf, V_inf, W_inf) on the inlet x = 1 and on both y walls,
      anti-bounce-back outlet at x = Nx, z periodic;
    * collision Coll with omega+- from Re, U_inf and Lambda (src/Parameters.jl:60-72);
    * solid boundary BC (SBB | IBB);
    * the lattice q (gen_lattice_sets, src/Utilities.jl:945-993: 19 or 27);
    * the body: the z-extruded cylinder of radius R, or the NACA airfoil of
      chord Chord pitched by theta0 about its quarter-chord pivot
      (init_solid_mask_state, src/Geometry.jl:26-39), centred at (x_c, y_c);
    * its heave law y(t) = y_c - A sin(w t), A = 0.2 (2R), w = 2 pi 0.154 U_inf / (2R)
      (src/MainCore.jl:72-80; the airfoil heaves with the same law, R included).

    The start the reference takes, the equilibrium of a velocity perturbed by up
    to 20 % per node, comes from perturbed_start(); the config alone starts
    uniform.  wall_point="reference" samples the moving-wall velocity where the
    reference does (src/SolidBC.jl:195-196).

    Returns (Config, [ConstantMotion]) for Solver(config).set_motion(laws).
    """
    from .bodies import ConstantMotion, X_INT, naca_params, rotation_z
    from .solver import Config, periodic, pressure, velocity

    h = float(p.h)
    U = (float(p.U_inf), float(p.V_inf), float(p.W_inf))
    coll = p.coll.upper()
    if coll not in ("BGK", "TRT"):
        raise ValueError(f"Unknown collision method: {p.coll}")
    cfg = Config(p.Nx, p.Ny, p.Nz, collision=coll, storage=storage, omega_plus=p.omega_p,
                 omega_minus=p.omega_m if coll == "TRT" else p.omega_p,
                 faces=[velocity(U), pressure(), velocity(U), velocity(U), periodic(), periodic()],
                 rho0=float(p.rho_inf), u0=U, wall_point=wall_point, refill=refill, solid_bc=p.bc.upper(),
                 q=int(p.q))
    A = 0.2 * 2.0 * float(p.R)
    w = 2.0 * math.pi * 0.154 * float(p.U_inf) / (2.0 * float(p.R))
    centre = (float(p.x_c) / h, float(p.y_c) / h, 0.5 * p.Nz)
    amp = (0.0, -A / h, 0.0)
    freq = w * float(p.dt)
    if p.shape == "CYL":
        law = ConstantMotion("cylinder", centre, (float(p.R) / h,), amp=amp, freq=freq)
    else:
        chord = float(p.chord) / h
        law = ConstantMotion("naca4", centre, (chord,), amp=amp, freq=freq, rot0=rotation_z(float(p.theta0)),
                             naca=naca_params(chord, p.naca_code or "0012", 0.25, X_INT))
    return cfg, [law]


def perturbed_start(cfg, seed: int = 0, fac: float = 0.2) -> tuple[np.ndarray, np.ndarray]:
    """(rho, u) of the reference's start (initialize_f_gpu, src/Initialization.jl:27-44,
    with rand_node of :70): rho = rho0 and u = u0 (1 + fac r) componentwise, one
    r = 2 rand - 1 in [-1, 1) per node shared by the three components, drawn in
    FP32 from numpy's PCG64 stream of `seed` (the reference draws CUDA.rand; its
    stream is not reproduced).  Arrays for Solver.init_equilibrium(rho, u)."""
    N = cfg.nx * cfg.ny * cfg.nz
    rng = np.random.Generator(np.random.PCG64(seed))
    r = (np.float32(2.0) * rng.random(N, dtype=np.float32) - np.float32(1.0)).astype(np.float64)
    u0 = np.asarray(cfg.u0, dtype=np.float64)
    rho = np.full(N, float(cfg.rho0))
    u = u0[:, None] * (1.0 + fac * r)[None, :]
    return rho, u
