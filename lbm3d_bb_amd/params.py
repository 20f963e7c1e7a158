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
