"""D3Q19 lattice in the reference's direction order (src/Utilities.jl:945-993), 0-based."""
from __future__ import annotations

from fractions import Fraction

import numpy as np

Q = 19
CX = np.array([1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 0], dtype=np.int64)
CY = np.array([0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 0], dtype=np.int64)
CZ = np.array([0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 0], dtype=np.int64)
C = np.stack([CX, CY, CZ], axis=1)
W_EXACT = [Fraction(1, 18)] * 6 + [Fraction(1, 36)] * 12 + [Fraction(1, 3)]
W = np.array([float(w) for w in W_EXACT])
# bbOpp (src/Utilities.jl:960) minus one
OPP = np.array([1, 0, 3, 2, 5, 4, 7, 6, 9, 8, 11, 10, 13, 12, 15, 14, 17, 16, 18], dtype=np.int64)
CS2 = 1.0 / 3.0  # src/Collision.jl:7


def equilibrium(rho, ux, uy, uz):
    """f_eq = w rho (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u^2) (src/Initialization.jl:37-42), shape (19, ...)."""
    rho = np.asarray(rho, dtype=np.float64)
    ux, uy, uz = (np.asarray(a, dtype=np.float64) for a in (ux, uy, uz))
    u2 = ux * ux + uy * uy + uz * uz
    out = np.empty((Q,) + np.broadcast(rho, ux).shape)
    for d in range(Q):
        cu = CX[d] * ux + CY[d] * uy + CZ[d] * uz
        out[d] = W[d] * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * u2)
    return out
