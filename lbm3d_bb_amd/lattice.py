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


# D3Q27 in the reference order (src/Utilities.jl:962-976), 0-based: the D3Q19
# moving directions, the eight corners, rest; weights by the number of non-zero
# velocity components (0: 8/27, 1: 2/27, 2: 1/54, 3: 1/216).
CX27 = np.array([1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 0],
                dtype=np.int64)
CY27 = np.array([0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1, 0],
                dtype=np.int64)
CZ27 = np.array([0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1, 0],
                dtype=np.int64)
C27 = np.stack([CX27, CY27, CZ27], axis=1)
W27_EXACT = [Fraction(2, 27)] * 6 + [Fraction(1, 54)] * 12 + [Fraction(1, 216)] * 8 + [Fraction(8, 27)]
W27 = np.array([float(w) for w in W27_EXACT])
# bbOpp (src/Utilities.jl:976) minus one
OPP27 = np.array([1, 0, 3, 2, 5, 4, 7, 6, 9, 8, 11, 10, 13, 12, 15, 14, 17, 16, 19, 18, 21, 20, 23, 22, 25, 24, 26],
                 dtype=np.int64)


def tables(q: int = 19):
    """(C (q, 3), W (q,), OPP (q,)) of D3Q19 or D3Q27."""
    if q == 19:
        return C, W, OPP
    if q == 27:
        return C27, W27, OPP27
    raise ValueError(f"Invalid q value: {q}")
