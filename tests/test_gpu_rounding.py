"""Stochastic rounding of the 16-bit step stores, pinned by the momentum budget.

A periodic box without bodies, driven by a body force g along x, gains exactly
M g of momentum per step (Guo forcing: the collision adds rho g to every cell's
momentum), and keeps its mass.  At g = 1e-6 in a mean flow of 0.01 along
x, the per-step change of a population, 3 w_d g ~ 1.7e-7, is about 0.02 of a
BF16 ulp of its deviation from the weight (|f - w| ~ 3 w u ~ 1.7e-3, ulp 2^-17)
and about 0.2 of an FP16 ulp (the deviation scaled by 2^12, ulp 2^-8 there, so
9.5e-7 unscaled).  Rounding to nearest would drop that change at every store
and the flow would stop accelerating; a rounding biased by a fraction of an ulp
would shift the gain by a comparable fraction of the increment.  Unbiased
stochastic rounding keeps every increment in expectation, so the momentum
gained over 400 steps matches 400 M g.  The rounding noise is a random walk of
independent store errors (about 0.4 ulp each, 400 steps x 32^3 cells x 10
directions with c_x != 0 for D3Q19): ~0.3 % of the gain for BF16 and ~0.04 %
for FP16.  The 1.5 % bar is five BF16 standard deviations and far below what a
dropped or biased increment gives.  The mass moves by the same noise only
(relative ~1e-6).
"""

from __future__ import annotations

import numpy as np
import pytest

from lbm3d_bb_amd.solver import Solver

from cases import periodic_box

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("q", [19, 27])
@pytest.mark.parametrize("storage", ["fp16s", "bf16s"])
def test_forced_periodic_box_momentum_gain(storage, q):
    steps, gx = 400, 1e-6
    cfg = periodic_box(n=32, storage=storage, collision="TRT", tau=0.6, g=(gx, 0.0, 0.0))
    cfg.q = q
    cfg.u0 = (0.01, 0.0, 0.0)
    s = Solver(cfg)
    s.init_equilibrium()
    s.step(1, forces=False)
    m0, p0 = s.conservation()
    s.step(steps, forces=False)
    m1, p1 = s.conservation()
    s.close()
    gain = (p1[0] - p0[0]) / (steps * m0 * gx)
    dm = abs(m1 - m0) / m0
    dyz = float(np.abs(p1[1:] - p0[1:]).max()) / (steps * m0 * gx)
    print(f"{storage} q{q}: momentum gain / (steps M g) = {gain:.5f}, y/z {dyz:.2e}, mass drift {dm:.2e}", flush=True)
    assert abs(gain - 1.0) < 1.5e-2, gain
    assert dyz < 1.5e-2, dyz
    assert dm < 1e-4, dm
