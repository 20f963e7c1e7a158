"""lbm_conservation_static, the allocation-free variant of lbm_conservation,
against lbm_conservation itself through the C ABI (GPU), to the parity bars:
1e-12 (FP64 storage) and 1e-5 (FP32) normwise relative.  Two lattices in a
row, the second smaller, share the library's device array; each context is
asked twice."""
from __future__ import annotations

import ctypes as ct

import numpy as np
import pytest

from lbm3d_bb_amd import _ffi
from lbm3d_bb_amd.bodies import ConstantMotion, sphere
from lbm3d_bb_amd.solver import Solver

from cases import TOL, channel, periodic_box

pytestmark = pytest.mark.gpu


def _static(s):
    m = ct.c_double(0.0)
    p = np.zeros(3)
    s._call("conservation_static", ct.byref(m), p.ctypes.data_as(_ffi._D))
    return m.value, p


def _case(kind, n, q, storage):
    if kind == "box":
        cfg = periodic_box(n=n, collision="TRT", storage=storage, g=(2e-5, -1e-5, 5e-6))
        cfg.u0 = (0.02, -0.01, 0.005)
        cfg.q = q
        s = Solver(cfg)
        s.set_bodies([sphere((0.4 * n, 0.5 * n, 0.45 * n), 0.2 * n)])
    else:
        cfg = channel(nx=2 * n, ny=n, nz=n, storage=storage)
        cfg.q = q
        s = Solver(cfg)
        s.set_motion([ConstantMotion("sphere", (0.5 * n, 0.5 * n, 0.5 * n), (0.2 * n,), vel=(0.02, 0.0, 0.0),
                                     omega=(0.0, 0.01, 0.0))])
    s.init_equilibrium()
    return s


@pytest.mark.parametrize("kind", ["box", "channel"])
@pytest.mark.parametrize("storage", ["fp64", "fp32"])
@pytest.mark.parametrize("q", [19, 27])
def test_conservation_static_matches_lbm_conservation(kind, storage, q):
    tol = TOL[storage]
    for n, steps in ((32, 23), (18, 7)):  # the second lattice is smaller
        s = _case(kind, n, q, storage)
        for k in range(2):
            if k:
                s.step(steps)
            m_a, p_a = s.conservation()
            m_s, p_s = _static(s)
            assert abs(m_s - m_a) <= tol * m_a, (n, k, m_s, m_a)
            scale = np.abs(p_a).max()
            assert scale > 1e-3 * m_a
            assert np.abs(p_s - p_a).max() <= tol * scale, (n, k, p_s, p_a)
        s.close()


def test_conservation_static_start_state_and_errors():
    cfg = periodic_box(n=12, collision="BGK", storage="fp64")
    s = Solver(cfg)
    with pytest.raises(_ffi.LbmError):  # before init: no state to sum
        _static(s)
    cfg.u0 = (0.03, 0.0, -0.01)
    s2 = Solver(cfg)
    s2.init_equilibrium()
    m, p = _static(s2)
    N = cfg.ncells
    # uniform start: M = N rho0, P = M (u0 + g / 2) exactly up to FP64 rounding
    assert abs(m - N) < 1e-12 * N
    assert np.abs(p - N * (np.asarray(cfg.u0) + 0.5 * np.asarray(cfg.g))).max() < 1e-12 * N
    s.close()
    s2.close()
