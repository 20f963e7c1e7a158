"""The split 16-bit D3Q27 step (k27_split: pair blocks over the plain cells,
per-cell blocks over the y/z face rows and the bodies' boxes) against the single
per-cell launch (LBM_Q27_SPLIT=0): the same populations and forces, bit for bit, on every x
treatment of the pair kernel (velocity inlet, anti-bounce-back outlet,
zero-gradient faces, periodic wrap), with static and moving bodies, simple and
interpolated bounce-back, and rows that do not fill a wave (nx = 40, 64).
The per-cell step itself is pinned to the oracle in tests/test_gpu_d3q27.py.
"""
from __future__ import annotations

import numpy as np
import pytest

from lbm3d_bb_amd.bodies import ConstantMotion, ellipsoid, rotation_matrix, sphere
from lbm3d_bb_amd.solver import Config, Solver, periodic, zero_gradient

from cases import channel, periodic_box, reference_channel

pytestmark = pytest.mark.gpu


def _run(cfg, setup, steps, split, monkeypatch):
    if split:
        monkeypatch.delenv("LBM_Q27_SPLIT", raising=False)
    else:
        monkeypatch.setenv("LBM_Q27_SPLIT", "0")
    s = Solver(cfg)
    try:
        setup(s)
        F = s.step(steps)
        return F, s.populations(), s.macros()[0]
    finally:
        s.close()
        monkeypatch.delenv("LBM_Q27_SPLIT", raising=False)


def _same(cfg, setup, steps, monkeypatch, what):
    Fa, fa, ra = _run(cfg, setup, steps, True, monkeypatch)
    Fb, fb, rb = _run(cfg, setup, steps, False, monkeypatch)
    assert np.isfinite(fa).all(), what
    d = float(np.abs(fa - fb).max())
    print(what, "max |split - single|", d, "force", float(np.abs(np.asarray(Fa) - np.asarray(Fb)).max()) if Fa is not None else None)
    assert np.array_equal(fa, fb), (what, d)
    assert np.array_equal(ra, rb), what
    if Fa is not None:
        assert np.array_equal(Fa, Fb), what


def _q27(cfg, storage):
    cfg.q = 27
    cfg.storage = storage
    return cfg


@pytest.mark.parametrize("storage", ["fp16s", "bf16s"])
@pytest.mark.parametrize("bc", ["SBB", "IBB"])
def test_split_channel_inlet_moving_sphere(storage, bc, monkeypatch):
    """Velocity inlet (rule lanes at x = 0), zero-gradient outlet (clamp at
    x = nx - 1), no-slip y/z walls (per-cell rows), a translating sphere (the
    per-cell blocks move with it), refill; nx = 64: two rows per wave."""
    cfg = _q27(channel(nx=64, ny=20, nz=18, storage="fp32", tau=0.6), storage)
    cfg.solid_bc = bc
    law = ConstantMotion("sphere", (20.0, 10.0, 9.0), (4.5,), vel=(0.02, 0.0, 0.0))

    def setup(s):
        s.set_motion([law])
        s.init_equilibrium()

    _same(cfg, setup, 60, monkeypatch, f"channel {storage} {bc}")


@pytest.mark.parametrize("storage", ["fp16s", "bf16s"])
def test_split_reference_faces_abb_outlet(storage, monkeypatch):
    """The reference's face set: velocity inlet, anti-bounce-back outlet (rule
    lanes at x = nx - 1 read u from the slot), moving y walls, periodic z;
    nx = 40: a wave spans parts of four rows; a spinning ellipsoid."""
    cfg = _q27(reference_channel(nx=40, ny=24, nz=10, storage="fp32"), storage)
    law = ConstantMotion("ellipsoid", (14.0, 12.0, 5.0), (5.0, 3.0, 2.5), vel=(0.0, 0.0, 0.0),
                         omega=(0.0, 0.0, 0.01), rot0=rotation_matrix((0, 0, 1), 0.3))

    def setup(s):
        s.set_motion([law])
        s.init_equilibrium()

    _same(cfg, setup, 60, monkeypatch, f"reference faces {storage}")


@pytest.mark.parametrize("collision", ["BGK", "TRT"])
def test_split_periodic_box_forced_static_bodies(collision, monkeypatch):
    """Periodic x (one wrapped 2-byte load per edge lane), body force, two
    static bodies."""
    cfg = _q27(periodic_box(n=32, collision=collision, storage="fp32", g=(2e-5, -1e-5, 0.0)), "fp16s")
    bodies = [sphere((9.5, 15.0, 16.2), 4.2), ellipsoid((24.0, 18.0, 10.0), (5.0, 3.0, 3.0),
                                                        rot=rotation_matrix((1, 1, 0), 0.4))]

    def setup(s):
        s.set_bodies(bodies)
        s.init_equilibrium()

    _same(cfg, setup, 50, monkeypatch, f"periodic box {collision}")


def test_split_zero_gradient_x_faces(monkeypatch):
    """Zero-gradient faces on both x ends (value clamps in the pair kernel),
    periodic y and z, body force, no body."""
    wp, wm = 1.0 / 0.6, 1.0 / (0.1875 / (0.6 - 0.5) + 0.5)
    cfg = Config(48, 16, 12, collision="TRT", storage="bf16s", omega_plus=wp, omega_minus=wm,
                 faces=[zero_gradient(), zero_gradient(), periodic(), periodic(), periodic(), periodic()],
                 g=(1e-5, 0.0, 0.0), u0=(0.02, 0.0, 0.0))
    cfg.q = 27

    def setup(s):
        s.init_equilibrium()

    _same(cfg, setup, 40, monkeypatch, "zero-gradient x")


def test_split_random_start_and_getters(monkeypatch):
    """A start from arbitrary populations (set_populations): the x-face slots are
    written by the start's collision, and the getters replay the split step."""
    cfg = _q27(channel(nx=32, ny=12, nz=10, storage="fp32", tau=0.58), "fp16s")
    rng = np.random.default_rng(7)
    from lbm3d_bb_amd.lattice import C27, W27

    u = 0.04 + 0.01 * rng.standard_normal((3, cfg.ncells))
    cu = C27.astype(np.float64) @ u
    f = W27[:, None] * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * (u * u).sum(axis=0))
    f *= 1.0 + 0.01 * rng.standard_normal(cfg.ncells)

    def setup(s):
        s.set_populations(f)

    _same(cfg, setup, 30, monkeypatch, "random start")
