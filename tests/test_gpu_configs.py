"""The benchmark configurations at their own sizes (GPU).

* 256x128x128 D3Q19 TRT channel, FP32, translating sphere R = 12 (the bench
  workload): HIP against the FP32 oracle for 50 steps; flags and links bit-exact,
  force, torque and state at the FP32 bar.
* 512x256x256 periodic box, 64 fixed spheres r in [6, 10] from seed 0: flags and
  links bit-exact against the oracle's geometry, HIP FP32 against the FP32 oracle
  for 5 steps; the FP64 mass and momentum budget of the FP64, FP32 and FP16S runs.
* 1024x384x384 channel, FP16S, 8 ellipsoids (16, 8, 8) rotating at 1e-3 rad/step:
  flags and links bit-exact against the oracle's geometry at the pose after 200
  steps; drag and torque of the FP16S run against the FP32 run.

Bars as in tests/cases.py.  Each test prints what it measured.
"""
from __future__ import annotations

import numpy as np
import pytest

import bench
from lbm3d_bb_amd import _ffi
from lbm3d_bb_amd.solver import Solver
from oracle.oracle import OracleSolver, body_at

from cases import check_forces, check_state

pytestmark = pytest.mark.gpu


def _case(name, storage=None):
    w = dict(bench.WORKLOADS[name])
    if storage:
        w["storage"] = storage
    return bench.make_case(w)


def _same_geometry(g, o, what):
    fg, fo = g.flags(), o.flags()
    assert np.array_equal(fg, fo), f"{what}: {(fg != fo).sum()} flags differ"
    cg, dg = g.links()
    co, do = o.links()
    assert len(co) > 0
    assert np.array_equal(cg, co) and np.array_equal(dg, do), what
    return len(co)


def test_channel256_d3q19_fp32_full_size_vs_oracle():
    cfg, laws, _, _ = _case("channel256")
    g, o = Solver(cfg), OracleSolver(cfg, "fp32")
    for s in (g, o):
        s.set_motion(laws)
        s.init_equilibrium()
    Fg, Fo = g.step(50), o.step(50)
    nl = _same_geometry(g, o, "channel256")
    mf = check_forces(Fg, Fo, o, "fp32", "channel256", shapes=laws)
    ms = check_state(g, o, "fp32", "channel256")
    print(f"channel256 fp32 50 steps: links {nl}, force rel {mf['F_rel']:.2e}, torque rel {mf['T_rel']:.2e}, "
          f"f {ms['f']:.2e}, rho {ms['rho']:.2e}, u {ms['u']:.2e}", flush=True)


def test_box512_spheres64_full_size_fp32_vs_oracle():
    """Started from a mean flow so that the drags of the first steps are well
    above the FP32 resolution of their link sums (from rest they are ~1e-8 of
    sum |t_link|)."""
    cfg, _, bodies, _ = _case("box512_spheres64")
    cfg.u0 = (0.02, 0.005, -0.003)
    g, o = Solver(cfg), OracleSolver(cfg, "fp32")
    for s in (g, o):
        s.set_bodies(bodies)
        s.init_equilibrium()
    Fg, Fo = g.step(5), o.step(5)
    nl = _same_geometry(g, o, "box512")
    mf = check_forces(Fg, Fo, o, "fp32", "box512", shapes=bodies)
    ms = check_state(g, o, "fp32", "box512")
    print(f"box512 fp32 5 steps: links {nl}, force rel {mf['F_rel']:.2e}, f {ms['f']:.2e}, rho {ms['rho']:.2e}, "
          f"u {ms['u']:.2e}", flush=True)
    g.close()
    o.close()


# Mass and momentum budget over 40 steps, from lbm_conservation's FP64 sums of
# the fluid cells: the mass is conserved, and the fluid momentum changes by
# M g - sum_bodies F_MEA each step (Guo forcing adds rho g at every cell; the
# momentum exchange takes F from the fluid).  FP64 storage holds both to FP64
# rounding.  FP32 arithmetic rounds every post-collision population to half an
# ulp of its size ~w_d, so per cell and step the mass may move by up to
# sum_d ulp(w_d)/2 and the momentum by up to sum_d |c_dx| ulp(w_d)/2, which at
# g = 1e-6 is ~1 % of the forcing increment rho g (FP16S collides in FP32 too).


def _fp32_rounding_bounds(gx: float, nsteps: int) -> tuple[float, float]:
    from lbm3d_bb_amd.lattice import CX, W

    half_ulp = np.spacing(W.astype(np.float32)).astype(np.float64) / 2.0
    return nsteps * float(half_ulp.sum()), float((np.abs(CX) * half_ulp).sum()) / gx


@pytest.mark.parametrize("storage", ["fp64", "fp32", "fp16s"])
def test_box512_spheres64_conservation(storage):
    cfg, _, bodies, _ = _case("box512_spheres64", storage)
    g = Solver(cfg)
    g.set_bodies(bodies)
    g.init_equilibrium()
    g.step(10)
    m0, p0 = g.conservation()
    F = g.step(40)
    m1, p1 = g.conservation()
    g.close()
    gx = np.asarray(cfg.g)
    dp_expected = 40 * m0 * gx - F[:, :, :3].sum(axis=(0, 1))
    dm = abs(m1 - m0) / m0
    # relative to the impulse the body force gives the fluid over the 40 steps
    dp = float(np.abs((p1 - p0) - dp_expected).max() / (40 * m0 * np.abs(gx).max()))
    tm, tp = (1e-13, 1e-9) if storage == "fp64" else _fp32_rounding_bounds(float(np.abs(gx).max()), 40)
    print(f"box512 {storage}: mass drift {dm:.2e} (bound {tm:.1e}), momentum budget {dp:.2e} (bound {tp:.1e})",
          flush=True)
    assert dm < tm, dm
    assert dp < tp, dp


def test_channel1024_ellipsoids8_geometry_and_fp16s_vs_fp32():
    nsteps = 200
    res = {}
    for storage in ("fp16s", "fp32"):
        cfg, laws, _, _ = _case("channel1024_ellipsoids8", storage)
        g = Solver(cfg)
        g.set_motion(laws)
        g.init_equilibrium()
        F = g.step(nsteps)
        res[storage] = F
        if storage == "fp16s":
            # the oracle's geometry (no stepping) at the mask pose of step nsteps
            o = OracleSolver(cfg, "fp32")
            posed = (_ffi.LbmBody * len(laws))(*[body_at(l, nsteps) for l in laws])
            o._call("set_bodies", len(laws), posed)
            o.nbodies = len(laws)
            o.update_geometry()
            nl = _same_geometry(g, o, "channel1024")
            o.close()
        g.close()
    F16, F32 = res["fp16s"], res["fp32"]
    # drag and torque magnitude per body over the last 100 steps
    last = slice(nsteps - 100, nsteps)
    dF = np.abs(F16[last, :, :3] - F32[last, :, :3]).max() / np.abs(F32[last, :, :3]).max()
    dT = np.abs(F16[last, :, 3:] - F32[last, :, 3:]).max() / np.abs(F32[last, :, 3:]).max()
    print(f"channel1024 ellipsoids8: links {nl}; fp16s vs fp32 over steps 101-200: force {dF:.2e}, "
          f"torque {dT:.2e}", flush=True)
    assert dF < 5e-3 and dT < 5e-3, (dF, dT)
    assert np.isfinite(F16).all()
