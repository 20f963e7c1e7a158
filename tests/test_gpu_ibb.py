"""The planar shapes and the interpolated bounce-back: HIP path vs the oracle (GPU).

Same bars as the other parity tests (tests/cases.py): flags and links
bit-exact; populations, rho and u within 1e-12 (FP64) or 1e-5 (FP32),
normwise; per-body force and torque within the same tolerance relative to
max|F_oracle| plus the rounding floor of the link sum.  The reference's
own input file (tests/golden/params_reference_default.txt: 1201 x 1201 x 3,
heaving cylinder, IBB, BGK) runs at full size.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from lbm3d_bb_amd.bodies import ConstantMotion, cylinder, ellipsoid, heave_motion, naca4, naca_params, rotation_matrix
from lbm3d_bb_amd.bodies import rotation_z, sphere
from lbm3d_bb_amd.lattice import C as LC, W as LW
from lbm3d_bb_amd.params import config_from_params, read_params_input
from lbm3d_bb_amd.solver import Config, Solver, periodic
from oracle.oracle import OracleSolver

from cases import TOL, channel, check_forces, check_state, periodic_box, reference_channel, rel_err

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
def _pair(cfg, prec):
    return Solver(cfg), OracleSolver(cfg, prec)


def _compare_state(g, o, tol, what=""):
    check_state(g, o, "fp64" if tol < 1e-6 else "fp32", what, tol=tol)


def _same_geometry(g, o):
    assert np.array_equal(g.flags(), o.flags())
    cg, dg = g.links()
    co, do = o.links()
    assert len(co) > 0
    assert np.array_equal(cg, co) and np.array_equal(dg, do)


def test_flags_and_links_bit_exact_cylinder_and_airfoil():
    cfg = reference_channel(nx=96, ny=64, nz=4, storage="fp32")
    for bodies in ([cylinder((30.3, 31.7, 2.0), 9.37)],
                   [naca4((40.2, 32.9, 2.0), 48.0, "2412", theta=0.17)],
                   [cylinder((20.0, 16.5, 2.0), 5.0), naca4((60.0, 40.0, 2.0), 30.0, "0015", theta=-0.4)]):
        g, o = _pair(cfg, "fp32")
        for s in (g, o):
            s.set_bodies(bodies)
            s.update_geometry()
        _same_geometry(g, o)


@pytest.mark.parametrize("collision", ["BGK", "TRT"])
def test_ibb_static_cylinder_reference_channel_fp64(collision):
    cfg = reference_channel(nx=64, ny=40, nz=4, storage="fp64")
    cfg.collision = collision
    cfg.solid_bc = "IBB"
    g, o = _pair(cfg, "fp64")
    for s in (g, o):
        s.set_bodies([cylinder((20.3, 19.6, 2.0), 6.37)])
        s.init_equilibrium()
    Fg, Fo = g.step(60), o.step(60)
    assert np.abs(Fo).max() > 0
    check_forces(Fg, Fo, o, "fp64", collision, shapes=[cylinder((20.3, 19.6, 2.0), 6.37)])
    _compare_state(g, o, TOL["fp64"], collision)


def test_ibb_pitched_airfoil_and_3d_bodies_fp64():
    cfg = periodic_box(n=40, collision="TRT", g=(2e-5, 5e-6, 0.0))
    cfg.solid_bc = "IBB"
    bodies = [naca4((14.2, 20.3, 20.0), 20.0, "2412", theta=0.14),
              sphere((31.3, 9.8, 10.4), 4.6),
              ellipsoid((30.1, 29.7, 28.2), (7.3, 4.1, 3.3), rot=rotation_matrix((0.2, 0.7, -0.4), 0.9))]
    g, o = _pair(cfg, "fp64")
    for s in (g, o):
        s.set_bodies(bodies)
        s.init_equilibrium()
    Fg, Fo = g.step(50), o.step(50)
    check_forces(Fg, Fo, o, "fp64", "mixed", shapes=bodies)
    _compare_state(g, o, TOL["fp64"], "mixed")


def test_ibb_heaving_cylinder_scaled_reference_case_fp64():
    """The reference's set-up (faces, BGK, IBB, heave law) on a small lattice."""
    p = read_params_input(os.path.join(HERE, "golden", "params_reference_default.txt"))
    p.Nx, p.Ny, p.Nz, p.R, p.x_c, p.y_c, p.Re = 120, 80, 3, 6.3, 40.5, 40.0, 20.0
    cfg, laws = config_from_params(p, storage="fp64", refill=True)
    # heave faster than the reference so that the mask moves within the run
    laws[0].freq = 0.05
    g, o = _pair(cfg, "fp64")
    for s in (g, o):
        s.set_motion(laws)
        s.init_equilibrium()
    Fg, Fo = g.step(40), o.step(40)
    _same_geometry(g, o)
    check_forces(Fg, Fo, o, "fp64", "heave", shapes=laws)
    _compare_state(g, o, TOL["fp64"], "heave")


@pytest.mark.parametrize("wall_point", ["physical", "reference"])
def test_ibb_moving_airfoil_and_spinning_ellipsoid_fp64(wall_point):
    cfg = channel(nx=64, ny=40, nz=36, storage="fp64")
    cfg.solid_bc = "IBB"
    cfg.wall_point = wall_point
    laws = [ConstantMotion("naca4", (20.3, 20.1, 18.0), (24.0,), vel=(0.02, 0.01, 0.0), omega=(0.0, 0.0, 0.004),
                           rot0=rotation_z(0.1), naca=naca_params(24.0, "4412")),
            ConstantMotion("ellipsoid", (45.2, 19.7, 17.6), (6.0, 3.5, 3.0), omega=(0.01, 0.02, 0.03),
                           rot0=rotation_matrix((1, 1, 0), 0.3))]
    g, o = _pair(cfg, "fp64")
    for s in (g, o):
        s.set_motion(laws)
        s.init_equilibrium()
    Fg, Fo = g.step(30), o.step(30)
    _same_geometry(g, o)
    check_forces(Fg, Fo, o, "fp64", wall_point, shapes=laws)
    _compare_state(g, o, TOL["fp64"], wall_point)


def test_ibb_cylinder_channel_fp32():
    cfg = channel(nx=80, ny=48, nz=8, storage="fp32")
    cfg.solid_bc = "IBB"
    g, o = _pair(cfg, "fp32")
    law = heave_motion("cylinder", (24.3, 24.2, 4.0), (7.4,), 7.4, 0.05)
    for s in (g, o):
        s.set_motion([law])
        s.init_equilibrium()
    Fg, Fo = g.step(50), o.step(50)
    check_forces(Fg, Fo, o, "fp32", "fp32", shapes=[law])
    _compare_state(g, o, TOL["fp32"], "fp32")


def test_ibb_uniform_flow_moving_with_bodies_stays_at_equilibrium():
    """Exact steady solution: every population stays f_eq(1, U) when the walls move at U."""
    U = (0.05, 0.02, -0.01)
    cfg = Config(48, 40, 32, collision="TRT", storage="fp64", omega_plus=1 / 0.7, omega_minus=1 / 1.1,
                 faces=[periodic()] * 6, u0=U, refill=True, solid_bc="IBB")
    laws = [ConstantMotion("cylinder", (12.3, 11.8, 8.0), (5.3,), vel=U),
            ConstantMotion("naca4", (30.2, 28.9, 8.0), (18.0,), vel=U, rot0=rotation_z(0.3),
                           naca=naca_params(18.0, "2412")),
            ConstantMotion("ellipsoid", (34.0, 12.0, 20.0), (6.0, 4.0, 3.0), vel=U,
                           rot0=rotation_matrix((1, 2, 3), 0.5))]
    g = Solver(cfg)
    g.set_motion(laws)
    g.init_equilibrium()
    g.step(12)
    cu = LC @ np.asarray(U)
    feq = LW * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * np.dot(U, U))
    assert np.abs(g.populations() - feq[:, None]).max() < 1e-13


@pytest.mark.parametrize("storage", ["fp16s", "bf16s"])
def test_ibb_16bit_storage_drag(storage):
    """16-bit storage with the interpolated rule: drag of a cylinder in a channel
    after 1500 steps within 1 % of the FP64 run."""
    def drag(st):
        cfg = channel(nx=96, ny=48, nz=4, storage=st, u_in=0.05, tau=0.6)
        cfg.faces[4] = cfg.faces[5] = periodic()
        cfg.solid_bc = "IBB"
        s = Solver(cfg)
        s.set_bodies([cylinder((30.3, 24.4, 2.0), 6.3)])
        s.init_equilibrium()
        F = s.step(1500)
        return F[-200:, 0, 0].mean()

    d64, d16 = drag("fp64"), drag(storage)
    assert abs(d16 / d64 - 1) < 0.01, (d16, d64)


def test_reference_default_case_full_size_fp32():
    """The reference's own input file, at its full 1201 x 1201 x 3 size, 3 steps, FP32."""
    p = read_params_input(os.path.join(HERE, "golden", "params_reference_default.txt"))
    cfg, laws = config_from_params(p, storage="fp32")
    g, o = _pair(cfg, "fp32")
    for s in (g, o):
        s.set_motion(laws)
        s.init_equilibrium()
    Fg, Fo = g.step(3), o.step(3)
    _same_geometry(g, o)
    check_forces(Fg, Fo, o, "fp32", "reference default", shapes=laws)
    _compare_state(g, o, TOL["fp32"], "reference default")


def test_run_reference_driver_outputs_gpu(tmp_path):
    """The reference's loop on the GPU (scaled input file, IBB, heave): its HDF5
    outputs hold the oracle's coefficients and snapshots within the FP32 bar."""
    h5io = pytest.importorskip("lbm3d_bb_amd.h5io")
    try:
        h5io.lib()
    except ImportError as e:
        pytest.skip(f"HDF5 C library not available: {e}")
    from lbm3d_bb_amd.params import aero_coefficients, perturbed_start
    from lbm3d_bb_amd.run_reference import run

    p = read_params_input(os.path.join(HERE, "golden", "params_reference_default.txt"))
    p.Nx, p.Ny, p.Nz, p.R, p.x_c, p.y_c, p.Re = 96, 64, 3, 7.3, 32.5, 32.2, 40.0
    out = str(tmp_path / "output")
    run("", out, n_max=50, n_save=20, params=p, log=lambda m: None)
    cfg, laws = config_from_params(p)
    o = OracleSolver(cfg, "fp32")
    o.set_motion(laws)
    o.init_equilibrium(*perturbed_start(cfg, 0))  # the driver's default start
    F, snaps = [], []
    for k in (20, 20, 10):
        F.append(o.step(k))
        if len(F) < 3:
            snaps.append(o.macros())
    F = np.concatenate(F)[:, 0, :]
    cl, cd, _ = aero_coefficients(F, p)
    with h5io.H5File(os.path.join(out, "aero_coeffs.h5")) as f:
        for name, ref in (("CD", cd), ("CL", cl)):
            assert rel_err(f.read(name), ref) < TOL["fp32"], (name, rel_err(f.read(name), ref))
    with h5io.H5File(os.path.join(out, "final_fields.h5")) as f:
        assert list(f.read("steps")) == [20, 40, 0]
        rho, u = f.read("rho"), f.read("u")
        for s, (r, uu) in enumerate(snaps):
            assert rel_err(rho[s].ravel(), r) < TOL["fp32"]
            assert rel_err(u[s].ravel(), uu[0]) < TOL["fp32"]
