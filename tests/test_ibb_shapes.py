"""Planar shapes (z-extruded cylinder, NACA 4-digit airfoil) and the interpolated
bounce-back, pinned on the oracle (no GPU).

Pins, each chosen so that a plausible slip fails it:
  * NACA thickness against the published NACA 0012 ordinates (Abbott & von
    Doenhoff, Theory of Wing Sections, App. I), the closed trailing edge of the
    reference's x_int squeeze, the camber line's maximum m at x = p;
  * cylinder and airfoil masks: z-invariance, cell counts against the areas
    (pi R^2; the integral of the thickness), the pitch direction;
  * level sets at points whose distance to the surface is known exactly;
  * the interpolated bounce-back keeps a uniform flow that moves with the body
    exactly at equilibrium for every wall fraction q (this fixes the sign of
    the moving-wall term and its 1/(2q) on the q >= 1/2 branch, and the weights
    2q, 1 - 2q and (2q - 1)/(2q));
  * Stokes drag of a square array of cylinders against Sangani & Acrivos
    (1982): within 2.5 % with the interpolated rule over sub-cell radius
    changes, where the simple bounce-back staircase misses by up to 5 %.
"""
from __future__ import annotations

import math
import os

import numpy as np
import pytest

from lbm3d_bb_amd.bodies import ConstantMotion, cylinder, ellipsoid, naca4, naca_params, rotation_matrix, sphere
from lbm3d_bb_amd.lattice import C as LC, W as LW
from lbm3d_bb_amd.params import config_from_params, read_params_input
from lbm3d_bb_amd.solver import Config, periodic
from oracle import oracle as orc
from oracle.oracle import OracleSolver

HERE = os.path.dirname(os.path.abspath(__file__))

# NACA 0012 half-thickness y/c at x/c (open trailing edge, x_int = 1)
NACA0012 = [(0.0125, 0.01894), (0.025, 0.02615), (0.05, 0.03555), (0.1, 0.04683), (0.2, 0.05737),
            (0.3, 0.06002), (0.4, 0.05803), (0.5, 0.05294), (0.6, 0.04563), (0.7, 0.03664),
            (0.8, 0.02623), (0.9, 0.01448), (0.95, 0.00807)]


def test_naca0012_thickness_matches_published_ordinates():
    c = 100.0
    for xc, yt in NACA0012:
        yl, yu = orc.naca_lines(xc * c, c, 0.0, 0.0, 0.12, 1.0)
        assert abs(yu / c - yt) < 2e-5, (xc, yu / c, yt)
        assert yl == -yu  # symmetric section: bit-exact mirror


def test_naca_trailing_edge_squeeze_and_camber():
    c = 120.0
    yl, yu = orc.naca_lines(c, c, 0.0, 0.0, 0.12, 1.0089304)
    assert abs(yu) < 1e-5 * c  # x_int closes the trailing edge
    yl1, yu1 = orc.naca_lines(c, c, 0.0, 0.0, 0.12, 1.0)
    assert abs(yu1 / c - 0.00126) < 2e-5  # open edge of the plain formula
    # 2412: the camber line (in chord fractions, as the reference keeps it) peaks at m
    m, p, t, _, _ = naca_params(c, "2412")
    assert (m, p) == (pytest.approx(0.02), pytest.approx(0.4)) and t == pytest.approx(0.12)
    yl, yu = orc.naca_lines(p * c, c, m, p, t, 1.0)
    assert abs(0.5 * (yl + yu) - m) < 1e-6
    # the surfaces straddle the camber line by the half thickness times cos(slope) (slope 0 at p)
    yl0, yu0 = orc.naca_lines(p * c, c, 0.0, 0.0, t, 1.0)
    assert abs((yu - yl) - (yu0 - yl0)) < 1e-4


def _flags(solver):
    cfg = solver.config
    return solver.flags().reshape(cfg.nz, cfg.ny, cfg.nx)


def test_cylinder_mask_extruded_and_gauss_count():
    cfg = Config(48, 40, 5, collision="BGK", storage="fp64", omega_plus=1.0, faces=[periodic()] * 6)
    o = OracleSolver(cfg, "fp64")
    R = 9.37
    o.set_bodies([cylinder((20.3, 19.6, 2.0), R)])
    o.update_geometry()
    f = _flags(o)
    for z in range(1, cfg.nz):
        assert np.array_equal(f[z], f[0])
    n = int((f[0] != 0).sum())
    assert abs(n - math.pi * R * R) < 2.0 * math.pi * R * 0.5  # Gauss circle bound
    # every cell centre strictly inside R is solid, strictly outside R + tol is fluid
    y, x = np.mgrid[0:cfg.ny, 0:cfg.nx]
    d2 = (x + 0.5 - 20.3) ** 2 + (y + 0.5 - 19.6) ** 2
    assert np.all(f[0][d2 < (R - 1e-3) ** 2] == 1)
    assert np.all(f[0][d2 > (R + 1e-3) ** 2] == 0)


def _naca_area(c, t, x_int, n=200001):
    x = np.linspace(0.0, c, n)
    xi = np.maximum(0.0, x_int * x / c)
    yt = 5 * t * c * (0.2969 * np.sqrt(xi) - 0.1260 * xi - 0.3516 * xi ** 2 + 0.2843 * xi ** 3 - 0.1015 * xi ** 4)
    return 2.0 * np.trapezoid(yt, x)


def test_naca_mask_area_and_pitch_direction():
    cfg = Config(96, 96, 3, collision="BGK", storage="fp64", omega_plus=1.0, faces=[periodic()] * 6)
    chord = 60.0
    pivot = (48.0, 48.0, 1.5)
    o = OracleSolver(cfg, "fp64")
    o.set_bodies([naca4(pivot, chord, "0012")])
    o.update_geometry()
    f0 = _flags(o)[0]
    area = _naca_area(chord, 0.12, 1.0089304)
    assert abs(int((f0 != 0).sum()) - area) < 0.05 * area
    # the chord runs along +x from the leading edge 0.25 c ahead of the pivot
    assert f0[48, int(48 - 0.25 * chord) + 2] and f0[48, int(48 + 0.75 * chord) - 8]
    assert not f0[48, int(48 - 0.25 * chord) - 2] and not f0[48, int(48 + 0.75 * chord) + 2]
    # pitch +90 deg about z turns the chord to +y (body frame R^T d, src/Geometry.jl:316-318)
    R90 = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    o.set_bodies([naca4(pivot, chord, "0012", rot=R90)])
    o.update_geometry()
    f90 = _flags(o)[0]
    assert f90[int(48 + 0.5 * chord), 48] and not f90[int(48 - 0.5 * chord), 48]
    # a symmetric section pitched 90 deg about a cell corner is the transposed mask
    assert np.array_equal(f90 != 0, (f0 != 0).T)


def test_level_set_closed_forms():
    per, L = (0, 0, 0), (1000, 1000, 1000)
    cyl = cylinder((10.5, 10.5, 3.0), 3.3)
    # radial points: distance minus radius, z ignored
    assert orc.phi(cyl, (14.5, 10.5, 0.5), per, L) == pytest.approx(0.7, abs=2e-6)
    assert orc.phi(cyl, (13.5, 10.5, 7.5), per, L) == pytest.approx(-0.3, abs=2e-6)
    assert orc.phi(cyl, (10.5, 14.5, 0.5), per, L) == pytest.approx(0.7, abs=2e-6)
    sph = sphere((5.0, 5.0, 5.0), 2.0)
    assert orc.phi(sph, (5.0, 5.0, 8.0), per, L) == pytest.approx(1.0, abs=1e-6)
    ell = ellipsoid((0.0, 0.0, 0.0), (4.0, 2.0, 1.0), rot=rotation_matrix((0, 0, 1), math.pi / 2),
                    rot_half=rotation_matrix((0, 0, 1), math.pi / 2))
    # rotated 90 deg about z: the long axis lies along y; the surface point (0, 4, 0)
    assert orc.phi(ell, (0.0, 4.0, 0.0), per, L) == pytest.approx(0.0, abs=1e-6)
    assert orc.phi(ell, (0.0, 8.0, 0.0), per, L) == pytest.approx(1.0, abs=1e-6)
    assert orc.phi(ell, (4.0, 0.0, 0.0), per, L) == pytest.approx(1.0, abs=1e-6)
    # airfoil: vertical distance over the chord, distance to the edge point beyond it
    c = 40.0
    af = naca4((0.0, 0.0, 0.0), c, "0012")
    _, yu = orc.naca_lines(0.5 * c, c, 0.0, 0.0, 0.12, 1.0089304)
    assert orc.phi(af, (0.25 * c, yu + 2.0, 0.0), per, L) == pytest.approx(2.0, abs=1e-5)
    assert orc.phi(af, (0.25 * c, 0.0, 0.0), per, L) == pytest.approx(-yu, abs=1e-5)
    assert orc.phi(af, (-0.25 * c - 3.0, 4.0, 0.0), per, L) == pytest.approx(5.0, abs=1e-5)


def _feq(rho, u):
    cu = LC @ np.asarray(u, dtype=np.float64)
    return LW * rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * np.dot(u, u))


@pytest.mark.parametrize("shape", ["sphere", "ellipsoid", "cylinder", "naca4"])
def test_ibb_keeps_a_uniform_flow_moving_with_the_body(shape):
    """f = f_eq(rho, U) everywhere is a steady solution when the wall moves at U:
    the interpolated rule must reproduce f_eq(U)_opp from f_eq(U) for any q."""
    U = (0.04, -0.025, 0.015)
    cfg = Config(32, 30, 28, collision="TRT", storage="fp64", omega_plus=1 / 0.7, omega_minus=1 / 1.1,
                 faces=[periodic()] * 6, u0=U, refill=True, solid_bc="IBB")
    semi = {"sphere": (6.3,), "ellipsoid": (7.7, 4.2, 3.6), "cylinder": (6.3,), "naca4": (18.0,)}[shape]
    rot0 = rotation_matrix((1, 2, 3), 0.7) if shape == "ellipsoid" else (
        rotation_matrix((0, 0, 1), 0.21) if shape == "naca4" else np.eye(3))
    law = ConstantMotion(shape, (15.3, 14.8, 13.9), semi, vel=U, rot0=rot0,
                         naca=naca_params(18.0, "2412") if shape == "naca4" else (0.0,) * 5)
    o = OracleSolver(cfg, "fp64")
    o.set_motion([law])
    o.init_equilibrium()
    o.step(6)
    f = o.populations()
    rho, _ = o.macros()
    assert np.abs(f - _feq(1.0, U)[:, None]).max() < 1e-13
    assert np.abs(rho - 1.0).max() < 1e-13


def _ibb_rule_counts(o, body):
    """Links per rule (0 simple, 1 q < 1/2, 2 q >= 1/2), from the oracle's level set."""
    cfg = o.config
    cells, dirs = o.links()
    per = tuple(1 if cfg.faces[2 * a].kind == "periodic" else 0 for a in range(3))
    L = (cfg.nx, cfg.ny, cfg.nz)
    flags = o.flags()
    counts = [0, 0, 0]
    for cell, d in zip(cells, dirs):
        x, y, z = cell % cfg.nx, (cell // cfg.nx) % cfg.ny, cell // (cfg.nx * cfg.ny)
        c = LC[d]
        pb = np.array([x + 0.5, y + 0.5, z + 0.5], dtype=np.float32)
        phb = orc.phi(body, pb, per, L)
        phs = orc.phi(body, pb + c.astype(np.float32), per, L)
        if not (phb > 0 and phs < 0):
            counts[0] += 1
            continue
        q = phb / (phb - phs)
        if q >= 0.5:
            counts[2] += 1
        else:
            s = [(x - c[0]) % cfg.nx, (y - c[1]) % cfg.ny, (z - c[2]) % cfg.nz]
            counts[1 if flags[s[0] + cfg.nx * (s[1] + cfg.ny * s[2])] == 0 else 0] += 1
    return counts


def _square_array_drag(R, bc, L=32, tau=0.8):
    """Steady drag coefficient F / (mu U) of a periodic square array of cylinders
    driven by a body force; F as a mean pressure gradient G exerts it, G L^2 =
    F_MEA / (1 - c); U the superficial velocity."""
    tm = 0.5 + (3.0 / 16.0) / (tau - 0.5)
    gx = 1e-6
    cfg = Config(L, L, 2, collision="TRT", storage="fp64", omega_plus=1 / tau, omega_minus=1 / tm,
                 faces=[periodic()] * 6, g=(gx, 0.0, 0.0), solid_bc=bc)
    o = OracleSolver(cfg, "fp64")
    body = cylinder((L / 2, L / 2, 1.0), R)
    o.set_bodies([body])
    o.init_equilibrium()
    last = None
    for _ in range(100):
        F = o.step(200)
        fx = F[-1, 0, 0]
        if last is not None and abs(fx - last) < 1e-7 * abs(fx):
            break
        last = fx
    rho, u = o.macros()
    fluid = o.flags() == 0
    c = math.pi * R * R / (L * L)
    U = u[0][fluid].sum() / (L * L * 2)
    nu = (tau - 0.5) / 3.0
    K = (fx / 2.0) / ((1.0 - c) * nu * U)
    K_sa = 4 * math.pi / (-0.5 * math.log(c) - 0.738 + c - 0.887 * c * c + 2.039 * c ** 3)
    return K / K_sa - 1.0, o, body


def test_ibb_square_array_drag_sangani_acrivos():
    errs = {"SBB": [], "IBB": []}
    for R in (5.2, 5.5, 5.8, 6.1, 6.4):
        for bc in errs:
            e, o, body = _square_array_drag(R, bc)
            errs[bc].append(e)
            if bc == "IBB" and R == 5.5:
                counts = _ibb_rule_counts(o, body)
                assert counts[1] > 0 and counts[2] > 0  # both interpolation branches ran
    assert max(abs(e) for e in errs["IBB"]) < 0.025, errs
    assert max(abs(e) for e in errs["SBB"]) > 2 * max(abs(e) for e in errs["IBB"]), errs


def test_reference_default_params_case():
    """The reference's own input file maps to its set-up: IBB, BGK, the heaving cylinder."""
    p = read_params_input(os.path.join(HERE, "golden", "params_reference_default.txt"))
    cfg, laws = config_from_params(p)
    assert (cfg.nx, cfg.ny, cfg.nz) == (1201, 1201, 3) and cfg.solid_bc == "IBB" and cfg.collision == "BGK"
    assert [f.kind for f in cfg.faces] == ["velocity", "pressure", "velocity", "velocity", "periodic", "periodic"]
    # nu = U 2R / Re; BGK takes omega+ = 1 / (1/2 + 3 nu)
    nu = 0.075 * 2 * 16.16 / 185
    assert cfg.omega_plus == pytest.approx(1 / (0.5 + 3 * nu), rel=1e-6)
    law = laws[0]
    assert law.shape == "cylinder" and law.semi[0] == pytest.approx(16.16)
    assert law.amp[1] == pytest.approx(-0.4 * 16.16, rel=1e-6)
    assert law.freq == pytest.approx(2 * math.pi * 0.154 * 0.075 / (2 * 16.16), rel=1e-6)
