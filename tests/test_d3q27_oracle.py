"""The oracle's D3Q27 lattice (src/Utilities.jl:962-976) pinned by what the
mathematics fixes: isotropy of the tables, equilibrium moments, exact
conservation, the uniform-flow fixed point with inflow/outflow faces, the
shear-wave viscosity, the TRT Poiseuille parabola, the single-cell momentum
exchange and the per-step momentum budget, and the 26-direction link list by
brute force."""
from __future__ import annotations

import numpy as np
import pytest

from lbm3d_bb_amd.bodies import Body, sphere
from lbm3d_bb_amd.lattice import C, C27, OPP27, W27, tables
from lbm3d_bb_amd.solver import Config, no_slip, periodic, pressure, velocity, zero_gradient
from oracle.oracle import OracleSolver

from cases import periodic_box, rel_err, trt_omegas


def _box(n, collision="TRT", g=(0.0, 0.0, 0.0), tau=0.8):
    cfg = periodic_box(n=n, collision=collision, g=g, tau=tau)
    cfg.q = 27
    return cfg


def _feq27(rho, u):
    cu = C27.astype(float) @ u
    return W27[:, None] * rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * (u * u).sum(0))


def test_d3q27_tables():
    Cf = C27.astype(float)
    assert abs(W27.sum() - 1.0) < 1e-15
    assert np.allclose(W27 @ Cf, 0.0, atol=1e-16)
    assert np.allclose(np.einsum("d,da,db->ab", W27, Cf, Cf), np.eye(3) / 3.0, atol=1e-16)
    I = np.eye(3)
    iso = (np.einsum("ab,cd->abcd", I, I) + np.einsum("ac,bd->abcd", I, I) + np.einsum("ad,bc->abcd", I, I)) / 9.0
    assert np.allclose(np.einsum("d,da,db,dc,de->abce", W27, Cf, Cf, Cf, Cf), iso, atol=1e-16)
    assert np.array_equal(C27[OPP27], -C27)
    # 27 distinct velocities, the first 18 those of D3Q19, the rest direction last
    assert len({tuple(c) for c in C27}) == 27
    assert np.array_equal(C27[:18], C[:18]) and not C27[26].any()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_d3q27_equilibrium_moments(prec):
    cfg = _box(5)
    o = OracleSolver(cfg, prec)
    rng = np.random.default_rng(0)
    rho = 1.0 + 0.05 * rng.standard_normal(cfg.ncells)
    u = 0.05 * rng.standard_normal((3, cfg.ncells))
    o.init_equilibrium(rho, u)
    f = o.populations()
    assert f.shape == (27, cfg.ncells)
    Cf = C27.astype(float)
    tol = 1e-13 if prec == "fp64" else 2e-6
    assert rel_err(f.sum(0), rho) < tol
    assert np.abs(Cf.T @ f - rho * u).max() < tol
    P = np.einsum("da,db,dn->abn", Cf, Cf, f)
    Pex = np.einsum("n,an,bn->abn", rho, u, u) + np.eye(3)[:, :, None] * rho / 3.0
    assert np.abs(P - Pex).max() < tol * 10


@pytest.mark.parametrize("collision", ["BGK", "TRT"])
def test_d3q27_conservation_and_body_force(collision):
    g = np.array([1e-4, -2e-4, 5e-5])
    cfg = _box(7, collision=collision, g=tuple(g))
    o = OracleSolver(cfg, "fp64")
    rng = np.random.default_rng(1)
    f = _feq27(np.ones(cfg.ncells), 0.05 * rng.standard_normal((3, cfg.ncells)))
    f += 1e-3 * rng.standard_normal(f.shape)
    o.set_populations(f)
    m0, p0 = o.conservation()
    o.step(20, forces=False)
    m1, p1 = o.conservation()
    assert abs(m1 - m0) < 1e-12 * m0
    assert np.abs((p1 - p0) - 20 * m0 * g).max() < 1e-12 * m0


@pytest.mark.parametrize("outlet", ["pressure", "zero_gradient"])
def test_d3q27_inlet_outlet_uniform_flow_fixed_point(outlet):
    U = 0.04
    wp, wm = trt_omegas(0.7)
    cfg = Config(10, 6, 6, collision="TRT", storage="fp64", omega_plus=wp, omega_minus=wm,
                 faces=[velocity((U, 0, 0)), pressure() if outlet == "pressure" else zero_gradient(),
                        velocity((U, 0, 0)), velocity((U, 0, 0)), periodic(), periodic()], u0=(U, 0, 0), q=27)
    o = OracleSolver(cfg, "fp64")
    o.init_equilibrium()
    f0 = o.populations()
    o.step(15, forces=False)
    assert rel_err(o.populations(), f0) < 1e-14


def test_d3q27_shear_wave_decay_rate():
    L, tau, A = 48, 0.9, 1e-4
    cfg = Config(3, L, 3, collision="BGK", storage="fp64", omega_plus=1 / tau, omega_minus=1 / tau,
                 faces=[periodic()] * 6, q=27)
    o = OracleSolver(cfg, "fp64")
    k = 2 * np.pi / L
    y = (np.arange(cfg.ncells) // 3) % L
    u = np.zeros((3, cfg.ncells))
    u[0] = A * np.sin(k * (y + 0.5))
    o.init_equilibrium(np.ones(cfg.ncells), u)
    n = 800
    o.step(n, forces=False)
    _, u1 = o.macros()
    amp = 2.0 / cfg.ncells * (u1[0] * np.sin(k * (y + 0.5))).sum()
    nu_fit = -np.log(amp / A) / (k * k * n)
    assert abs(nu_fit / ((tau - 0.5) / 3.0) - 1.0) < 2e-3


def test_d3q27_poiseuille_exact_trt_magic():
    H, gx, tau = 10, 1e-5, 0.8
    wp, wm = trt_omegas(tau)
    nu = (tau - 0.5) / 3.0
    cfg = Config(3, H, 3, collision="TRT", storage="fp64", omega_plus=wp, omega_minus=wm,
                 faces=[periodic(), periodic(), no_slip(), no_slip(), periodic(), periodic()], g=(gx, 0, 0), q=27)
    o = OracleSolver(cfg, "fp64")
    o.init_equilibrium()
    o.step(6000, forces=False)
    _, u = o.macros()
    ux = u[0].reshape(3, H, 3)[1, :, 1]
    yv = np.arange(H) + 0.5
    assert rel_err(ux, gx * yv * (H - yv) / (2.0 * nu)) < 1e-9
    assert np.abs(u[1:]).max() < 1e-15


def test_d3q27_mea_single_cell():
    """One solid cell: F = 2 rho u in uniform flow, F = -2 rho u_w when it moves in fluid at rest."""
    n, u0 = 8, np.array([0.03, -0.01, 0.02])
    cfg = _box(n)
    cfg.u0 = tuple(u0)
    o = OracleSolver(cfg, "fp64")
    flags = np.zeros(n ** 3, dtype=np.uint8)
    flags[4 + n * (4 + n * 4)] = 1
    o.set_flags(flags)
    o.init_equilibrium()
    F = o.step(1)
    assert np.abs(F[0, 0, :3] - 2.0 * u0).max() < 1e-15
    uw = np.array([0.02, 0.01, -0.03])
    o = OracleSolver(_box(9), "fp64")
    o.set_bodies([Body("sphere", (4.5, 4.5, 4.5), (0.4, 0.4, 0.4), vel=tuple(uw))])
    o.init_equilibrium()
    F = o.step(1)
    assert (o.flags() != 0).sum() == 1
    assert len(o.links()[0]) == 26
    assert np.abs(F[0, 0, :3] + 2.0 * uw).max() < 1e-15


def test_d3q27_momentum_budget_and_links():
    n = 16
    g = np.array([2e-5, 1e-5, 0.0])
    cfg = _box(n, g=tuple(g))
    o = OracleSolver(cfg, "fp64")
    bodies = [sphere((5.2, 8.1, 7.7), 3.0), sphere((12.5, 11.0, 4.0), 2.2)]
    o.set_bodies(bodies)
    o.init_equilibrium()
    o.step(5)
    m0, p0 = o.conservation()
    F = o.step(1)
    m1, p1 = o.conservation()
    Ftot = F[0, :, :3].sum(0)
    assert abs(m1 - m0) < 1e-13 * m0
    assert np.abs((p1 - p0) - (m0 * g - Ftot)).max() < 1e-13 * (np.abs(p0).max() + m0 * np.abs(g).max() + 1.0)
    # links: every (fluid cell, direction) whose neighbour across the periodic box is solid
    flags = o.flags().reshape(n, n, n)
    cells, dirs = o.links()
    Cq, _, _ = tables(27)
    want = []
    for cell in np.nonzero(flags.ravel() == 0)[0]:
        x, y, z = cell % n, (cell // n) % n, cell // (n * n)
        for d in range(26):
            if flags[(z + Cq[d, 2]) % n, (y + Cq[d, 1]) % n, (x + Cq[d, 0]) % n]:
                want.append((cell, d))
    assert [(int(c), int(d)) for c, d in zip(cells, dirs)] == want


@pytest.mark.parametrize("shape", ["sphere", "ellipsoid", "cylinder", "naca4"])
def test_d3q27_ibb_keeps_a_uniform_flow_moving_with_the_body(shape):
    """f = f_eq(rho, U) is a steady solution when every wall moves at U: the
    interpolated rule on the 26 link directions must reproduce f_eq(U)_opp from
    f_eq(U) for any q; both interpolation branches occur."""
    from lbm3d_bb_amd.bodies import ConstantMotion, naca_params, rotation_matrix
    from oracle import oracle as orc

    U = (0.04, -0.025, 0.015)
    cfg = Config(32, 30, 28, collision="TRT", storage="fp64", omega_plus=1 / 0.7, omega_minus=1 / 1.1,
                 faces=[periodic()] * 6, u0=U, refill=True, solid_bc="IBB", q=27)
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
    assert np.abs(f - _feq27(1.0, np.asarray(U)[:, None])).max() < 1e-13
    assert np.abs(rho - 1.0).max() < 1e-13
    # the links' wall fractions: both branches (q < 1/2, q >= 1/2) present (the
    # mask of step 7 is the pose at t = 6)
    semi3 = tuple(semi) + (semi[0],) * (3 - len(semi))
    body = Body(shape, tuple(np.asarray(law.center0) + 6.0 * np.asarray(U)), semi3, rot=rot0, naca=law.naca)
    cells, dirs = o.links()
    L = (cfg.nx, cfg.ny, cfg.nz)
    qs = []
    for cell, d in zip(cells[::7], dirs[::7]):
        x, y, z = cell % L[0], (cell // L[0]) % L[1], cell // (L[0] * L[1])
        pb = np.array([x + 0.5, y + 0.5, z + 0.5], dtype=np.float32)
        phb = orc.phi(body, pb, (1, 1, 1), L)
        phs = orc.phi(body, pb + C27[d].astype(np.float32), (1, 1, 1), L)
        if phb > 0 and phs < 0:
            qs.append(phb / (phb - phs))
    assert min(qs) < 0.5 <= max(qs)
