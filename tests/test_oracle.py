"""Pins the oracle (CPU) against what the mathematics fixes.

The reference ships no tests, fixtures or golden vectors for this path and its
Julia/CUDA.jl code cannot run here, so every part of the oracle is checked by a
closed form, an exact invariant, a brute-force enumeration or an independent
library routine, each chosen so that a plausible slip (a wrong weight, sign,
direction index, force factor or wall term) fails it.
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from lbm3d_bb_amd.bodies import ConstantMotion, Body, ellipsoid, rotation_matrix, sphere
from lbm3d_bb_amd.lattice import CX, CY, CZ, OPP, W, equilibrium
from lbm3d_bb_amd.solver import Config, no_slip, periodic, pressure, velocity, zero_gradient
from oracle.oracle import OracleSolver, body_at, planted

from cases import periodic_box, rel_err, trt_omegas


# ---------------------------------------------------------------- lattice ----
def test_lattice_moments():
    C = np.stack([CX, CY, CZ], 1).astype(float)
    assert abs(W.sum() - 1.0) < 1e-15
    assert np.allclose(W @ C, 0.0, atol=1e-16)
    assert np.allclose(np.einsum("d,da,db->ab", W, C, C), np.eye(3) / 3.0, atol=1e-16)
    # fourth-order isotropy: sum w c_a c_b c_c c_d = (d_ab d_cd + d_ac d_bd + d_ad d_bc) / 9
    I = np.eye(3)
    iso = (np.einsum("ab,cd->abcd", I, I) + np.einsum("ac,bd->abcd", I, I) + np.einsum("ad,bc->abcd", I, I)) / 9.0
    assert np.allclose(np.einsum("d,da,db,dc,de->abce", W, C, C, C, C), iso, atol=1e-16)
    assert np.array_equal(C[OPP], -C)
    assert np.array_equal(OPP[OPP], np.arange(19))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_oracle_equilibrium_moments(prec):
    cfg = periodic_box(n=6)
    o = OracleSolver(cfg, prec)
    rng = np.random.default_rng(0)
    rho = 1.0 + 0.05 * rng.standard_normal(cfg.ncells)
    u = 0.05 * rng.standard_normal((3, cfg.ncells))
    o.init_equilibrium(rho, u)
    f = o.populations()
    C = np.stack([CX, CY, CZ], 1).astype(float)
    tol = 1e-13 if prec == "fp64" else 2e-6
    assert rel_err(f.sum(0), rho) < tol
    assert np.abs(C.T @ f - rho * u).max() < tol
    # second moment of the quadratic equilibrium: rho (u u + I/3)
    P = np.einsum("da,db,dn->abn", C, C, f)
    Pex = np.einsum("n,an,bn->abn", rho, u, u) + np.eye(3)[:, :, None] * rho / 3.0
    assert np.abs(P - Pex).max() < tol * 10


# ------------------------------------------------------------ invariants ----
@pytest.mark.parametrize("collision", ["BGK", "TRT"])
def test_mass_momentum_conserved_periodic(collision):
    cfg = periodic_box(n=8, collision=collision, g=(0.0, 0.0, 0.0))
    o = OracleSolver(cfg, "fp64")
    rng = np.random.default_rng(1)
    f = equilibrium(np.ones(cfg.ncells), *(0.05 * rng.standard_normal((3, cfg.ncells))))
    f += 1e-3 * rng.standard_normal(f.shape)
    o.set_populations(f)
    m0, p0 = o.conservation()
    o.step(20, forces=False)
    m1, p1 = o.conservation()
    assert abs(m1 - m0) < 1e-12 * m0
    assert np.abs(p1 - p0).max() < 1e-12 * m0


def test_body_force_adds_exact_momentum():
    """Guo forcing adds F = rho g per node and step: d(sum rho u)/dt = M g exactly."""
    g = np.array([1e-4, -2e-4, 5e-5])
    cfg = periodic_box(n=8, collision="TRT", g=tuple(g))
    o = OracleSolver(cfg, "fp64")
    rng = np.random.default_rng(2)
    o.init_equilibrium(1.0 + 0.01 * rng.standard_normal(cfg.ncells), 0.01 * rng.standard_normal((3, cfg.ncells)))
    m0, p0 = o.conservation()
    o.step(25, forces=False)
    m1, p1 = o.conservation()
    assert np.abs((p1 - p0) - 25 * m0 * g).max() < 1e-13 * m0


@pytest.mark.parametrize("collision", ["BGK", "TRT"])
def test_uniform_flow_is_fixed_point(collision):
    cfg = periodic_box(n=6, collision=collision, g=(0.0, 0.0, 0.0))
    cfg.u0 = (0.03, -0.02, 0.01)
    o = OracleSolver(cfg, "fp64")
    o.init_equilibrium()
    f0 = o.populations()
    o.step(10, forces=False)
    assert rel_err(o.populations(), f0) < 1e-15


@pytest.mark.parametrize("outlet", ["pressure", "zero_gradient"])
def test_inlet_outlet_uniform_flow_fixed_point(outlet):
    """Uniform flow through a velocity inlet (src/OtherBC.jl:138-149) and an
    anti-bounce-back (src/OtherBC.jl:152-166) or zero-gradient outlet stays exact."""
    U = 0.04
    wp, wm = trt_omegas(0.7)
    cfg = Config(12, 6, 6, collision="TRT", storage="fp64", omega_plus=wp, omega_minus=wm,
                 faces=[velocity((U, 0, 0)), pressure() if outlet == "pressure" else zero_gradient(),
                        periodic(), periodic(), periodic(), periodic()], u0=(U, 0, 0))
    o = OracleSolver(cfg, "fp64")
    o.init_equilibrium()
    f0 = o.populations()
    o.step(15, forces=False)
    assert rel_err(o.populations(), f0) < 1e-14


# --------------------------------------------------------- closed forms ----
def test_poiseuille_exact_trt_magic():
    """TRT with Lambda = 3/16, Guo forcing and halfway bounce-back walls: the
    discrete steady state is the parabola u = g y (H - y) / (2 nu) exactly."""
    H, gx, tau = 10, 1e-5, 0.8
    wp, wm = trt_omegas(tau)
    nu = (tau - 0.5) / 3.0
    cfg = Config(3, H, 3, collision="TRT", storage="fp64", omega_plus=wp, omega_minus=wm,
                 faces=[periodic(), periodic(), no_slip(), no_slip(), periodic(), periodic()], g=(gx, 0, 0))
    o = OracleSolver(cfg, "fp64")
    o.init_equilibrium()
    o.step(6000, forces=False)
    _, u = o.macros()
    ux = u[0].reshape(3, H, 3)[1, :, 1]
    y = np.arange(H) + 0.5
    exact = gx * y * (H - y) / (2.0 * nu)
    assert rel_err(ux, exact) < 1e-9
    assert np.abs(u[1:]).max() < 1e-15


def test_couette_exact_moving_wall():
    """Linear shear between a wall at rest and one moving at U (moving-wall bounce-back)."""
    H, U = 8, 0.02
    wp, wm = trt_omegas(0.9)
    cfg = Config(3, H, 3, collision="TRT", storage="fp64", omega_plus=wp, omega_minus=wm,
                 faces=[periodic(), periodic(), no_slip(), velocity((U, 0, 0)), periodic(), periodic()])
    o = OracleSolver(cfg, "fp64")
    o.init_equilibrium()
    o.step(4000, forces=False)
    _, u = o.macros()
    ux = u[0].reshape(3, H, 3)[1, :, 1]
    exact = U * (np.arange(H) + 0.5) / H
    assert rel_err(ux, exact) < 1e-9


def test_shear_wave_decay_rate():
    """u_x = A sin(k y) decays as exp(-nu k^2 t), nu = (tau - 1/2)/3 (pins viscosity and streaming)."""
    L, tau, A = 64, 0.9, 1e-4
    cfg = Config(4, L, 4, collision="BGK", storage="fp64", omega_plus=1 / tau, omega_minus=1 / tau,
                 faces=[periodic()] * 6)
    o = OracleSolver(cfg, "fp64")
    k = 2 * np.pi / L
    y = (np.arange(cfg.ncells) // 4) % L
    u = np.zeros((3, cfg.ncells))
    u[0] = A * np.sin(k * (y + 0.5))
    o.init_equilibrium(np.ones(cfg.ncells), u)
    nsteps = 1600
    o.step(nsteps, forces=False)
    _, u1 = o.macros()
    ratio = np.abs(u1[0]).max() / A
    nu = (tau - 0.5) / 3.0
    assert abs(ratio - np.exp(-nu * k * k * nsteps)) / np.exp(-nu * k * k * nsteps) < 5e-3


# --------------------------------------------------------------- geometry ----
def _numpy_flags(bodies, n, per=(True, True, True)):
    """Brute-force FP32 inside test with numpy, same operation order."""
    L = np.array([n, n, n])
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    pos = [(x.astype(np.float32) + np.float32(0.5)), (y.astype(np.float32) + np.float32(0.5)),
           (z.astype(np.float32) + np.float32(0.5))]
    flags = np.zeros((n, n, n), dtype=np.uint8)
    for bi, b in reversed(list(enumerate(bodies))):
        d = []
        for a in range(3):
            da = (pos[a] - np.float32(b.center[a])).astype(np.float32)
            if per[a]:
                h = np.float32(0.5) * np.float32(L[a])
                da = np.where(da > h, da - np.float32(L[a]), np.where(da < -h, da + np.float32(L[a]), da))
            d.append(da.astype(np.float32))
        R = np.float32(b.semi[0])
        r2 = np.float32(R * R)
        s = (d[0] * d[0] + d[1] * d[1]).astype(np.float32)
        s = (s + d[2] * d[2]).astype(np.float32)
        inside = s <= np.float32(r2 + np.float32(1e-6) * r2)
        flags[inside] = bi + 1
    return flags.reshape(-1)


def _numpy_links(flags, n):
    f3 = flags.reshape(n, n, n)
    cells, dirs = [], []
    fluid = f3 == 0
    hit = np.zeros((19, n, n, n), dtype=bool)
    for d in range(19):
        nb = np.roll(f3, shift=(-CZ[d], -CY[d], -CX[d]), axis=(0, 1, 2))
        hit[d] = fluid & (nb != 0)
    idx = np.nonzero(hit.reshape(19, -1).T)
    return idx[0].astype(np.int32), idx[1].astype(np.int8)


def test_sphere_flags_and_links_brute_force():
    n = 20
    bodies = [sphere((7.3, 9.1, 10.6), 4.2), sphere((16.5, 3.2, 18.9), 3.1)]
    o = OracleSolver(periodic_box(n=n), "fp64")
    o.set_bodies(bodies)
    o.update_geometry()
    flags = o.flags()
    assert np.array_equal(flags, _numpy_flags(bodies, n))
    c, d = o.links()
    c2, d2 = _numpy_links(flags, n)
    assert np.array_equal(c, c2) and np.array_equal(d, d2)


def test_sphere_cell_count_closed_form():
    """Centre on a lattice node corner: offsets are half-integers, so the count is the
    number of odd (i, j, k) with i^2 + j^2 + k^2 <= 4 R^2 (integer arithmetic)."""
    n, R = 24, 5
    o = OracleSolver(periodic_box(n=n), "fp64")
    o.set_bodies([sphere((12.0, 12.0, 12.0), float(R))])
    o.update_geometry()
    odd = np.arange(-2 * R - 1, 2 * R + 2, 2)
    I, J, K = np.meshgrid(odd, odd, odd, indexing="ij")
    assert int((o.flags() != 0).sum()) == int(((I * I + J * J + K * K) <= 4 * R * R).sum())


def test_ellipsoid_axis_aligned_matches_formula():
    n = 24
    semi = (7.0, 4.0, 3.0)
    o = OracleSolver(periodic_box(n=n), "fp64")
    o.set_bodies([ellipsoid((12.0, 12.0, 12.0), semi)])
    o.update_geometry()
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    q = ((x + 0.5 - 12) / semi[0]) ** 2 + ((y + 0.5 - 12) / semi[1]) ** 2 + ((z + 0.5 - 12) / semi[2]) ** 2
    ref = (q <= 1.0).reshape(-1)
    # cells within 1e-5 of the surface may differ by the FP32 tolerance; there are none here
    assert np.abs(q - 1.0).min() > 1e-5
    assert np.array_equal(o.flags() != 0, ref)


def test_motion_law_pose():
    ax = np.array([0.2, -0.7, 0.4])
    om = 0.013 * ax / np.linalg.norm(ax)
    R0 = rotation_matrix((1, 0, 0), 0.3)
    law = ConstantMotion("ellipsoid", (10.0, 11.0, 12.0), (5.0, 3.0, 2.0), vel=(0.01, -0.02, 0.005),
                         omega=tuple(om), rot0=R0)
    for n in (1, 7, 250):
        b = body_at(law, n)
        t = n - 1
        assert np.allclose(list(b.center), [10 + 0.01 * t, 11 - 0.02 * t, 12 + 0.005 * t], atol=1e-5)
        assert np.allclose(list(b.center_half), [10 + 0.01 * (t + .5), 11 - 0.02 * (t + .5), 12 + 0.005 * (t + .5)])
        Rex = Rotation.from_rotvec(om * t).as_matrix() @ R0
        assert np.abs(np.array(list(b.rot)).reshape(3, 3) - Rex).max() < 1e-6


# --------------------------------------------------- momentum exchange ----
def test_mea_single_cell_uniform_flow():
    """One solid cell in uniform equilibrium flow: F = 2 rho u at the first step."""
    n, u0 = 8, np.array([0.03, -0.01, 0.02])
    cfg = periodic_box(n=n, g=(0, 0, 0))
    cfg.u0 = tuple(u0)
    o = OracleSolver(cfg, "fp64")
    flags = np.zeros(n ** 3, dtype=np.uint8)
    flags[4 + n * (4 + n * 4)] = 1
    o.set_flags(flags)
    o.init_equilibrium()
    F = o.step(1)
    assert np.abs(F[0, 0, :3] - 2.0 * u0).max() < 1e-15


def test_mea_single_cell_moving_wall():
    """One solid cell moving at u_w in fluid at rest: F = -2 rho u_w."""
    n, uw = 9, np.array([0.02, 0.01, -0.03])
    o = OracleSolver(periodic_box(n=n, g=(0, 0, 0)), "fp64")
    # a sphere of radius 0.4 centred on a node covers exactly that node
    o.set_bodies([Body("sphere", (4.5, 4.5, 4.5), (0.4, 0.4, 0.4), vel=tuple(uw))])
    o.init_equilibrium()
    F = o.step(1)
    assert (o.flags() != 0).sum() == 1
    assert np.abs(F[0, 0, :3] + 2.0 * uw).max() < 1e-15


# A solid cell at s (a sphere of radius 0.4 on the node covers it alone) whose
# half-step centre X differs from s, translating at V and spinning at Omega, in
# fluid at rest (rho = 1).  At the first step every link term is
# t_d = 2 w_d rho - 6 w_d rho c_d.u_w(x_w), u_w = V + Omega x (x_w - X), with the
# wall point x_w = s - (1 - h) c_d on the link; c_d.(Omega x c_d) = 0, so
# c_d.u_w = c_d.U with U = V + Omega x (s - X), and sum_d w_d c_d c_d = I/3 gives
#     F = -2 rho U,        T = sum_d (x_w - X) x c_d t_d = (s - X) x F
# (the arm's part along c_d drops out of c_d x c_d).  A planted sign flip of
# Omega x r moves F, a flipped arm moves T; a mirrored arm (x - h c_d, the
# reference's wall sampling point, src/SolidBC.jl:195-196) provably does not.
_S = np.array([4.5, 4.5, 4.5])
_X = np.array([4.8, 4.3, 5.2])
_V = np.array([0.011, -0.007, 0.004])
_OM = np.array([0.003, 0.005, -0.002])


def _spinning_cell_step(q, solid_bc="SBB", wall_point="physical", X=_X):
    cfg = periodic_box(n=9, g=(0, 0, 0))
    cfg.q, cfg.solid_bc, cfg.wall_point = q, solid_bc, wall_point
    o = OracleSolver(cfg, "fp64")
    o.set_bodies([Body("sphere", tuple(_S), (0.4, 0.4, 0.4), center_half=tuple(X), vel=tuple(_V),
                       omega=tuple(_OM))])
    o.init_equilibrium()
    F = o.step(1)[0, 0]
    assert (o.flags() != 0).sum() == 1
    return F[:3], F[3:]


@pytest.mark.parametrize("q", [19, 27])
def test_mea_moving_spinning_cell_force_and_torque_closed_form(q):
    U = _V + np.cross(_OM, _S - _X)
    Fex = -2.0 * U
    F, T = _spinning_cell_step(q)
    assert np.abs(F - Fex).max() < 1e-15
    assert np.abs(T - np.cross(_S - _X, Fex)).max() < 1e-15
    # the reference's mirrored wall point gives the same force and torque
    Fr, Tr = _spinning_cell_step(q, wall_point="reference")
    assert np.abs(Fr - Fex).max() < 1e-15 and np.abs(Tr - T).max() < 1e-15


@pytest.mark.parametrize("q", [19, 27])
def test_mea_spinning_cell_pins_fail_on_planted_mistakes(q):
    U = _V + np.cross(_OM, _S - _X)
    Fex = -2.0 * U
    Tex = np.cross(_S - _X, Fex)
    with planted("spin_sign"):
        F, T = _spinning_cell_step(q)
    assert np.abs(F - Fex).max() > 1e-3 * np.abs(Fex).max()
    with planted("arm_sign"):
        F, T = _spinning_cell_step(q)
    assert np.abs(F - Fex).max() < 1e-15 and np.abs(T - Tex).max() > 0.5 * np.abs(Tex).max()
    with planted("wall_half"):
        F, T = _spinning_cell_step(q)
    assert np.abs(F - Fex).max() > 0.4 * np.abs(Fex).max()
    # a mirrored torque arm is invisible: the arm's change is parallel to c_d
    with planted("arm_mirror"):
        F, T = _spinning_cell_step(q)
    assert np.abs(T - Tex).max() < 1e-15
    F, T = _spinning_cell_step(q)
    assert np.abs(F - Fex).max() < 1e-15 and np.abs(T - Tex).max() < 1e-15


@pytest.mark.parametrize("q", [19, 27])
def test_mea_spinning_cell_interpolated_closed_form(q):
    """Interpolated bounce-back on the same cell, centred on the node (the level
    set |x - X| - 0.4 is measured from the half-step centre X = s): every link
    is on the q >= 1/2 branch with q = (|c| - 0.4)/|c|, so at rest
    t_d = 2 w_d rho - 6 w_d rho c_d.V / (2 q_d) and
    F = -3 rho sum_d w_d c_d c_d V / q_d, i.e. -(rho/3)(1/q_a + 2/q_e) V on
    D3Q19 (axis and edge directions) and -(rho/9)(4/q_a + 4/q_e + 1/q_c) V on
    D3Q27 (with the corners); the spin adds nothing (Omega x (s - X) = 0) and
    T = 0.  q = 1/2 everywhere would give the simple rule's -2 rho V."""
    qa, qe, qc = 0.6, 1.0 - 0.4 / np.sqrt(2.0), 1.0 - 0.4 / np.sqrt(3.0)
    k = (1.0 / qa + 2.0 / qe) / 3.0 if q == 19 else (4.0 / qa + 4.0 / qe + 1.0 / qc) / 9.0
    Fex = -k * _V
    F, T = _spinning_cell_step(q, solid_bc="IBB", X=_S)
    # q is formed from the FP32 level set: relative rounding ~1e-7
    assert np.abs(F - Fex).max() < 1e-6 * np.abs(Fex).max()
    assert np.abs(T).max() < 1e-15
    with planted("wall_half"):
        F, T = _spinning_cell_step(q, solid_bc="IBB", X=_S)
    assert np.abs(F - Fex).max() > 0.4 * np.abs(Fex).max()


@pytest.mark.parametrize("q", [19, 27])
def test_mea_static_cell_uniform_flow_torque(q):
    """A static solid cell at s in uniform equilibrium flow u: F = 2 rho u and
    T = (s - X) x F about a half-step centre X != s."""
    u0 = np.array([0.03, -0.01, 0.02])
    cfg = periodic_box(n=9, g=(0, 0, 0))
    cfg.q, cfg.u0 = q, tuple(u0)
    o = OracleSolver(cfg, "fp64")
    o.set_bodies([Body("sphere", tuple(_S), (0.4, 0.4, 0.4), center_half=tuple(_X))])
    o.init_equilibrium()
    F = o.step(1)[0, 0]
    assert np.abs(F[:3] - 2.0 * u0).max() < 1e-15
    assert np.abs(F[3:] - np.cross(_S - _X, 2.0 * u0)).max() < 1e-15


def test_mea_momentum_budget():
    """Fluid momentum changes by sum(rho g) - F_MEA each step (static bodies, FP64)."""
    n = 16
    g = np.array([2e-5, 1e-5, 0.0])
    cfg = periodic_box(n=n, collision="TRT", g=tuple(g))
    o = OracleSolver(cfg, "fp64")
    o.set_bodies([sphere((5.2, 8.1, 7.7), 3.0), sphere((12.5, 11.0, 4.0), 2.2)])
    o.init_equilibrium()
    o.step(5)
    m0, p0 = o.conservation()
    F = o.step(1)
    m1, p1 = o.conservation()
    assert abs(m1 - m0) < 1e-13 * m0
    Ftot = F[0, :, :3].sum(0)
    # exact identity up to the rounding of the FP64 sums over cells (|p| ~ 1e-1 here)
    assert np.abs(Ftot).max() > 1e-3
    assert np.abs((p1 - p0) - (m0 * g - Ftot)).max() < 1e-13 * (np.abs(p0).max() + m0 * np.abs(g).max() + 1.0)


def test_refill_is_equilibrium_of_neighbour_average():
    n = 14
    cfg = periodic_box(n=n, collision="TRT", g=(0, 0, 0))
    cfg.refill = True
    cfg.u0 = (0.01, 0.0, 0.0)
    o = OracleSolver(cfg, "fp64")
    law = ConstantMotion("sphere", (6.3, 7.0, 7.0), (3.0,), vel=(0.25, 0.0, 0.0))
    o.set_motion([law])
    o.init_equilibrium()
    o.step(3)
    flags_before = o.flags().copy()
    rho0, u0 = o.macros()
    o.step(1)
    flags_after = o.flags()
    fresh = np.nonzero((flags_before != 0) & (flags_after == 0))[0]
    assert len(fresh) > 0
    f = o.populations()
    for c in fresh:
        x, y, z = c % n, (c // n) % n, c // (n * n)
        acc_r, acc_u, cnt = 0.0, np.zeros(3), 0
        for d in range(18):
            q = ((x + CX[d]) % n, (y + CY[d]) % n, (z + CZ[d]) % n)
            nc = q[0] + n * (q[1] + n * q[2])
            if flags_after[nc] == 0 and not (flags_before[nc] != 0):
                acc_r += rho0[nc]
                acc_u += u0[:, nc]
                cnt += 1
        feq = equilibrium(acc_r / cnt, *(acc_u / cnt))
        assert np.abs(f[:, c] - feq).max() < 1e-15
