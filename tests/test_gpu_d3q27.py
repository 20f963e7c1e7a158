"""D3Q27 (src/Utilities.jl:962-976): HIP path vs the oracle (GPU).

Same bars as the D3Q19 parity tests (tests/cases.py): flags and the
26-direction link list bit-exact; populations, rho and u within 1e-12 (FP64)
or 1e-5 (FP32), normwise; per-body force and torque within the same tolerance
relative to max|F_oracle| plus the rounding floor of the link sum.  The
oracle's D3Q27 step is pinned separately (tests/test_d3q27_oracle.py).
"""
from __future__ import annotations

import numpy as np
import pytest

from lbm3d_bb_amd.bodies import ConstantMotion, ellipsoid, rotation_matrix, sphere
from lbm3d_bb_amd.lattice import C27, W27
from lbm3d_bb_amd.solver import Solver
from oracle.oracle import OracleSolver

from cases import TOL, channel, check_forces, check_state, periodic_box, reference_channel, rel_err, sphere_pack

pytestmark = pytest.mark.gpu

def _q27(cfg):
    cfg.q = 27
    return cfg


def _pair(cfg, prec):
    return Solver(cfg), OracleSolver(cfg, prec)


def _compare_state(g, o, tol, what=""):
    check_state(g, o, "fp64" if tol < 1e-6 else "fp32", what, tol=tol)


def test_d3q27_flags_and_links_bit_exact():
    cfg = _q27(channel(nx=40, ny=26, nz=22, storage="fp32"))
    bodies = [sphere((12.3, 13.1, 10.8), 5.2), ellipsoid((27.0, 12.0, 11.0), (6.0, 4.0, 3.0),
                                                         rot=rotation_matrix((1, 2, 0), 0.5))]
    g, o = _pair(cfg, "fp32")
    for s in (g, o):
        s.set_bodies(bodies)
        s.update_geometry()
    assert np.array_equal(g.flags(), o.flags())
    cg, dg = g.links()
    co, do = o.links()
    assert len(co) > 0 and int(do.max()) > 17  # edge and corner links present
    assert np.array_equal(cg, co) and np.array_equal(dg, do)
    # every link points at a solid node from a fluid node
    fl = o.flags()
    x, y, z = cg % 40, (cg // 40) % 26, cg // (40 * 26)
    c = C27[dg.astype(np.int64)]
    assert (fl[cg] == 0).all()
    assert (fl[(x + c[:, 0]) + 40 * ((y + c[:, 1]) + 26 * (z + c[:, 2]))] != 0).all()


def test_d3q27_start_state_and_raw_getters():
    cfg = _q27(periodic_box(n=12, collision="TRT"))
    g, o = _pair(cfg, "fp64")
    rng = np.random.default_rng(3)
    rho = 1.0 + 0.02 * rng.standard_normal(cfg.ncells)
    u = 0.03 * rng.standard_normal((3, cfg.ncells))
    for s in (g, o):
        s.init_equilibrium(rho, u)
    _compare_state(g, o, 1e-15, "init")
    f = o.populations() + 1e-4 * rng.standard_normal((27, cfg.ncells))
    g.set_populations(f)
    assert np.array_equal(g.populations(), f)


@pytest.mark.parametrize("collision", ["BGK", "TRT"])
def test_d3q27_periodic_sphere_forced_fp64(collision):
    cfg = _q27(periodic_box(n=24, collision=collision, g=(2e-5, -1e-5, 5e-6)))
    g, o = _pair(cfg, "fp64")
    for s in (g, o):
        s.set_bodies([sphere((11.7, 12.2, 12.9), 5.3)])
        s.init_equilibrium()
    Fg, Fo = g.step(40), o.step(40)
    assert np.abs(Fo).max() > 0
    check_forces(Fg, Fo, o, "fp64", collision, shapes=[sphere((11.7, 12.2, 12.9), 5.3)])
    _compare_state(g, o, TOL["fp64"], collision)


def test_d3q27_multi_body_wrapping_fp64():
    """Several spheres, some across the periodic boundaries, over more than one chunk of steps."""
    cfg = _q27(periodic_box(n=28, collision="TRT"))
    bodies = sphere_pack(6, 28, 2.5, 3.5, seed=4)
    assert any(min(b.center[a], 28 - b.center[a]) < b.semi[0] for b in bodies for a in range(3))
    g, o = _pair(cfg, "fp64")
    for s in (g, o):
        s.set_bodies(bodies)
        s.init_equilibrium()
    Fg, Fo = g.step(70), o.step(70)
    assert Fg.shape == (70, len(bodies), 6)
    check_forces(Fg, Fo, o, "fp64", "multi", shapes=bodies)
    _compare_state(g, o, TOL["fp64"], "multi")


def test_d3q27_channel_translating_sphere_refill_fp64():
    """Velocity inlet, zero-gradient outlet, no-slip walls (edges and corners of
    the 27 directions), a translating sphere with fresh-node refill."""
    cfg = _q27(channel(nx=40, ny=20, nz=20, storage="fp64"))
    g, o = _pair(cfg, "fp64")
    law = ConstantMotion("sphere", (12.3, 10.1, 9.8), (4.0,), vel=(0.05, 0.0, 0.0))
    for s in (g, o):
        s.set_motion([law])
        s.init_equilibrium()
    Fg, Fo = g.step(50), o.step(50)
    check_forces(Fg, Fo, o, "fp64", "channel", shapes=[law])
    _compare_state(g, o, TOL["fp64"], "channel")
    assert np.array_equal(g.flags(), o.flags())


def test_d3q27_reference_face_set_moving_wall_fp64():
    """The reference's faces (velocity inlet and y-walls at U_inf, anti-bounce-back
    outlet, z periodic), its wall sampling point, a sphere moving in y and spinning."""
    cfg = _q27(reference_channel(nx=40, ny=24, nz=8, storage="fp64"))
    g, o = _pair(cfg, "fp64")
    law = ConstantMotion("sphere", (14.0, 12.0, 4.0), (3.5,), vel=(0.0, 0.02, 0.0), omega=(0.0, 0.0, 0.01))
    for s in (g, o):
        s.set_motion([law])
        s.init_equilibrium()
    Fg, Fo = g.step(40), o.step(40)
    assert np.abs(Fo[:, :, 5]).max() > 0
    check_forces(Fg, Fo, o, "fp64", "refchannel", shapes=[law])
    _compare_state(g, o, TOL["fp64"], "refchannel")


def test_d3q27_rotating_ellipsoid_refill_fp64():
    cfg = _q27(channel(nx=36, ny=28, nz=28, storage="fp64", u_in=0.02))
    g, o = _pair(cfg, "fp64")
    law = ConstantMotion("ellipsoid", (17.0, 14.0, 14.0), (8.0, 4.0, 4.0), omega=(0.0, 0.02, 0.03),
                         rot0=rotation_matrix((1, 1, 0), 0.4))
    for s in (g, o):
        s.set_motion([law])
        s.init_equilibrium()
    Fg, Fo = g.step(30), o.step(30)
    assert np.abs(Fo[:, :, 3:]).max() > 0
    check_forces(Fg, Fo, o, "fp64", "ellipsoid", shapes=[law])
    _compare_state(g, o, TOL["fp64"], "ellipsoid")
    assert np.array_equal(g.flags(), o.flags())


def test_d3q27_channel_translating_sphere_fp32():
    cfg = _q27(channel(nx=40, ny=20, nz=20, storage="fp32"))
    g, o = _pair(cfg, "fp32")
    law = ConstantMotion("sphere", (12.3, 10.1, 9.8), (4.0,), vel=(0.01, 0.0, 0.0))
    for s in (g, o):
        s.set_motion([law])
        s.init_equilibrium()
    Fg, Fo = g.step(30), o.step(30)
    check_forces(Fg, Fo, o, "fp32", "channel fp32", shapes=[law])
    _compare_state(g, o, TOL["fp32"], "fp32")


def test_d3q27_explicit_flags_fp64():
    """Geometry from lbm_set_flags (no analytic extent): every cell checks its flags."""
    cfg = _q27(periodic_box(n=16, collision="BGK"))
    o = OracleSolver(cfg, "fp64")
    o.set_bodies([sphere((8.2, 7.9, 8.4), 4.1)])
    o.update_geometry()
    flags = o.flags()
    g = Solver(cfg)
    g.set_flags(flags)
    o2 = OracleSolver(cfg, "fp64")
    o2.set_flags(flags)
    for s in (g, o2):
        s.init_equilibrium()
    Fg, Fo = g.step(20), o2.step(20)
    check_forces(Fg, Fo, o2, "fp64", "flags")
    _compare_state(g, o2, TOL["fp64"], "flags")


def test_d3q27_forces_deterministic_and_mass_conserved():
    cfg = _q27(periodic_box(n=24, collision="TRT", g=(1e-4, 0.0, 0.0)))
    bodies = sphere_pack(4, 24, 2.5, 3.5, seed=2)
    runs = []
    for _ in range(2):
        g = Solver(cfg)
        g.set_bodies(bodies)
        g.init_equilibrium()
        runs.append(g.step(20))
        g.close()
    assert np.array_equal(runs[0], runs[1])
    g = Solver(cfg)
    g.init_equilibrium()
    m0, p0 = g.conservation()
    g.step(50)
    m1, p1 = g.conservation()
    assert abs(m1 - m0) / m0 < 1e-13
    assert abs((p1[0] - p0[0]) - 50 * m0 * 1e-4) < 1e-10 * m0


def test_d3q27_moving_bodies_across_periodic_boundary_fp64():
    """Two moving bodies, one crossing the periodic x boundary: the moved flags are
    rebuilt in boxes around each body's old and new pose (wrapping)."""
    cfg = _q27(periodic_box(n=24, collision="TRT", g=(1e-5, 0.0, 0.0)))
    laws = [ConstantMotion("sphere", (22.6, 11.3, 12.4), (3.2,), vel=(0.06, 0.0, 0.01)),
            ConstantMotion("ellipsoid", (9.0, 12.0, 11.0), (5.0, 3.0, 2.5), vel=(0.0, -0.02, 0.0),
                           omega=(0.01, 0.0, 0.02), rot0=rotation_matrix((0, 1, 1), 0.3))]
    g, o = _pair(cfg, "fp64")
    for s in (g, o):
        s.set_motion(laws)
        s.init_equilibrium()
    Fg, Fo = g.step(70), o.step(70)
    assert np.array_equal(g.flags(), o.flags())
    check_forces(Fg, Fo, o, "fp64", "moving-wrap", shapes=laws)
    _compare_state(g, o, TOL["fp64"], "moving-wrap")


# --- interpolated bounce-back on the D3Q27 links ---------------------------------------

def _feq27(U):
    cu = C27.astype(np.float64) @ np.asarray(U, dtype=np.float64)
    return W27 * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * np.dot(U, U))


@pytest.mark.parametrize("shape", ["sphere", "ellipsoid", "cylinder", "naca4"])
def test_d3q27_ibb_uniform_flow_moving_with_the_body(shape):
    """f_eq(U) stays a steady solution when the wall moves at U (every q, both branches)."""
    from lbm3d_bb_amd.bodies import naca_params
    from lbm3d_bb_amd.solver import Config, periodic

    U = (0.04, -0.025, 0.015)
    cfg = Config(32, 30, 28, collision="TRT", storage="fp64", omega_plus=1 / 0.7, omega_minus=1 / 1.1,
                 faces=[periodic()] * 6, u0=U, refill=True, solid_bc="IBB", q=27)
    semi = {"sphere": (6.3,), "ellipsoid": (7.7, 4.2, 3.6), "cylinder": (6.3,), "naca4": (18.0,)}[shape]
    rot0 = rotation_matrix((1, 2, 3), 0.7) if shape == "ellipsoid" else (
        rotation_matrix((0, 0, 1), 0.21) if shape == "naca4" else np.eye(3))
    law = ConstantMotion(shape, (15.3, 14.8, 13.9), semi, vel=U, rot0=rot0,
                         naca=naca_params(18.0, "2412") if shape == "naca4" else (0.0,) * 5)
    g, o = _pair(cfg, "fp64")
    for s in (g, o):
        s.set_motion([law])
        s.init_equilibrium()
    Fg, Fo = g.step(6), o.step(6)
    assert np.abs(g.populations() - _feq27(U)[:, None]).max() < 1e-13
    assert np.array_equal(g.flags(), o.flags())
    check_forces(Fg, Fo, o, "fp64", shape, shapes=[law])


@pytest.mark.parametrize("collision", ["BGK", "TRT"])
def test_d3q27_ibb_static_cylinder_reference_channel_fp64(collision):
    cfg = _q27(reference_channel(nx=64, ny=40, nz=4, storage="fp64"))
    cfg.collision = collision
    cfg.solid_bc = "IBB"
    g, o = _pair(cfg, "fp64")
    from lbm3d_bb_amd.bodies import cylinder

    for s in (g, o):
        s.set_bodies([cylinder((20.3, 19.6, 2.0), 6.37)])
        s.init_equilibrium()
    Fg, Fo = g.step(60), o.step(60)
    assert np.abs(Fo).max() > 0
    check_forces(Fg, Fo, o, "fp64", collision, shapes=[cylinder((20.3, 19.6, 2.0), 6.37)])
    _compare_state(g, o, TOL["fp64"], collision)


def test_d3q27_ibb_heaving_cylinder_from_params_fp64():
    """The reference's own input file with q = 27 (faces, BGK, IBB, heave law) on a
    small lattice, the heave quickened so that the mask moves within the run."""
    import dataclasses
    import os

    from lbm3d_bb_amd.params import config_from_params, read_params_input

    here = os.path.dirname(os.path.abspath(__file__))
    p = read_params_input(os.path.join(here, "golden", "params_reference_default.txt"))
    p = dataclasses.replace(p, Nx=120, Ny=80, Nz=3, R=6.3, x_c=40.5, y_c=40.0, Re=20.0, q=27)
    cfg, laws = config_from_params(p, storage="fp64", refill=True)
    assert cfg.q == 27
    laws[0].freq = 0.05
    g, o = _pair(cfg, "fp64")
    for s in (g, o):
        s.set_motion(laws)
        s.init_equilibrium()
    Fg, Fo = g.step(40), o.step(40)
    assert np.array_equal(g.flags(), o.flags())
    cg, dg = g.links()
    co, do = o.links()
    assert np.array_equal(cg, co) and np.array_equal(dg, do)
    check_forces(Fg, Fo, o, "fp64", "heave", shapes=laws)
    _compare_state(g, o, TOL["fp64"], "heave")


@pytest.mark.parametrize("wall_point", ["physical", "reference"])
def test_d3q27_ibb_moving_airfoil_and_spinning_ellipsoid_fp32(wall_point):
    from lbm3d_bb_amd.bodies import naca_params, rotation_z

    cfg = _q27(channel(nx=64, ny=40, nz=36, storage="fp32"))
    cfg.solid_bc = "IBB"
    cfg.wall_point = wall_point
    laws = [ConstantMotion("naca4", (24.3, 20.2, 18.0), (22.0,), vel=(0.0, 0.01, 0.0), rot0=rotation_z(0.12),
                           naca=naca_params(22.0, "2412")),
            ConstantMotion("ellipsoid", (46.0, 20.0, 18.0), (6.0, 3.5, 3.0), omega=(0.01, 0.02, 0.015),
                           rot0=rotation_matrix((1, 0, 1), 0.5))]
    g, o = _pair(cfg, "fp32")
    for s in (g, o):
        s.set_motion(laws)
        s.init_equilibrium()
    Fg, Fo = g.step(30), o.step(30)
    check_forces(Fg, Fo, o, "fp32", wall_point, shapes=laws)
    _compare_state(g, o, TOL["fp32"], "fp32")


# --- full size -----------------------------------------------------------------------

def test_d3q27_periodic_box_64_bgk_sphere_full_size_fp64():
    """64^3 periodic box, BGK tau = 0.8, FP64, body force 1e-6 along x, fixed sphere
    R = 8 at the centre, 300 steps: the drag every 10 steps and the final state
    against the FP64 oracle."""
    cfg = _q27(periodic_box(n=64, collision="BGK", storage="fp64", tau=0.8, g=(1e-6, 0.0, 0.0)))
    g, o = _pair(cfg, "fp64")
    for s in (g, o):
        s.set_bodies([sphere((32.0, 32.0, 32.0), 8.0)])
        s.init_equilibrium()
    Fg, Fo = g.step(300), o.step(300)
    assert np.abs(Fo[9::10, 0, 0]).min() > 0
    check_forces(Fg[9::10], Fo[9::10], o, "fp64", "box64", shapes=[sphere((32.0, 32.0, 32.0), 8.0)])
    _compare_state(g, o, TOL["fp64"], "box64")


def test_d3q27_channel256_full_size_fp32_against_fp64():
    """The 256x128x128 D3Q27 channel with the translating sphere R = 12 at its full
    size, where the oracle is too slow: the FP32 run keeps the FP64 run's flags and
    links bit for bit, its drag to FP32 rounding, and repeats bitwise."""
    import bench

    runs = {}
    for storage in ("fp64", "fp32", "fp32"):
        w = dict(bench.WORKLOADS["channel256_q27"], storage=storage)
        cfg, laws, _, _ = bench.make_case(w)
        s = Solver(cfg)
        s.set_motion(laws)
        s.init_equilibrium()
        F = s.step(150)
        runs.setdefault(storage, []).append((F, s.flags(), s.links()))
        s.close()
    (F64, fl64, (c64, d64)), = runs["fp64"]
    (F32, fl32, (c32, d32)), (F32b, _, _) = runs["fp32"]
    assert np.array_equal(fl64, fl32) and np.array_equal(c64, c32) and np.array_equal(d64, d32)
    assert np.array_equal(F32, F32b)
    drag64, drag32 = F64[-50:, 0, 0], F32[-50:, 0, 0]
    assert np.abs(drag32 - drag64).max() < 1e-4 * np.abs(drag64).max(), np.abs(drag32 - drag64).max()


# --- 16-bit population storage -----------------------------------------------------------

def _mean_drag(F, last=100):
    return np.asarray(F)[-last:, :, 0].mean(axis=0)


@pytest.mark.parametrize("storage,tol", [("fp16s", 5e-3), ("bf16s", 5e-2)])
def test_d3q27_shifted_storage_drag_vs_fp64(storage, tol):
    """Periodic box, body force, fixed sphere: the 16-bit storages' drag against the
    FP64 oracle (the same bars as on D3Q19)."""
    cfg64 = _q27(periodic_box(n=24, collision="TRT", storage="fp64", g=(1e-5, 0, 0)))
    cfgs = _q27(periodic_box(n=24, collision="TRT", storage=storage, g=(1e-5, 0, 0)))
    o = OracleSolver(cfg64, "fp64")
    g = Solver(cfgs)
    for s in (g, o):
        s.set_bodies([sphere((12.0, 12.0, 12.0), 4.0)])
        s.init_equilibrium()
    Fg, Fo = g.step(200), o.step(200)
    err = abs(Fg[-1, 0, 0] - Fo[-1, 0, 0]) / abs(Fo[-1, 0, 0])
    print(storage, "drag error vs FP64", err)
    assert err < tol, err


def test_d3q27_shifted_storage_channel_drag_vs_fp64():
    """Channel (inlet, zero-gradient outlet, no-slip walls) with a translating
    sphere: mean drag of the last 100 of 1000 steps, FP16S and BF16S against the
    FP64 oracle; the simple and the interpolated bounce-back."""
    law = ConstantMotion("sphere", (20.0, 16.0, 16.0), (5.0,), vel=(0.01, 0.0, 0.0))
    for bc in ("SBB", "IBB"):
        cfg = _q27(channel(nx=64, ny=32, nz=32, storage="fp64", tau=0.6))
        cfg.solid_bc = bc
        o = OracleSolver(cfg, "fp64")
        o.set_motion([law])
        o.init_equilibrium()
        ref = _mean_drag(o.step(1000))
        o.close()
        for storage, tol in (("fp16s", 5e-3), ("bf16s", 5e-2)):
            cfg = _q27(channel(nx=64, ny=32, nz=32, storage=storage, tau=0.6))
            cfg.solid_bc = bc
            g = Solver(cfg)
            g.set_motion([law])
            g.init_equilibrium()
            d = _mean_drag(g.step(1000))
            g.close()
            err = float(abs(d[0] - ref[0]) / abs(ref[0]))
            print(bc, storage, "channel drag error vs FP64", err)
            assert err < tol, (bc, storage, err)


def test_d3q27_shifted_storage_start_state():
    """The shifted formats hold f - w with the D3Q27 weights: a uniform start
    reads back as f_eq to the format's resolution, the rest state as the FP32
    weights exactly (a zero deviation)."""
    # half an FP16 ulp of a deviation near 7e-3 scaled by 2^12 is 1.9e-6; half a BF16 ulp 1.5e-5
    for storage, tol in (("fp16s", 4e-6), ("bf16s", 5e-5)):
        cfg = _q27(periodic_box(n=8, collision="TRT", storage=storage, g=(0.0, 0.0, 0.0)))
        cfg.u0 = (0.0, 0.0, 0.0)
        g = Solver(cfg)
        g.init_equilibrium()
        w32 = W27.astype(np.float32).astype(np.float64)
        assert np.array_equal(g.populations(), np.repeat(w32[:, None], cfg.ncells, axis=1))
        g.close()
        cfg.u0 = (0.03, -0.02, 0.01)
        g = Solver(cfg)
        g.init_equilibrium()
        assert np.abs(g.populations() - _feq27(cfg.u0)[:, None]).max() < tol
        g.close()


def test_d3q27_shape_change_between_steps_fp64():
    """lbm_set_bodies between steps with a smaller shape at a shifted pose: the
    flags of the larger old shape are cleared (the box rebuild takes the larger
    reach of the two poses)."""
    cfg = _q27(periodic_box(n=24, collision="TRT", g=(1e-5, 0.0, 0.0)))
    cfg.u0 = (0.02, -0.01, 0.005)  # u well above its FP64 rounding floor
    g, o = _pair(cfg, "fp64")
    for s in (g, o):
        s.set_bodies([sphere((12.0, 12.0, 12.0), 9.0)])
        s.init_equilibrium()
        s.step(3)
        s.set_bodies([sphere((12.4, 12.0, 12.0), 2.5)])
        s.step(3)
        # the flags buffer swapped in holds the R = 9 pose, far beyond the new reach
        s.set_bodies([sphere((12.8, 11.7, 12.0), 2.0)])
    Fg, Fo = g.step(4), o.step(4)
    assert np.array_equal(g.flags(), o.flags())
    check_forces(Fg, Fo, o, "fp64", "shape change", shapes=[sphere((12.8, 11.7, 12.0), 2.0)])
    _compare_state(g, o, TOL["fp64"], "shape change")
