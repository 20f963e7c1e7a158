"""HIP path vs the oracle on the same inputs (GPU).

Bars (tests/cases.py): solid flags and boundary links bit-exact; populations,
rho and u within a normwise relative tolerance of 1e-12 (FP64 storage vs the
FP64 oracle) or 1e-5 (FP32 vs the FP32 oracle), normwise = max|gpu - oracle|
/ max|oracle| over the whole array; per-body force and torque within the same
tolerance relative to max|F_oracle| plus the rounding floor of a sum of n
cancelling link terms (2 gamma_n sum|t_link|, Higham), which every run logs
next to the error relative to |F|.
"""
from __future__ import annotations

import numpy as np
import pytest

from lbm3d_bb_amd.bodies import ConstantMotion, ellipsoid, rotation_matrix, sphere
from lbm3d_bb_amd.solver import Solver
from oracle.oracle import OracleSolver

from cases import TOL, channel, check_forces, check_state, periodic_box, reference_channel, rel_err, sphere_pack

pytestmark = pytest.mark.gpu

def _pair(cfg, prec):
    return Solver(cfg), OracleSolver(cfg, prec)


def _compare_state(g, o, tol, what=""):
    check_state(g, o, "fp64" if tol < 1e-6 else "fp32", what, tol=tol)


def test_flags_and_links_bit_exact_sphere():
    cfg = periodic_box(n=32)
    g, o = _pair(cfg, "fp64")
    b = [sphere((15.7, 16.2, 16.9), 6.3), sphere((3.0, 3.5, 30.5), 4.0)]  # second one wraps
    for s in (g, o):
        s.set_bodies(b)
        s.update_geometry()
    assert np.array_equal(g.flags(), o.flags())
    cg, dg = g.links()
    co, do = o.links()
    assert len(co) > 0
    assert np.array_equal(cg, co) and np.array_equal(dg, do)


def test_flags_and_links_bit_exact_ellipsoid_channel():
    cfg = channel(nx=48, ny=30, nz=28, storage="fp32")
    g, o = _pair(cfg, "fp32")
    R = rotation_matrix((0.3, -0.5, 0.8), 0.7)
    b = [ellipsoid((20.3, 14.9, 13.2), (9.0, 4.5, 4.0), rot=R)]
    for s in (g, o):
        s.set_bodies(b)
        s.update_geometry()
    assert np.array_equal(g.flags(), o.flags())
    cg, dg = g.links()
    co, do = o.links()
    assert np.array_equal(cg, co) and np.array_equal(dg, do)


def test_initial_state_matches():
    cfg = periodic_box(n=16)
    g, o = _pair(cfg, "fp64")
    rng = np.random.default_rng(3)
    rho = 1.0 + 0.01 * rng.standard_normal(cfg.ncells)
    u = 0.02 * rng.standard_normal((3, cfg.ncells))
    for s in (g, o):
        s.init_equilibrium(rho, u)
    _compare_state(g, o, 1e-14, "init")


@pytest.mark.parametrize("collision", ["BGK", "TRT"])
def test_periodic_sphere_fp64(collision):
    """Periodic box with one fixed sphere (the 64^3 BGK sphere case at 24^3): body force, halfway bounce-back, FP64."""
    cfg = periodic_box(n=24, collision=collision)
    g, o = _pair(cfg, "fp64")
    for s in (g, o):
        s.set_bodies([sphere((12.0, 12.0, 12.0), 4.0)])
        s.init_equilibrium()
    Fg = g.step(60)
    Fo = o.step(60)
    assert np.abs(Fo).max() > 0
    check_forces(Fg, Fo, o, "fp64", collision, shapes=[sphere((12.0, 12.0, 12.0), 4.0)])
    _compare_state(g, o, TOL["fp64"], collision)


def test_multi_body_static_fp64():
    cfg = periodic_box(n=28, collision="TRT", g=(2e-5, 1e-5, 0.0))
    bodies = sphere_pack(4, 28, 2.5, 4.0, seed=0)
    g, o = _pair(cfg, "fp64")
    for s in (g, o):
        s.set_bodies(bodies)
        s.init_equilibrium()
    Fg = g.step(40)
    Fo = o.step(40)
    check_forces(Fg, Fo, o, "fp64", "multi", shapes=bodies)
    _compare_state(g, o, TOL["fp64"], "multi")


def test_channel_translating_sphere_fp64():
    """Channel (inlet, zero-gradient outlet, no-slip walls) with a translating sphere and refill."""
    cfg = channel(nx=40, ny=20, nz=20, storage="fp64")
    g, o = _pair(cfg, "fp64")
    law = ConstantMotion("sphere", (12.3, 10.1, 9.8), (4.0,), vel=(0.05, 0.0, 0.0))
    for s in (g, o):
        s.set_motion([law])
        s.init_equilibrium()
    Fg = g.step(50)
    Fo = o.step(50)
    check_forces(Fg, Fo, o, "fp64", "channel", shapes=[law])
    _compare_state(g, o, TOL["fp64"], "channel")
    assert np.array_equal(g.flags(), o.flags())


def test_channel_translating_sphere_fp32():
    cfg = channel(nx=40, ny=20, nz=20, storage="fp32")
    g, o = _pair(cfg, "fp32")
    law = ConstantMotion("sphere", (12.3, 10.1, 9.8), (4.0,), vel=(0.01, 0.0, 0.0))
    for s in (g, o):
        s.set_motion([law])
        s.init_equilibrium()
    Fg = g.step(30)
    Fo = o.step(30)
    check_forces(Fg, Fo, o, "fp32", "channel fp32", shapes=[law])
    _compare_state(g, o, TOL["fp32"], "fp32")


def test_reference_face_set_fp64():
    """The reference's faces: velocity inlet and y-walls at U_inf, ABB outlet, z periodic,
    with the reference's wall-velocity sampling point and a sphere moving in y."""
    cfg = reference_channel(nx=40, ny=24, nz=8, storage="fp64")
    g, o = _pair(cfg, "fp64")
    law = ConstantMotion("sphere", (14.0, 12.0, 4.0), (3.5,), vel=(0.0, 0.02, 0.0), omega=(0.0, 0.0, 0.01))
    for s in (g, o):
        s.set_motion([law])
        s.init_equilibrium()
    Fg = g.step(40)
    Fo = o.step(40)
    check_forces(Fg, Fo, o, "fp64", "refchannel", shapes=[law])
    _compare_state(g, o, TOL["fp64"], "refchannel")


def test_rotating_ellipsoid_refill_fp64():
    cfg = channel(nx=36, ny=28, nz=28, storage="fp64", u_in=0.02)
    g, o = _pair(cfg, "fp64")
    law = ConstantMotion("ellipsoid", (17.0, 14.0, 14.0), (8.0, 4.0, 4.0), omega=(0.0, 0.02, 0.03),
                         rot0=rotation_matrix((1, 1, 0), 0.4))
    for s in (g, o):
        s.set_motion([law])
        s.init_equilibrium()
    Fg = g.step(30)
    Fo = o.step(30)
    assert np.abs(Fo[:, :, 3:]).max() > 0
    check_forces(Fg, Fo, o, "fp64", "ellipsoid", shapes=[law])
    _compare_state(g, o, TOL["fp64"], "ellipsoid")
    assert np.array_equal(g.flags(), o.flags())


@pytest.mark.parametrize("storage,tol", [("fp16s", 5e-3), ("bf16s", 5e-2)])
def test_shifted_storage_drag_vs_fp64(storage, tol):
    cfg64 = periodic_box(n=24, collision="TRT", storage="fp64", g=(1e-5, 0, 0))
    cfgs = periodic_box(n=24, collision="TRT", storage=storage, g=(1e-5, 0, 0))
    o = OracleSolver(cfg64, "fp64")
    g = Solver(cfgs)
    for s in (g, o):
        s.set_bodies([sphere((12.0, 12.0, 12.0), 4.0)])
        s.init_equilibrium()
    Fg = g.step(200)
    Fo = o.step(200)
    err = abs(Fg[-1, 0, 0] - Fo[-1, 0, 0]) / abs(Fo[-1, 0, 0])
    assert err < tol, err


def test_set_populations_roundtrip_and_step():
    cfg = periodic_box(n=16, collision="TRT")
    g, o = _pair(cfg, "fp64")
    rng = np.random.default_rng(7)
    from lbm3d_bb_amd.lattice import equilibrium

    rho = 1.0 + 0.01 * rng.standard_normal(cfg.ncells)
    u = 0.03 * rng.standard_normal((3, cfg.ncells))
    f = equilibrium(rho, *u) + 1e-4 * rng.standard_normal((19, cfg.ncells))
    for s in (g, o):
        s.set_bodies([sphere((8.0, 8.0, 8.0), 3.0)])
        s.set_populations(f)
    assert rel_err(g.populations(), f) < 1e-15
    g.step(10)
    o.step(10)
    _compare_state(g, o, TOL["fp64"], "setpop")


def _cells_linking_two_bodies(o, q=19):
    """Fluid cells with links into more than one body (oracle geometry)."""
    from lbm3d_bb_amd.lattice import C, C27

    cfg = o.config
    cells, dirs = o.links()
    c = (C27 if q == 27 else C)[dirs.astype(np.int64)]
    x, y, z = cells % cfg.nx, (cells // cfg.nx) % cfg.ny, cells // (cfg.nx * cfg.ny)
    nb = ((x + c[:, 0]) % cfg.nx) + cfg.nx * (((y + c[:, 1]) % cfg.ny) + cfg.ny * ((z + c[:, 2]) % cfg.nz))
    body = o.flags()[nb].astype(np.int64)
    pairs = np.unique(np.stack([cells.astype(np.int64), body], 1), axis=0)
    cnt = np.bincount(pairs[:, 0])
    return int((cnt >= 2).sum())


@pytest.mark.parametrize("q", [19, 27])
def test_forces_deterministic(q):
    """Bitwise equal forces on repeated runs, on a pack where fluid cells sit in
    one-cell gaps between spheres and link to two bodies (their second body's
    links go through the wave-ordered sums, the rest through the per-thread
    accumulators), static and moving."""
    cfg = periodic_box(n=28, collision="TRT")
    cfg.q = q
    bodies = [sphere((7.0, 14.0, 14.0), 4.6), sphere((17.5, 14.0, 14.0), 4.6), sphere((12.3, 22.9, 14.0), 4.6),
              sphere((12.0, 14.0, 3.0), 2.8)]
    o = OracleSolver(cfg, "fp64")
    o.set_bodies(bodies)
    o.update_geometry()
    two = _cells_linking_two_bodies(o, q)
    assert two > 10, two
    laws = [ConstantMotion("sphere", tuple(b.center), tuple(b.semi), vel=(0.01 * k, -0.005, 0.0))
            for k, b in enumerate(bodies)]
    for setup in ("bodies", "motion"):
        runs = []
        for _ in range(3):
            g = Solver(cfg)
            if setup == "bodies":
                g.set_bodies(bodies)
            else:
                g.set_motion(laws)
            g.init_equilibrium()
            runs.append(g.step(30))
            g.close()
        assert np.abs(runs[0]).max() > 0
        assert np.array_equal(runs[0], runs[1]) and np.array_equal(runs[0], runs[2]), setup


def test_mass_conserved_periodic_no_body():
    cfg = periodic_box(n=20, collision="TRT", storage="fp64", g=(1e-4, 0, 0))
    g = Solver(cfg)
    rng = np.random.default_rng(0)
    g.init_equilibrium(1.0 + 0.01 * rng.standard_normal(cfg.ncells), 0.01 * rng.standard_normal((3, cfg.ncells)))
    m0, p0 = g.conservation()
    g.step(100)
    m1, p1 = g.conservation()
    assert abs(m1 - m0) / m0 < 1e-13
    # momentum grows by exactly sum(F) = M g per step
    assert abs((p1[0] - p0[0]) - 100 * m0 * 1e-4) < 1e-10 * m0


def _mean_drag(F, last=100):
    return np.asarray(F)[-last:, :, 0].mean(axis=0)


def test_shifted_storage_channel_drag_vs_fp64():
    """Channel shape (inlet, zero-gradient outlet, no-slip walls) with a
    translating sphere: drag of the 16-bit storages against the FP64 oracle,
    averaged over the last 100 of 1000 steps."""
    law = ConstantMotion("sphere", (20.0, 16.0, 16.0), (5.0,), vel=(0.01, 0.0, 0.0))
    o = OracleSolver(channel(nx=64, ny=32, nz=32, storage="fp64", tau=0.6), "fp64")
    o.set_motion([law])
    o.init_equilibrium()
    ref = _mean_drag(o.step(1000))
    o.close()
    for storage, tol in (("fp16s", 5e-3), ("bf16s", 5e-2)):
        g = Solver(channel(nx=64, ny=32, nz=32, storage=storage, tau=0.6))
        g.set_motion([law])
        g.init_equilibrium()
        d = _mean_drag(g.step(1000))
        g.close()
        err = float(abs(d[0] - ref[0]) / abs(ref[0]))
        print(storage, "channel drag error vs FP64", err)
        assert err < tol, (storage, err)


def test_shifted_storage_per_sphere_drag_vs_fp64():
    """Periodic box with a seeded sphere pack driven by a body force (the
    shape of the 64-sphere benchmark, 48^3 with 6 spheres): per-sphere drag
    of FP32, FP16S and BF16S storage against the FP64 oracle."""
    bodies = sphere_pack(6, 48, 3.0, 5.0, seed=0)

    def cfg(storage):
        return periodic_box(n=48, collision="TRT", storage=storage, tau=0.6, g=(2e-5, 0.0, 0.0))

    o = OracleSolver(cfg("fp64"), "fp64")
    o.set_bodies(bodies)
    o.init_equilibrium()
    ref = _mean_drag(o.step(1500))
    o.close()
    for storage, tol in (("fp32", 1e-4), ("fp16s", 5e-3), ("bf16s", 5e-2)):
        g = Solver(cfg(storage))
        g.set_bodies(bodies)
        g.init_equilibrium()
        d = _mean_drag(g.step(1500))
        g.close()
        err = float(np.abs(d - ref).max() / np.abs(ref).max())
        print(storage, "per-sphere drag error vs FP64", err)
        assert err < tol, (storage, err)


def test_periodic_box_64_bgk_sphere_full_size_fp64():
    """The reference-runnable case at its full size: 64^3 periodic box, BGK
    tau = 0.8, FP64, body force g = 1e-6 along x, one fixed sphere of radius 8
    centred at (32, 32, 32), 1000 steps; the drag every 10 steps and the final
    state against the FP64 oracle."""
    cfg = periodic_box(n=64, collision="BGK", storage="fp64", tau=0.8, g=(1e-6, 0.0, 0.0))
    g, o = _pair(cfg, "fp64")
    for s in (g, o):
        s.set_bodies([sphere((32.0, 32.0, 32.0), 8.0)])
        s.init_equilibrium()
    Fg = g.step(1000)
    Fo = o.step(1000)
    assert np.abs(Fo[9::10, 0, 0]).min() > 0
    check_forces(Fg[9::10], Fo[9::10], o, "fp64", "box64", shapes=[sphere((32.0, 32.0, 32.0), 8.0)])
    _compare_state(g, o, TOL["fp64"], "box64")


def test_shape_change_between_steps_fp64():
    """lbm_set_bodies between steps with a smaller shape at a shifted pose: the
    cells of the larger old shape turn fluid (the body region spans the larger
    reach of the two poses), also once the body stops moving."""
    cfg = periodic_box(n=24, collision="TRT", g=(1e-5, 0.0, 0.0))
    cfg.u0 = (0.02, -0.01, 0.005)  # u well above its FP64 rounding floor
    g, o = _pair(cfg, "fp64")
    for s in (g, o):
        s.set_bodies([sphere((12.0, 12.0, 12.0), 9.0)])
        s.init_equilibrium()
        s.step(3)
        s.set_bodies([sphere((12.8, 11.7, 12.0), 2.0)])
    Fg, Fo = g.step(4), o.step(4)
    assert np.array_equal(g.flags(), o.flags())
    check_forces(Fg, Fo, o, "fp64", "shape change", shapes=[sphere((12.8, 11.7, 12.0), 2.0)])
    _compare_state(g, o, TOL["fp64"], "shape change")


@pytest.mark.parametrize("q", [19, 27])
def test_bodies_removed_between_steps_fp64(q):
    """lbm_set_bodies with no bodies after steps with one: the flags clear and the
    step continues as the oracle's (no links, no force rows)."""
    cfg = periodic_box(n=16, collision="TRT", g=(1e-5, 0.0, 0.0))
    cfg.q = q
    cfg.u0 = (0.02, -0.01, 0.005)  # u well above its FP64 rounding floor
    g, o = _pair(cfg, "fp64")
    for s in (g, o):
        s.set_bodies([sphere((8.0, 8.0, 8.0), 3.0)])
        s.init_equilibrium()
        s.step(3)
        s.set_bodies([])
        s.step(3)
    assert not g.flags().any() and not o.flags().any()
    _compare_state(g, o, TOL["fp64"], f"removed q{q}")


@pytest.mark.parametrize("q", [19, 27])
@pytest.mark.parametrize("dims,faces", [
    ((2, 3, 5), "pvz"),   # the smallest periodic axis (2 cells) and the smallest walled one (3)
    ((3, 4, 3), "vav"),   # every axis non-periodic at its minimum size
    ((5, 2, 2), "zpp"),
])
def test_tiny_domains_fp64(q, dims, faces):
    """Lattices at the minimum sizes lbm_create accepts, with mixed faces per axis
    (p periodic, v velocity walls, a velocity inlet + anti-bounce-back outlet,
    z velocity inlet + zero-gradient outlet): HIP path against the oracle."""
    from lbm3d_bb_amd.solver import Config, periodic, pressure, velocity, zero_gradient

    fs = []
    for k in faces:
        fs += {"p": [periodic(), periodic()], "v": [velocity((0.02, -0.01, 0.0)), velocity((0.0, 0.0, 0.0))],
               "a": [velocity((0.03, 0.0, 0.01)), pressure()], "z": [velocity((0.02, 0.0, 0.0)), zero_gradient()]}[k]
    cfg = Config(*dims, collision="TRT", storage="fp64", omega_plus=1 / 0.7, omega_minus=1 / 1.1, faces=fs,
                 u0=(0.01, 0.0, 0.0), q=q)
    g, o = _pair(cfg, "fp64")
    for s in (g, o):
        s.init_equilibrium()
    g.step(7)
    o.step(7)
    _compare_state(g, o, TOL["fp64"], f"tiny {dims} {faces}")


@pytest.mark.parametrize("q", [19, 27])
def test_non_finite_force_returns_numeric_error(q):
    """A NaN density next to a body reaches its links: lbm_step returns
    LBM_ERR_NUMERIC (-5) naming the body and the first step whose record is not
    finite, and hands the records over as they are."""
    from lbm3d_bb_amd import _ffi

    cfg = periodic_box(n=16, collision="TRT")
    cfg.q = q
    g = Solver(cfg)
    g.set_bodies([sphere((4.0, 4.0, 4.0), 2.0), sphere((11.0, 11.0, 11.0), 2.6)])
    rho = np.ones(cfg.ncells)
    rho[14 + 16 * (10 + 16 * 10)] = np.nan  # a fluid cell with links into the second sphere
    g.init_equilibrium(rho, None)
    out = np.zeros((3, 2, 6))
    with pytest.raises(_ffi.LbmError) as e:
        g._call("step", 3, out.ctypes.data_as(_ffi._D))
    assert e.value.code == -5
    assert "body 1 at step 1" in str(e.value), str(e.value)
    assert np.isfinite(out[0, 0]).all() and not np.isfinite(out[0, 1]).all()
    g.close()
