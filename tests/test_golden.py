"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the oracle): the oracle must keep reproducing them (CPU), and the HIP
path must match them (GPU) to the FP64 parity bar.  Flags and links are
integer data and must match exactly."""
from __future__ import annotations

import importlib.util
import os

import numpy as np
import pytest

from cases import EPS, rel_err, u_bar

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def _maker():
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLDEN, "make_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


MG = _maker()
NAMES = sorted(MG.CASES)


def _load(name):
    with np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _check(s, gold, steps, tol):
    """Flags and links exact; populations and rho normwise relative below tol, u
    within the velocity bar of tests/cases.py; force and torque below tol of their maximum plus the rounding floor of the
    FP64 link sums (tests/cases.py: 2 gamma_n sum|t_link|, times the arm bound
    half the domain diagonal for the torque)."""
    from lbm3d_bb_amd.lattice import W, W27

    F = s.step(steps)
    assert np.array_equal(s.flags(), gold["flags"])
    cells, dirs = s.links()
    assert np.array_equal(cells, gold["link_cell"]) and np.array_equal(dirs, gold["link_dir"])
    rho, u = s.macros()
    w = W27 if gold["f"].shape[0] == 27 else W
    n = len(dirs)
    A = float((2.0 * w[dirs.astype(np.int64)] * np.abs(gold["rho"][cells.astype(np.int64)])).sum())
    floor = 2.0 * n * EPS["fp64"] / (1.0 - n * EPS["fp64"]) * A
    arm = 0.5 * float(np.linalg.norm([s.config.nx, s.config.ny, s.config.nz]))
    Fg = gold["forces"]
    assert np.abs(F[..., :3] - Fg[..., :3]).max() <= tol * np.abs(Fg[..., :3]).max() + floor
    assert np.abs(F[..., 3:] - Fg[..., 3:]).max() <= tol * np.abs(Fg[..., 3:]).max() + floor * arm
    assert rel_err(rho, gold["rho"]) < tol
    f = s.populations()
    assert rel_err(f, gold["f"]) < tol
    df = float(np.abs(f - gold["f"]).max())
    assert np.abs(u - gold["u"]).max() <= u_bar(gold["u"], gold["rho"], df, f.shape[0], "fp64", tol)


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_golden(name):
    cfg, o = MG.build(name)
    o.init_equilibrium()
    _check(o, _load(name), MG.CASES[name]["steps"], 1e-14)
    o.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_hip_matches_golden(name):
    """The HIP path against the fixtures at the FP64 bar."""
    from lbm3d_bb_amd.solver import Solver

    cfg, o = MG.build(name)
    o.close()
    s = Solver(cfg)
    MG.setup(name, s)
    s.init_equilibrium()
    _check(s, _load(name), MG.CASES[name]["steps"], 1e-12)
    s.close()


def test_benchmark_body_lists_match_fixtures():
    """The bodies the benchmark configurations draw from seed 0 (bench.make_case)
    are the committed lists (tests/golden/make_config_bodies.py wrote them)."""
    import json

    spec = importlib.util.spec_from_file_location("make_config_bodies", os.path.join(GOLDEN, "make_config_bodies.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    for name, fname in (("box512_spheres64", "box512_spheres64_seed0.json"),
                        ("channel1024_ellipsoids8", "channel1024_ellipsoids8_seed0.json")):
        with open(os.path.join(GOLDEN, fname), encoding="utf-8") as fh:
            assert mod.bodies_of(name) == json.load(fh), name
    with open(os.path.join(GOLDEN, "box512_spheres64_seed0.json"), encoding="utf-8") as fh:
        spheres = np.array(json.load(fh)["spheres"])
    # 64 spheres, radius in [6, 10], pairwise gaps of at least 2 cells across the periodic box
    assert spheres.shape == (64, 4) and (spheres[:, 3] >= 6).all() and (spheres[:, 3] <= 10).all()
    L = np.array([512.0, 256.0, 256.0])
    d = np.abs(spheres[:, None, :3] - spheres[None, :, :3])
    d = np.linalg.norm(np.minimum(d, L - d), axis=-1)
    gap = d - spheres[:, None, 3] - spheres[None, :, 3]
    assert gap[~np.eye(64, dtype=bool)].min() >= 2.0
