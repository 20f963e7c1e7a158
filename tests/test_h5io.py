"""The reference's HDF5 outputs (src/I_O.jl:296-405, 587-662), written through the
HDF5 C library and read back with it; the run driver checked on the oracle."""
from __future__ import annotations

import math
import os

import numpy as np
import pytest

h5io = pytest.importorskip("lbm3d_bb_amd.h5io")

try:
    h5io.lib()
except ImportError as e:  # pragma: no cover - depends on the image
    pytest.skip(f"HDF5 C library not available: {e}", allow_module_level=True)

HERE = os.path.dirname(os.path.abspath(__file__))


def test_field_snapshots_layout_and_roundtrip(tmp_path):
    nx, ny, nz, snaps = 7, 5, 3, 4
    path = str(tmp_path / "final_fields.h5")
    fs = h5io.FieldSnapshots(path, nx, ny, nz, snaps)
    cell = np.arange(nx * ny * nz, dtype=np.float64)  # cell = x + nx (y + ny z)
    u = np.stack([cell + 0.25, -cell, 2 * cell])
    fs.write(2, 20, 1.0 + 1e-3 * cell, u)
    with h5io.H5File(path) as f:
        assert f.info("rho") == ((snaps, nz, ny, nx), "float", 4)  # Julia (Nx, Ny, Nz, S), x fastest
        assert f.info("steps") == ((snaps,), "integer", 4)
        rho = f.read("rho")
        v = f.read("v")
        steps = f.read("steps")
    # element (x, y, z) of snapshot 2 is Julia rho[x+1, y+1, z+1, 2]
    x, y, z = 5, 3, 2
    c = x + nx * (y + ny * z)
    assert rho[1, z, y, x] == np.float32(1.0 + 1e-3 * c)
    assert v[1, z, y, x] == np.float32(-c)
    # unwritten snapshots read as the zeros the reference creates
    assert not rho[0].any() and not rho[3].any()
    assert list(steps) == [0, 20, 0, 0]


def test_aero_coefficients_blocks(tmp_path):
    path = str(tmp_path / "aero_coeffs.h5")
    ac = h5io.AeroCoefficients(path, 10)
    ac.write(1, np.arange(4), np.arange(4) + 10, np.zeros(4))
    ac.write(5, [7.5] * 6, [8.5] * 6, [9.5] * 6)
    with h5io.H5File(path) as f:
        cl, cd = f.read("CL"), f.read("CD")
        assert f.info("CZ") == ((10,), "float", 4)
    assert list(cl) == [0, 1, 2, 3] + [7.5] * 6
    assert list(cd) == [10, 11, 12, 13] + [8.5] * 6


def test_run_driver_on_the_oracle(tmp_path):
    """The reference's loop (scaled input file) on the oracle: the files hold the
    oracle's own moments at every n_save step and its coefficients at every step."""
    from lbm3d_bb_amd.params import aero_coefficients, config_from_params, perturbed_start, read_params_input
    from lbm3d_bb_amd.run_reference import run
    from oracle.oracle import OracleSolver

    p = read_params_input(os.path.join(HERE, "golden", "params_reference_default.txt"))
    p.Nx, p.Ny, p.Nz, p.R, p.x_c, p.y_c, p.Re = 48, 32, 3, 4.3, 16.5, 16.2, 30.0
    out = str(tmp_path / "output")
    run("", out, n_max=25, n_save=10, params=p, make_solver=lambda cfg: OracleSolver(cfg, "fp32"), log=lambda m: None)
    cfg, laws = config_from_params(p)
    o = OracleSolver(cfg, "fp32")
    o.set_motion(laws)
    o.init_equilibrium(*perturbed_start(cfg, 0))  # the driver's default start
    F, snaps = [], []
    for k in (10, 10, 5):
        F.append(o.step(k))
        if len(F) < 3:
            snaps.append(o.macros())
    F = np.concatenate(F)[:, 0, :]
    cl, cd, cz = aero_coefficients(F, p)
    with h5io.H5File(os.path.join(out, "aero_coeffs.h5")) as f:
        assert np.array_equal(f.read("CL"), cl.astype(np.float32))
        assert np.array_equal(f.read("CD"), cd.astype(np.float32))
    with h5io.H5File(os.path.join(out, "final_fields.h5")) as f:
        assert f.info("u")[0] == (math.ceil(25 / 10), p.Nz, p.Ny, p.Nx)
        assert list(f.read("steps")) == [10, 20, 0]
        rho, u = f.read("rho"), f.read("u")
        for s, (r, uu) in enumerate(snaps):
            assert np.array_equal(rho[s].ravel(), r.astype(np.float32))
            assert np.array_equal(u[s].ravel(), uu[0].astype(np.float32))
