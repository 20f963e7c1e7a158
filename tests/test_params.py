"""params.txt reader and LBMParams derivation (src/Parameters.jl:49-164)."""
from __future__ import annotations

import math
import os

import numpy as np
import pytest

from lbm3d_bb_amd.params import read_params_input, trt_rates

HERE = os.path.dirname(os.path.abspath(__file__))


def test_read_case_file():
    p = read_params_input(os.path.join(HERE, "golden", "params_sphere_case.txt"))
    assert (p.Nx, p.Ny, p.Nz, p.q, p.n_max, p.n_save) == (256, 128, 128, 19, 20000, 1000)
    assert p.bc == "SBB" and p.coll == "TRT" and p.shape == "CYL" and p.naca_code == "2412"
    assert p.chord == 60.0 and p.R == 12.0
    # The reference takes the first number on each line (src/Parameters.jl:124), which on
    # "Theta0: | 15" is the 0 of the key itself: theta0 reads as 0 whatever the value.
    assert p.theta0 == 0.0
    # nu = U (2R) / Re = 0.05 * 24 / 60 = 0.02; tau+ = 0.56; tau- = 0.5 + 0.1875/0.06
    assert abs(p.nu - 0.02) < 1e-8
    assert abs(p.omega_p - 1 / 0.56) < 1e-6
    assert abs(p.omega_m - 1 / (0.5 + 0.1875 / 0.06)) < 1e-6


def test_trt_magic_product():
    # Lambda = (1/omega+ - 1/2)(1/omega- - 1/2)
    for nu in (0.01, 0.05, 0.2):
        wp, wm = trt_rates(nu, 0.1875, dtype=np.float64)
        assert abs((1 / wp - 0.5) * (1 / wm - 0.5) - 0.1875) < 1e-12


def test_parse_errors(tmp_path):
    f = tmp_path / "bad.txt"
    f.write_text("Nx: | 10\nBC: | SBB\n")
    with pytest.raises(ValueError, match="Expected 19"):
        read_params_input(str(f))
    lines = [f"v{i}: | {i + 1}" for i in range(19)]
    f.write_text("\n".join(lines) + "\n")
    with pytest.raises(ValueError, match="Missing BC"):
        read_params_input(str(f))
    f.write_text("\n".join(lines) + "\nBC: | SBB\nShape: | BOX\n")
    with pytest.raises(ValueError, match="CYL or AIR"):
        read_params_input(str(f))


def test_airfoil_defaults_naca():
    import tempfile

    lines = [f"v{i}: | {i + 1}" for i in range(19)] + ["BC: | IBB", "Shape: | AIR"]
    with tempfile.NamedTemporaryFile("w", suffix=".txt", delete=False) as fh:
        fh.write("\n".join(lines) + "\n")
    p = read_params_input(fh.name)
    os.unlink(fh.name)
    assert p.naca_code == "0012" and p.coll == "BGK" and p.shape == "AIR"


def test_lattice_q_reaches_the_config():
    """gen_lattice_sets(q) (src/Utilities.jl:945-993): the input file's q selects the lattice."""
    import dataclasses

    from lbm3d_bb_amd.params import config_from_params

    p = read_params_input(os.path.join(HERE, "golden", "params_sphere_case.txt"))
    cfg, _ = config_from_params(p)
    assert cfg.q == 19 and cfg.to_c().q == 19
    cfg27, _ = config_from_params(dataclasses.replace(p, q=27))
    assert cfg27.q == 27 and cfg27.to_c().q == 27
    with pytest.raises(ValueError):
        config_from_params(dataclasses.replace(p, q=15))[0].to_c()


def test_perturbed_start_is_the_reference_form():
    """initialize_f's start (src/Initialization.jl:27-44,70): u = U (1 + 0.2 r)
    with one r in [-1, 1) per node for all three components, rho = rho_inf;
    seeded, so the same seed gives the same field."""
    import numpy as np

    from lbm3d_bb_amd.params import perturbed_start
    from lbm3d_bb_amd.solver import Config

    cfg = Config(20, 16, 12, u0=(0.075, -0.02, 0.01), rho0=1.02)
    rho, u = perturbed_start(cfg, seed=3)
    assert rho.shape == (cfg.ncells,) and (rho == 1.02).all()
    assert u.shape == (3, cfg.ncells)
    r = u / np.asarray(cfg.u0)[:, None]
    assert np.allclose(r[0], r[1]) and np.allclose(r[0], r[2])  # one r per node
    assert r.min() >= 0.8 - 1e-12 and r.max() < 1.2
    assert abs(r.mean() - 1.0) < 0.01 and r.std() > 0.1  # uniform on [0.8, 1.2): std 0.2/sqrt(3)
    rho2, u2 = perturbed_start(cfg, seed=3)
    assert np.array_equal(u, u2)
    assert not np.array_equal(u, perturbed_start(cfg, seed=4)[1])
