"""Warm-start workload host logic (CPU): packs, encodings, the fallback rule."""
from __future__ import annotations

import numpy as np

from lbm3d_bb_amd.packs import COUNT_RANGE, GAP, make_pack
from lbm3d_bb_amd.warmstart import decode_output, encode_input, encode_target, plausible
from oracle.oracle import OracleSolver


def test_pack_is_seeded_and_non_overlapping():
    a, b = make_pack(11, 64), make_pack(11, 64)
    assert np.array_equal(a.centers, b.centers) and a.re == b.re
    assert COUNT_RANGE[0] <= len(a.bodies) <= COUNT_RANGE[1]
    assert 1.0 <= a.re <= 100.0
    L = a.L
    for i in range(len(a.bodies)):
        for j in range(i):
            d = np.abs(a.centers[i] - a.centers[j])
            d[1:] = np.minimum(d[1:], L - d[1:])
            assert np.linalg.norm(d) >= a.radii[i] + a.radii[j] + GAP - 1e-9
    assert np.all(a.centers[:, 0] - a.radii >= 0.15 * L - 1e-9)


def test_pack_mask_matches_oracle_flags():
    case = make_pack(5, 32)
    o = OracleSolver(case.config("fp64"), "fp64")
    o.set_bodies(case.bodies)
    o.update_geometry()
    flags = o.flags().reshape(32, 32, 32)
    o.close()
    assert np.array_equal(flags != 0, case.mask())


def test_reynolds_sets_viscosity():
    c = make_pack(2, 128, re=50.0)
    assert abs(c.u_in * 17.0 / c.nu - 50.0) < 1e-12
    cfg = c.config()
    assert abs(1.0 / cfg.omega_plus - (0.5 + 3.0 * c.nu)) < 1e-12
    # TRT magic product (tau+ - 1/2)(tau- - 1/2) = 3/16
    assert abs((1.0 / cfg.omega_plus - 0.5) * (1.0 / cfg.omega_minus - 0.5) - 3.0 / 16.0) < 1e-12


def test_encode_decode_roundtrip():
    rng = np.random.default_rng(0)
    shape = (4, 5, 6)
    N = 4 * 5 * 6
    rho = 1.0 + 0.01 * rng.standard_normal(N)
    u = 0.05 * rng.standard_normal((3, N))
    y = encode_target(rho, u, 0.05, shape)
    r2, u2 = decode_output(y, 0.05)
    assert np.abs(r2 - rho).max() < 1e-7 and np.abs(u2 - u).max() < 1e-8
    x = encode_input(rng.random(shape) < 0.2, 100.0)
    assert x.shape == (2,) + shape and np.all(x[1] == 1.0)
    x3 = encode_input(rng.random(shape) < 0.2, 10.0, 3)
    assert x3.shape == (3,) + shape and np.allclose(x3[1], 0.5)
    assert np.all(x3[2, :, :, 0] == 0.0) and np.all(x3[2, :, :, -1] == 1.0) and np.all(np.diff(x3[2], axis=2) > 0)


def test_garbage_prediction_falls_back():
    N = 64
    good_r, good_u = np.ones(N), np.zeros((3, N))
    assert plausible(good_r, good_u, 0.05)[0]
    bad = good_r.copy()
    bad[3] = np.nan
    assert not plausible(bad, good_u, 0.05)[0]
    assert not plausible(good_r * 2.0, good_u, 0.05)[0]
    fast = good_u.copy()
    fast[0, 0] = 1.0
    assert not plausible(good_r, fast, 0.05)[0]
