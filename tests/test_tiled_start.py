"""Tile-wise warm-start (CPU): the tile placement along one axis, the stitching
of tile cores into one lattice, and the start it hands to a solver.

A stand-in network made of local operations of radius r (box sums and
shifts of the mask, with edge padding) gives, on every core cell of a tile
with halo >= r, exactly what it gives on the whole lattice at once; so the
tiled prediction must equal the whole-lattice one bit for bit, and a halo
smaller than r must not (the check has teeth).  Shifts of different sign on
the three axes make a transposed or mirrored tile show."""
from __future__ import annotations

import numpy as np
import pytest

from lbm3d_bb_amd.warmstart import plausible, predict_tiled, tile_spans, warm_start_tiled


def test_tile_spans_cover_each_axis():
    for n in range(1, 41):
        for t in range(1, n + 1):
            for halo in range(0, (t + 1) // 2):
                spans = tile_spans(n, t, halo)
                assert spans[0][1] == 0 and spans[-1][2] == n
                for k, (s, lo, hi) in enumerate(spans):
                    assert 0 <= s and s + t <= n
                    assert s <= lo < hi <= s + t
                    assert lo - s >= min(halo, lo)            # left context
                    assert s + t - hi >= min(halo, n - hi)     # right context
                    if k:
                        assert lo == spans[k - 1][2]
                # no more tiles than the cores need
                assert len(spans) <= 1 + -(-max(n - (t - halo), 0) // (t - 2 * halo))


def test_tile_spans_reject_bad_tiles():
    with pytest.raises(ValueError):
        tile_spans(10, 11, 0)
    with pytest.raises(ValueError):
        tile_spans(10, 8, 4)
    assert tile_spans(16, 16, 4) == [(0, 0, 16)]


class LocalNet:
    """Input (B, 2, tz, ty, tx) of (mask, Re channel) -> (B, 4, ...): box sum
    of the mask of radius r, the mask shifted by +r along x, by -r along y,
    and the Re channel times (1 + the mask shifted by +1 along z)."""

    def __init__(self, shape, r=2, batch=3):
        self.shape, self.r, self.batch, self.cin = tuple(shape), r, batch, 2
        self.calls = 0

    def forward(self, x):
        assert x.shape[1:] == (2,) + self.shape and 1 <= x.shape[0] <= self.batch
        self.calls += 1
        r = self.r
        y = np.empty((x.shape[0], 4) + self.shape, dtype=np.float32)
        for b in range(x.shape[0]):
            m = x[b, 0]
            p = np.pad(m, r, mode="edge")
            tz, ty, tx = self.shape
            box = np.zeros(self.shape, dtype=np.float32)
            for dz in range(2 * r + 1):
                for dy in range(2 * r + 1):
                    for dx in range(2 * r + 1):
                        box += p[dz:dz + tz, dy:dy + ty, dx:dx + tx]
            y[b, 0] = box
            y[b, 1] = p[r:r + tz, r:r + ty, 2 * r:2 * r + tx]
            y[b, 2] = p[r:r + tz, 0:ty, r:r + tx]
            y[b, 3] = x[b, 1] * (1.0 + p[r + 1:r + 1 + tz, r:r + ty, r:r + tx])
        return y


def _mask(shape, seed=0):
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    m = np.zeros(shape, dtype=bool)
    for _ in range(6):
        c = rng.uniform(0, 1, 3) * np.array([nz, ny, nx])
        r = rng.uniform(2.0, 5.0)
        m |= (z - c[0]) ** 2 + (y - c[1]) ** 2 + (x - c[2]) ** 2 <= r * r
    return m


@pytest.mark.parametrize("tile", [(12, 10, 16), (8, 22, 9)])
def test_tiled_prediction_equals_whole_lattice(tile):
    shape = (20, 22, 37)
    m = _mask(shape)
    re, u_ref = 37.0, 0.05
    whole = LocalNet(shape, batch=1)
    rho0, u0 = predict_tiled(whole, m, re, u_ref, halo=2)
    net = LocalNet(tile)
    rho, u = predict_tiled(net, m, re, u_ref, halo=2)
    ntiles = np.prod([len(tile_spans(n, t, 2)) for n, t in zip(shape, tile)])
    assert ntiles > 1 and net.calls == -(-ntiles // net.batch)
    assert np.array_equal(rho, rho0) and np.array_equal(u, u0)
    assert np.all(u[:, m.reshape(-1)] == 0.0)
    # a halo below the network's reach changes the result near the tile seams
    rho1, u1 = predict_tiled(LocalNet(tile), m, re, u_ref, halo=1)
    assert not (np.array_equal(rho1, rho0) and np.array_equal(u1, u0))


def test_tiled_start_initialises_the_solver():
    from lbm3d_bb_amd.bodies import sphere
    from lbm3d_bb_amd.solver import Config, periodic, velocity, zero_gradient

    from oracle.oracle import OracleSolver

    for q in (19, 27):
        cfg = Config(24, 16, 16, collision="TRT", omega_plus=1.0 / 0.6, omega_minus=1.0 / 0.9, q=q,
                     faces=[velocity((0.05, 0, 0)), zero_gradient(), periodic(), periodic(), periodic(), periodic()],
                     u0=(0.05, 0.0, 0.0))
        o = OracleSolver(cfg)
        o.set_bodies([sphere((11.3, 8.2, 7.6), 3.4)])
        net = LocalNet((8, 8, 16), r=1, batch=2)
        # small outputs: scale the stand-in so that (rho, u) stays plausible
        fwd = net.forward
        net.forward = lambda x: 0.01 * fwd(x)
        info = warm_start_tiled(o, net, re=20.0, u_ref=0.05, halo=2)
        assert info["start"] == "net" and info["tiles"] == 3 * 3 * 2
        m = (o.flags() != 0)
        rho_p, u_p = predict_tiled(net, m.reshape(16, 16, 24), 20.0, 0.05, halo=2)
        assert plausible(rho_p, u_p, 0.05)[0]
        rho, u = o.macros()
        fl = ~m
        assert np.abs(rho[fl] - rho_p[fl]).max() < 1e-13
        assert np.abs(u[:, fl] - u_p[:, fl]).max() < 1e-13
        o.close()


def test_implausible_tiled_prediction_falls_back_to_uniform():
    from lbm3d_bb_amd.solver import Config

    from oracle.oracle import OracleSolver

    o = OracleSolver(Config(16, 8, 8, u0=(0.02, 0.0, 0.0)))
    net = LocalNet((8, 8, 8), r=1)
    fwd = net.forward
    net.forward = lambda x: np.full_like(fwd(x), np.nan)
    info = warm_start_tiled(o, net, re=10.0, u_ref=0.02, halo=1)
    assert info["start"] == "uniform" and "finite" in info["reason"]
    rho, u = o.macros()
    assert np.abs(rho - 1.0).max() < 1e-14 and np.abs(u[0] - 0.02).max() < 1e-14
    o.close()


@pytest.mark.gpu
@pytest.mark.parametrize("q", [19, 27])
def test_gpu_tiled_start_with_the_unet(q):
    """The library's U-Net applied tile-wise (two tiles per forward call) on a
    channel with a tilted ellipsoid: each core equals a single-tile forward,
    the HIP solver starts from the stitched prediction, and from that start
    it stays at the FP64 parity bar against the oracle for 30 steps."""
    from lbm3d_bb_amd.bodies import ellipsoid, rotation_matrix
    from lbm3d_bb_amd.solver import Config, Solver, no_slip, velocity, zero_gradient
    from lbm3d_bb_amd.warmstart import UNet, encode_input

    from cases import rel_err
    from oracle.oracle import OracleSolver

    nz, ny, nx = 24, 24, 40
    u_ref, re = 0.05, 30.0
    cfg = Config(nx, ny, nz, collision="TRT", omega_plus=1.0 / 0.6, omega_minus=1.0 / 0.9, q=q, storage="fp64",
                 faces=[velocity((u_ref, 0, 0)), zero_gradient(), no_slip(), no_slip(), no_slip(), no_slip()],
                 u0=(u_ref, 0.0, 0.0))
    body = ellipsoid((17.4, 12.3, 11.8), (7.0, 4.0, 3.5), rotation_matrix((0.3, 1.0, 0.2), 0.7))
    net = UNet((16, 16, 16), batch=2, levels=4, base=16, seed=5)
    s = Solver(cfg)
    o = OracleSolver(cfg)
    try:
        s.set_bodies([body])
        o.set_bodies([body])
        assert np.array_equal(s.flags(), o.flags())
        info = warm_start_tiled(s, net, re, u_ref, halo=4)
        assert info["start"] == "net", info
        assert info["tiles"] == 2 * 2 * 4
        mask = (s.flags() != 0).reshape(nz, ny, nx)
        rho_p, u_p = predict_tiled(net, mask, re, u_ref, halo=4)
        # a core against the network run on its tile alone
        (sz, lz, hz), (sy, ly, hy), (sx, lx, hx) = (tile_spans(nz, 16, 4)[1], tile_spans(ny, 16, 4)[1],
                                                    tile_spans(nx, 16, 4)[1])
        y = net.forward(encode_input(mask[sz:sz + 16, sy:sy + 16, sx:sx + 16], re)[None])[0]
        core = 1.0 + u_ref * y[0, lz - sz:hz - sz, ly - sy:hy - sy, lx - sx:hx - sx].astype(np.float64)
        assert np.abs(rho_p.reshape(nz, ny, nx)[lz:hz, ly:hy, lx:hx] - core).max() < 1e-6 * u_ref
        fl = ~mask.reshape(-1)
        rho, u = s.macros()
        assert np.abs(rho[fl] - rho_p[fl]).max() < 1e-6 * u_ref
        assert np.abs(u[:, fl] - u_p[:, fl]).max() < 1e-6 * u_ref
        assert np.abs(u_p[:, fl]).max() > 1e-3 * u_ref  # the start is not the uniform one
        o.set_populations(s.populations())
        F = s.step(30)
        Fo = o.step(30)
        assert np.abs(F - Fo).max() < 1e-11 * np.abs(Fo).max()
        assert rel_err(s.populations(), o.populations()) < 1e-12
    finally:
        s.close()
        o.close()
        net.close()
