"""Seeded sphere packs and the flow case the warm-start network learns.

A pack is 4-32 non-overlapping spheres with radii uniform in [5, 12] cells at
a 128^3 lattice (both scaled with the lattice size L / 128), centres drawn by
rejection sampling from one numpy Generator seeded per sample.  The flow case
is a channel along x: velocity inlet u_in = 0.08 at x = 0, zero-gradient
outlet at x = L - 1, periodic y and z, D3Q19 TRT with Lambda = 3/16, FP32
populations.  Neither this outlet nor the reference's anti-bounce-back one
(src/OtherBC.jl:137-166, which takes the node's own density) fixes the
density level: the steady state's mass, and the drag with it, carries the
history of the start, and a uniform start keeps drifting by ~0.1 % per 30k
steps after its drag residual first falls below 1e-4.  (The anti-bounce-back
outlet also diverged on these packs within 2k steps at Re 22.)  Its Reynolds number Re = u_in D_ref / nu with D_ref =
2 * 8.5 * L / 128 (the mean diameter of the radius range) fixes nu, hence
tau+ = 1/2 + 3 nu.  Re is drawn uniformly in [1, 100], capped where nu
would fall below NU_MIN (tau+ < 0.53, where the scheme is no longer stable
on these packs): the cap only binds on lattices coarser than 128^3.

The reference solves one body at a time from a uniform start
(src/Initialization.jl:57-83); packs, the channel faces and the Reynolds
parametrisation are this build's own workload definitions.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .bodies import Body, sphere
from .solver import Config, periodic, velocity, zero_gradient

RADIUS_RANGE = (5.0, 12.0)   # at L = 128
COUNT_RANGE = (4, 32)
RE_RANGE = (1.0, 100.0)
D_REF_128 = 17.0             # 2 x mean radius of RADIUS_RANGE
LAMBDA = 3.0 / 16.0
GAP = 2.0                    # minimum surface-to-surface distance (cells)
U_IN = 0.08
NU_MIN = 0.01
X_MARGIN = 0.15              # sphere-free fraction of the domain at the inlet and the outlet


@dataclass
class PackCase:
    seed: int
    L: int
    re: float
    u_in: float
    bodies: list[Body]
    centers: np.ndarray
    radii: np.ndarray

    @property
    def nu(self) -> float:
        return self.u_in * d_ref(self.L) / self.re

    @property
    def tau_plus(self) -> float:
        return 0.5 + 3.0 * self.nu

    def config(self, storage: str = "fp32") -> Config:
        tp = self.tau_plus
        tm = 0.5 + LAMBDA / (tp - 0.5)
        return Config(self.L, self.L, self.L, collision="TRT", storage=storage, omega_plus=1.0 / tp,
                      omega_minus=1.0 / tm,
                      faces=[velocity((self.u_in, 0, 0)), zero_gradient(), periodic(), periodic(), periodic(),
                             periodic()],
                      u0=(self.u_in, 0.0, 0.0))

    def mask(self) -> np.ndarray:
        """Solid indicator (nz, ny, nx) from the same FP32 test the solver uses
        (generate_mask_s_local!, src/Geometry.jl:66-73, with y/z minimum image)."""
        L = self.L
        c = (np.arange(L, dtype=np.float32) + np.float32(0.5))
        m = np.zeros((L, L, L), dtype=bool)
        half = np.float32(0.5 * L)
        for b in self.bodies:
            cx, cy, cz = (np.float32(v) for v in b.center)
            r = np.float32(b.semi[0])
            r2 = r * r
            dx = c - cx
            dy = c - cy
            dy = np.where(dy > half, dy - np.float32(L), np.where(dy < -half, dy + np.float32(L), dy))
            dz = c - cz
            dz = np.where(dz > half, dz - np.float32(L), np.where(dz < -half, dz + np.float32(L), dz))
            s = (dx * dx)[None, None, :] + (dy * dy)[None, :, None]
            s = s + (dz * dz)[:, None, None]
            m |= s <= r2 + np.float32(1e-6) * r2
        return m


def d_ref(L: int) -> float:
    return D_REF_128 * L / 128.0


def re_max(L: int, u_in: float = U_IN) -> float:
    """Largest Reynolds number with nu >= NU_MIN on an L^3 lattice."""
    return min(RE_RANGE[1], u_in * d_ref(L) / NU_MIN)


def make_pack(seed: int, L: int = 128, u_in: float = U_IN, re: float | None = None) -> PackCase:
    """A seeded pack; Re is drawn uniformly in [1, re_max(L)] unless given."""
    rng = np.random.default_rng(seed)
    s = L / 128.0
    n = int(rng.integers(COUNT_RANGE[0], COUNT_RANGE[1] + 1))
    if re is None:
        re = float(rng.uniform(RE_RANGE[0], re_max(L, u_in)))
    centers, radii = [], []
    tries = 0
    while len(centers) < n:
        tries += 1
        if tries > 200000:
            break  # dense draws keep the spheres placed so far
        r = rng.uniform(*RADIUS_RANGE) * s
        lo = X_MARGIN * L + r
        hi = (1.0 - X_MARGIN) * L - r
        c = np.array([rng.uniform(lo, hi), rng.uniform(0, L), rng.uniform(0, L)])
        ok = True
        for c2, r2 in zip(centers, radii):
            d = np.abs(c - c2)
            d[1:] = np.minimum(d[1:], L - d[1:])
            if np.linalg.norm(d) < r + r2 + GAP:
                ok = False
                break
        if ok:
            centers.append(c)
            radii.append(r)
    bodies = [sphere(tuple(float(v) for v in c), float(r)) for c, r in zip(centers, radii)]
    return PackCase(seed, L, float(re), u_in, bodies, np.array(centers), np.array(radii))
